"""Learned warm start for the n = 54 spin-wave family (geneig_learn.h).

A multilayer perceptron (8 features -> 256 -> 256 -> 256, tanh -> 2*54*54)
maps the parameters of a pair -- wavevector q, coupling scale s, lambda_min
target t -- to a complex 54 x 54 matrix whose Cholesky-QR is an approximate
eigenbasis of the reduced matrix C = L^-1 A L^-H.  Training and inference run
on the library's own HIP kernels; this module only sequences them (torch
tensors are device buffers, nothing else):

  forward     geneig_learn_mlp_forward (FP32 GEMM + bias + tanh)
  loss        geneig_learn_loss: Q = Z R^-1 (Cholesky-QR),
              loss = ||off(Q^H C Q)||_F^2 / ||C||_F^2, and dloss/dZ
  backward    geneig_learn_gemm / _tanh_backward / _colsum
  update      geneig_learn_adam
  solve       geneig_learned_n54_batched_launch: network seed -> FP32 Jacobi
              sweeps -> back-transform -> FP64 refinement -> FP64 fallback

The training matrices C are the solver's own (geneig_learn_reduce_n54: FP64
Cholesky of H, complex64 reduction).
"""
import ctypes as C

import numpy as np

from . import (GeneigError, _check_hip, _ptr, _stream, _torch, empty_outputs, gen_spinwave, lib,
               spinwave_params)

N = 54


def _dims():
    L = lib()
    return L.geneig_learn_num_features(), L.geneig_learn_hidden(), L.geneig_learn_num_outputs()


def features(params, stream=None):
    """(batch, 5) parameters -> (batch, 8) network inputs, on the device."""
    torch = _torch()
    F, _, _ = _dims()
    X = torch.empty((params.shape[0], F), dtype=torch.float32, device=params.device)
    _check_hip(lib().geneig_learn_features(_ptr(params), _ptr(X), int(params.shape[0]), _stream(stream)),
               "geneig_learn_features")
    return X


def reduced_matrices(A, B, stream=None):
    """The solver's reduced matrices C = L^-1 A L^-H (complex64, (batch, 54, 54)
    column-major per pair as stored, full Hermitian) and the zhegv info."""
    torch = _torch()
    batch = A.shape[0]
    L = torch.empty_like(B)
    info = torch.zeros(batch, dtype=torch.int32, device=A.device)
    Cm = torch.empty((batch, N, N), dtype=torch.complex64, device=A.device)
    _check_hip(lib().geneig_learn_reduce_n54(_ptr(A), _ptr(B), _ptr(L), _ptr(info), _ptr(Cm), int(batch),
                                             _stream(stream)), "geneig_learn_reduce_n54")
    return Cm, info


def spinwave_family(batch, seed, grid, first_q=0, device="cuda"):
    """Pairs of the n = 54 spin-wave family with a coupling scale per q,
    s ~ U(0.8, 1.2): (A, B, params)."""
    A, B = gen_spinwave(27, batch, seed=seed, grid=grid, first_q=first_q, jscale=-1.0, device=device)
    P = spinwave_params(batch, seed, grid=grid, first_q=first_q, jscale=-1.0, device=device)
    return A, B, P


class WarmStart:
    """Network weights (one flat FP32 device buffer) and Adam state."""

    def __init__(self, seed=0, device="cuda"):
        torch = _torch()
        F, H, O = _dims()
        self.shapes = [(H, F), (H,), (H, H), (H,), (H, H), (H,), (O, H), (O,)]
        sizes = [int(np.prod(s)) for s in self.shapes]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        total = int(self.offsets[-1])
        host = np.empty(total, dtype=np.float32)
        rng = np.random.default_rng(seed)
        dims = [F, H, H, H, O]
        for l in range(4):
            fi, fo = dims[l], dims[l + 1]
            # Glorot-uniform; the output layer small with bias vec(I), so the
            # untrained network proposes Q = I (plain Jacobi) to start with
            a = np.sqrt(6.0 / (fi + fo)) * (0.05 if l == 3 else 1.0)
            w = host[self.offsets[2 * l]:self.offsets[2 * l + 1]]
            w[:] = rng.uniform(-a, a, size=fo * fi)
            b = host[self.offsets[2 * l + 1]:self.offsets[2 * l + 2]]
            if l == 3:
                eye = np.zeros((N * N, 2), dtype=np.float32)
                eye[np.arange(N) * (N + 1), 0] = 1.0
                b[:] = eye.reshape(-1)
            else:
                b[:] = 0.0
        self.device = torch.device(device)
        self.flat = torch.from_numpy(host).to(self.device)
        self.grad = torch.zeros_like(self.flat)
        self.m = torch.zeros_like(self.flat)
        self.v = torch.zeros_like(self.flat)
        self.step = 0
        self._ptrs = None

    def view(self, buf, k):
        return buf[int(self.offsets[k]):int(self.offsets[k + 1])].view(*self.shapes[k])

    def weights(self):
        return [self.view(self.flat, k) for k in range(8)]

    def weight_ptrs(self):
        if self._ptrs is None:
            self._ptrs = (C.c_void_p * 8)(*[self.view(self.flat, k).data_ptr() for k in range(8)])
        return C.cast(self._ptrs, C.c_void_p)

    def state_numpy(self):
        return self.flat.cpu().numpy().copy()

    def load_numpy(self, arr):
        torch = _torch()
        self.flat.copy_(torch.from_numpy(np.asarray(arr, dtype=np.float32)).to(self.device))

    def forward(self, X, stream=None):
        """Z (batch, 5832) and the hidden activations (3, batch, 256)."""
        torch = _torch()
        F, H, O = _dims()
        B = X.shape[0]
        Hs = torch.empty((3, B, H), dtype=torch.float32, device=self.device)
        Z = torch.empty((B, O), dtype=torch.float32, device=self.device)
        _check_hip(lib().geneig_learn_mlp_forward(self.weight_ptrs(), _ptr(X), int(B), _ptr(Hs), _ptr(Z),
                                                  _stream(stream)), "geneig_learn_mlp_forward")
        return Z, Hs

    def loss(self, Z, Cset, idx=None, scale=1.0, stream=None):
        """Per-sample loss and dloss/dZ (scaled) for rows Z against Cset[idx]."""
        torch = _torch()
        B = Z.shape[0]
        loss = torch.empty(B, dtype=torch.float32, device=self.device)
        dZ = torch.empty_like(Z)
        _check_hip(lib().geneig_learn_loss(_ptr(Z), _ptr(Cset), _ptr(idx), int(B), float(scale), _ptr(loss),
                                           _ptr(dZ), _stream(stream)), "geneig_learn_loss")
        return loss, dZ

    def backward(self, X, Hs, dZ, stream=None):
        """Fills self.grad with the gradient of sum(dZ * Z) (dZ already holds the
        loss derivative); every product is a library GEMM."""
        torch = _torch()
        F, H, O = _dims()
        B = X.shape[0]
        s = _stream(stream)
        L = lib()
        g = [self.view(self.grad, k) for k in range(8)]
        w = self.weights()
        inputs = [X, Hs[0], Hs[1], Hs[2]]
        widths = [F, H, H, H]
        d = dZ
        dout = O
        for l in (3, 2, 1, 0):
            fin = widths[l]
            _check_hip(L.geneig_learn_gemm(1, 0, dout, fin, B, _ptr(d), dout, _ptr(inputs[l]), fin, _ptr(g[2 * l]),
                                           fin, None, 0, s), "gemm dW")
            _check_hip(L.geneig_learn_colsum(_ptr(d), B, dout, _ptr(g[2 * l + 1]), s), "colsum db")
            if l > 0:
                dn = torch.empty((B, fin), dtype=torch.float32, device=self.device)
                _check_hip(L.geneig_learn_gemm(0, 0, B, fin, dout, _ptr(d), dout, _ptr(w[2 * l]), fin, _ptr(dn), fin,
                                               None, 0, s), "gemm dX")
                _check_hip(L.geneig_learn_tanh_backward(_ptr(inputs[l]), _ptr(dn), B * fin, s), "tanh backward")
                d, dout = dn, fin

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        self.step += 1
        _check_hip(lib().geneig_learn_adam(_ptr(self.flat), _ptr(self.grad), _ptr(self.m), _ptr(self.v),
                                           int(self.flat.numel()), float(lr), float(beta1), float(beta2), float(eps),
                                           int(self.step), _stream(stream)), "geneig_learn_adam")

    def train_step(self, X, Cset, idx, lr, stream=None):
        """One Adam step on the mean loss over the rows; returns the per-sample
        losses (device tensor; NaN marks a rank-deficient network output)."""
        Z, Hs = self.forward(X, stream)
        loss, dZ = self.loss(Z, Cset, idx, scale=1.0 / X.shape[0], stream=stream)
        self.backward(X, Hs, dZ, stream)
        self.adam(lr, stream=stream)
        return loss

    def train(self, X, Cset, epochs, batch=256, lr=1e-3, lr_final=1e-4, seed=0, log=None, spike=1.5,
              nan_spike=0.01):
        """Adam over shuffled mini-batches, learning rate decayed geometrically
        from lr to lr_final.  Returns the mean loss per epoch over the samples
        whose loss is finite (``nan_counts`` holds the others per epoch).

        Loss spikes: a network output close to rank deficiency makes the
        Cholesky-QR gradient huge (and an exactly rank-deficient one a NaN
        loss), and one such step can throw the weights far off (the loss then
        stays high).  The weights and Adam moments of the best epoch so far are
        kept; an epoch whose mean loss exceeds `spike` times the best, or whose
        fraction of NaN losses exceeds `nan_spike`, returns to them with the
        learning rate halved from then on and is never kept as the best; NaN
        samples never enter a mean.  Training ends on the best epoch's weights
        (``train_best`` is its mean loss; ``rollbacks`` counts the returns)."""
        torch = _torch()
        P = X.shape[0]
        gen = torch.Generator(device="cpu").manual_seed(seed)
        steps = epochs * ((P + batch - 1) // batch)
        k = 0
        hist = []
        best = None
        factor = 1.0
        self.rollbacks = 0
        self.nan_counts = []
        for ep in range(epochs):
            perm = torch.randperm(P, generator=gen).to(torch.int32).to(self.device)
            tot = torch.zeros((), dtype=torch.float64, device=self.device)
            nans = torch.zeros((), dtype=torch.int64, device=self.device)
            cnt = 0
            for off in range(0, P, batch):
                idx = perm[off:off + batch].contiguous()
                rate = factor * lr * (lr_final / lr) ** (k / max(1, steps - 1))
                l = self.train_step(X.index_select(0, idx.long()).contiguous(), Cset, idx, rate).double()
                bad = torch.isnan(l)
                tot += torch.where(bad, torch.zeros_like(l), l).sum()
                nans += bad.sum()
                cnt += idx.numel()
                k += 1
            n_nan = int(nans.item())
            self.nan_counts.append(n_nan)
            hist.append(float((tot / (cnt - n_nan)).item()) if cnt > n_nan else float("nan"))
            degenerate = n_nan > nan_spike * cnt
            if log:
                log("epoch %d: mean loss %.5f (%d NaN of %d)" % (ep, hist[-1], n_nan, cnt))
            if not degenerate and (best is None or hist[-1] < best[0]):
                best = (hist[-1], self.flat.clone(), self.m.clone(), self.v.clone(), self.step)
            elif best is not None and (degenerate or hist[-1] > spike * best[0]):
                self._restore(best)
                factor *= 0.5
                self.rollbacks += 1
                if log:
                    log("epoch %d: loss spike (best %.5f, %d NaN): back to the best weights, learning rate x%.3g"
                        % (ep, best[0], n_nan, factor))
        if best is not None and not (hist[-1] <= best[0] and self.nan_counts[-1] <= nan_spike * P):
            self._restore(best)
        self.train_best = best[0] if best is not None else float("nan")
        return hist

    def _restore(self, state):
        _, flat, m, v, step = state
        self.flat.copy_(flat)
        self.m.copy_(m)
        self.v.copy_(v)
        self.step = step


def learned_n54_batched(H, S, W, U, info, X=None, model=None, max_sweeps=12, jacobi_tol=1e-6, max_iter=0, tol=0.0,
                        sweeps=None, iters=None, batch=None, stream=None):
    """geneig_learned_n54_batched_launch: the learned solve (X and model given)
    or the same pipeline from Q = I (X None).  Returns the number of pairs
    re-solved in FP64."""
    batch = H.shape[0] if batch is None else batch
    if (X is None) != (model is None):
        raise GeneigError("features and model go together")
    fb = C.c_int(0)
    _check_hip(lib().geneig_learned_n54_batched_launch(
        _ptr(H), _ptr(S), _ptr(W), _ptr(U), _ptr(info), int(batch), _ptr(X),
        model.weight_ptrs() if model is not None else None, int(max_sweeps), float(jacobi_tol), int(max_iter),
        float(tol), _ptr(sweeps), _ptr(iters), C.byref(fb), _stream(stream)), "geneig_learned_n54_batched_launch")
    return fb.value
