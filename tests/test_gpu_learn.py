"""GPU tests of the learned warm start (geneig_learn.h): the hand-written
GEMM, loss and backward kernels against the CPU oracle (oracle/learn_oracle.py,
float64), the learned solve against the FP64 oracle, and what a trained
network gains over the same pipeline started from Q = I."""
import numpy as np
import pytest

from test_gpu_parity import check_pairs, to_host

pytestmark = pytest.mark.gpu

FAMILY_SEED = 11


@pytest.fixture(scope="module")
def L(geneig):
    return geneig.learn


@pytest.fixture(scope="module")
def lo():
    import learn_oracle

    return learn_oracle


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_vs_numpy(geneig, device, ta, tb):
    import torch

    rng = np.random.default_rng(ta * 2 + tb)
    M, N, K = 70, 133, 45
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    dA, dB, db = (torch.from_numpy(x).to(device) for x in (A, B, bias))
    for act in (0, 1):
        out = torch.empty((M, N), dtype=torch.float32, device=device)
        geneig.lib().geneig_learn_gemm(ta, tb, M, N, K, geneig._ptr(dA), A.shape[1], geneig._ptr(dB), B.shape[1],
                                       geneig._ptr(out), N, geneig._ptr(db), act, geneig._stream(None))
        torch.cuda.synchronize()
        ref = (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64) + bias
        if act:
            ref = np.tanh(ref)
        assert np.abs(to_host(out) - ref).max() <= 1e-4 * max(1.0, np.abs(ref).max())


def family(L, batch, grid, first_q=0):
    A, B, P = L.spinwave_family(batch, FAMILY_SEED, grid, first_q)
    Cm, info = L.reduced_matrices(A, B)
    assert int(info.abs().max().item()) == 0
    return A, B, P, Cm


def test_reduced_matrices_are_the_pencil(L, oracle, device):
    """C = L^-1 A L^-H has the pencil's eigenvalues (complex64 accuracy)."""
    A, B, P, Cm = family(L, 16, (4, 2, 2))
    Ah, Bh = to_host(A), to_host(B)
    Wo, _, _ = oracle.zhegv(Ah, Bh)
    Ch = to_host(Cm)
    for b in range(16):
        c = Ch[b].T   # storage is the transpose (column-major per pair)
        assert np.abs(c - c.conj().T).max() == 0.0
        w = np.linalg.eigvalsh(c.astype(np.complex128))
        assert np.abs(w - Wo[b]).max() <= 1e-4 * np.abs(Wo[b]).max()


def test_network_gradients_vs_oracle(L, lo, device):
    """Forward, loss, dloss/dZ and every weight gradient of the GPU kernels
    against the float64 oracle on the same weights and inputs."""
    import torch

    A, B, P, Cm = family(L, 8, (2, 2, 2), first_q=3)
    X = L.features(P)
    m = L.WarmStart(seed=3, device=device)
    # a non-trivial network: scale the output layer up so Q is far from I
    m.view(m.flat, 6).mul_(40.0)
    Z, Hs = m.forward(X)
    loss, dZ = m.loss(Z, Cm, None, scale=1.0 / 8)
    m.backward(X, Hs, dZ)
    torch.cuda.synchronize()
    w = [to_host(t).astype(np.float64) for t in m.weights()]
    Xh = to_host(X).astype(np.float64)
    assert np.abs(Xh - lo.features(to_host(P))).max() < 1e-5
    hidden, Zo = lo.mlp_forward(Xh, w)
    assert np.abs(to_host(Z) - Zo).max() <= 1e-4 * np.abs(Zo).max()
    Ch = to_host(Cm)
    dZo = np.zeros_like(Zo)
    losses = []
    for b in range(8):
        Zb = lo.z_from_row(Zo[b], 54)
        lb, Gb = lo.cholqr_loss_grad(Zb, Ch[b].T.astype(np.complex128))
        losses.append(lb)
        dZo[b] = lo.row_from_z(Gb) / 8
    print("losses", np.round(losses, 4), flush=True)
    assert np.abs(to_host(loss) - losses).max() <= 2e-3 * max(losses)
    assert np.abs(to_host(dZ) - dZo).max() <= 5e-3 * np.abs(dZo).max()
    grads = lo.mlp_backward(Xh, w, hidden, dZo)
    for k in range(8):
        gk = to_host(m.view(m.grad, k))
        err = np.abs(gk - grads[k]).max() / np.abs(grads[k]).max()
        print("grad %d rel err %.2e" % (k, err), flush=True)
        assert err <= 1e-2, k


def test_learned_solve_matches_oracle(L, geneig, oracle, device):
    """The learned pipeline (untrained network and Q = I) meets the FP64 bar."""
    import torch

    A, B, P, Cm = family(L, 64, (4, 4, 4), first_q=7)
    X = L.features(P)
    m = L.WarmStart(seed=1, device=device)
    Ah, Bh = to_host(A), to_host(B)
    Wo, Xo, _ = oracle.zhegv(Ah, Bh)
    for use in (False, True):
        W, U, info = geneig.empty_outputs(54, 64, device)
        sw = torch.zeros(64, dtype=torch.int32, device=device)
        fb = L.learned_n54_batched(A, B, W, U, info, X=X if use else None, model=m if use else None, sweeps=sw)
        torch.cuda.synchronize()
        print("network %s: fallback %d, sweeps mean %.2f" % (use, fb, to_host(sw).mean()), flush=True)
        assert (to_host(info) == 0).all()
        check_pairs(oracle, Ah, Bh, to_host(W), to_host(U), Wo, Xo, 54, "learned pipeline network=%s" % use)


def test_training_saves_jacobi_sweeps(L, geneig, oracle, device):
    """A short training run lowers the loss, and on unseen q the trained network
    needs fewer Jacobi sweeps than the same pipeline from Q = I, with both
    meeting the FP64 bar."""
    import torch

    A, B, P, Cm = family(L, 2048, (16, 16, 8))
    X = L.features(P)
    del A, B
    m = L.WarmStart(seed=0, device=device)
    hist = m.train(X, Cm, epochs=25, batch=128, lr=2e-3, lr_final=2e-4, seed=0)
    print("training loss", np.round(hist, 4), flush=True)
    assert hist[-1] < 0.7 * hist[0]
    Ae, Be, Pe, _ = family(L, 256, (8, 8, 4), first_q=0)     # q not on the training grid in general
    Xe = L.features(Pe)
    Ah, Bh = to_host(Ae), to_host(Be)
    Wo, Xo, _ = oracle.zhegv(Ah, Bh)
    mean = {}
    for use in (False, True):
        W, U, info = geneig.empty_outputs(54, 256, device)
        sw = torch.zeros(256, dtype=torch.int32, device=device)
        fb = L.learned_n54_batched(Ae, Be, W, U, info, X=Xe if use else None, model=m if use else None, sweeps=sw,
                                   max_sweeps=30)
        torch.cuda.synchronize()
        mean[use] = float(to_host(sw).mean())
        print("network %s: sweeps mean %.2f, fallback %d" % (use, mean[use], fb), flush=True)
        check_pairs(oracle, Ah, Bh, to_host(W), to_host(U), Wo, Xo, 54, "eval network=%s" % use)
    assert mean[True] < mean[False]
