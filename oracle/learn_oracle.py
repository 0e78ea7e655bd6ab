"""CPU oracle of the learned warm start (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's checks; the
product library never imports it.  It restates, in plain numpy, what the HIP
kernels of sunny-batched-geneig_amd/csrc/geneig_learn.hip compute:

  features(params)              cos/sin(2 pi q_k), 5 (s - 1), log10 t + 2
  mlp_forward / mlp_backward    8 -> 256 -> 256 -> 256 (tanh) -> 2 n^2, row-major W[out][in]
  cholqr_loss_grad(Z, C)        Q = Z R^-1 with R^H R = Z^H Z (R upper, real positive
                                diagonal), loss = ||off(Q^H C Q)||_F^2 / ||C||_F^2 and
                                dloss/dZ in the complex (Wirtinger) form G with
                                dloss = Re tr(G^H dZ), i.e. dloss/dRe Z = Re G,
                                dloss/dIm Z = Im G
  jacobi(C, Q0, tol, max_sweeps) round-robin cyclic Jacobi on Q0^H C Q0

The reference has no counterpart (the learned warm start is an addition of
this library); the loss gradient and the network backward are pinned by finite
differences (tests/test_learn_oracle.py), the Jacobi by numpy.linalg.eigh.
"""
import numpy as np


def features(params):
    p = np.asarray(params, dtype=np.float64)
    q = p[:, :3]
    out = np.empty((p.shape[0], 8))
    out[:, 0:6:2] = np.cos(2 * np.pi * q)
    out[:, 1:6:2] = np.sin(2 * np.pi * q)
    out[:, 6] = (p[:, 3] - 1.0) * 5.0
    out[:, 7] = np.log10(p[:, 4]) + 2.0
    return out


def mlp_forward(X, weights):
    """weights = [W1, b1, W2, b2, W3, b3, W4, b4]; returns (hidden list, Z rows)."""
    h = np.asarray(X, dtype=np.float64)
    hidden = []
    for l in range(3):
        h = np.tanh(h @ weights[2 * l].T + weights[2 * l + 1])
        hidden.append(h)
    return hidden, h @ weights[6].T + weights[7]


def mlp_backward(X, weights, hidden, dout):
    """Gradients of sum(dout * out) with respect to every weight."""
    grads = [None] * 8
    d = np.asarray(dout, dtype=np.float64)
    inputs = [np.asarray(X, dtype=np.float64)] + hidden
    for l in (3, 2, 1, 0):
        grads[2 * l] = d.T @ inputs[l]
        grads[2 * l + 1] = d.sum(axis=0)
        if l > 0:
            d = (d @ weights[2 * l]) * (1.0 - hidden[l - 1] ** 2)
    return grads


def z_from_row(row, n):
    """Network output row (2 n^2 reals) -> n x n complex, Z(i, j) at 2 (i + n j)."""
    v = np.asarray(row, dtype=np.float64).reshape(n * n, 2)
    return (v[:, 0] + 1j * v[:, 1]).reshape(n, n, order="F")


def row_from_z(Z):
    v = np.asarray(Z).reshape(-1, order="F")
    return np.stack([v.real, v.imag], axis=1).reshape(-1)


def cholqr(Z):
    G = Z.conj().T @ Z
    R = np.linalg.cholesky(G).conj().T          # upper, real positive diagonal
    Q = np.linalg.solve(R.T, Z.T).T             # Z R^-1
    return Q, R


def cholqr_loss_grad(Z, C):
    """(loss, G): G is dloss/dZ in the form dloss = Re tr(G^H dZ)."""
    Q, R = cholqr(Z)
    cn = np.vdot(C, C).real
    M = Q.conj().T @ C @ Q
    off = M - np.diag(np.diag(M))
    loss = np.vdot(off, off).real / cn
    Gq = (4.0 / cn) * (C @ Q @ off)
    A = Q.conj().T @ Gq
    S = A - A.conj().T
    K = np.tril(S, -1) + np.diag(np.diag(S)) / 2.0
    T = Gq + Q @ (K - A)
    G = np.linalg.solve(R.conj(), T.T).T       # T R^-H
    return loss, G


def loss_only(Z, C):
    Q, _ = cholqr(Z)
    M = Q.conj().T @ C @ Q
    off = M - np.diag(np.diag(M))
    return np.vdot(off, off).real / np.vdot(C, C).real


def round_robin(n, r):
    """Disjoint (p, q) pairs of round r of n - 1 (circle method, n even)."""
    pos = [0] + [((i - 1 + r) % (n - 1)) + 1 for i in range(1, n)]
    return [tuple(sorted((pos[k], pos[n - 1 - k]))) for k in range(n // 2)]


def jacobi(C, Q0=None, tol=1e-6, max_sweeps=30):
    """Cyclic Jacobi on Q0^H C Q0 with the kernel's rotation (symmetric Schur 2 x 2
    after removing the phase of the off-diagonal); returns (eigenvalues
    unsorted, eigenvectors Q0 V, sweeps)."""
    n = C.shape[0]
    Q = np.eye(n, dtype=np.complex128) if Q0 is None else np.array(Q0, dtype=np.complex128)
    M = Q.conj().T @ C @ Q
    M = np.tril(M) + np.tril(M, -1).conj().T
    cn = np.vdot(C, C).real
    sweeps = 0
    while True:
        off = np.vdot(M, M).real - np.sum(np.abs(np.diag(M)) ** 2)
        if not off > tol * tol * cn or sweeps >= max_sweeps:
            break
        sweeps += 1
        for r in range(n - 1):
            U = np.eye(n, dtype=np.complex128)
            for p, q in round_robin(n, r):
                a, d, b = M[p, p].real, M[q, q].real, M[p, q]
                ab = abs(b)
                if ab > 1e-300:
                    tau = (d - a) / (2 * ab)
                    t = (1.0 if tau >= 0 else -1.0) / (abs(tau) + np.sqrt(1 + tau * tau))
                    c = 1 / np.sqrt(1 + t * t)
                    s = t * c
                    ph = np.conj(b) / ab                 # e^{-i phi}
                    U[p, p], U[p, q], U[q, p], U[q, q] = c, s, -s * ph, c * ph
            M = U.conj().T @ M @ U
            Q = Q @ U
    return np.diag(M).real.copy(), Q, sweeps
