"""CPU pinning of the learned-warm-start oracle (oracle/learn_oracle.py):
the loss gradient and the network backward against central finite
differences, the Jacobi against numpy.linalg.eigh, the round-robin schedule
against its definition (every pair exactly once per sweep)."""
import numpy as np
import pytest

import learn_oracle as L


def herm(rng, n):
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return M + M.conj().T


@pytest.mark.parametrize("n,seed", [(4, 1), (6, 2), (9, 3)])
def test_loss_gradient_finite_differences(n, seed):
    rng = np.random.default_rng(seed)
    Z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    C = herm(rng, n)
    loss, G = L.cholqr_loss_grad(Z, C)
    assert abs(loss - L.loss_only(Z, C)) < 1e-14
    h = 1e-6
    FD = np.zeros((n, n), complex)
    for i in range(n):
        for j in range(n):
            for unit in (1.0, 1j):
                Zp, Zm = Z.copy(), Z.copy()
                Zp[i, j] += h * unit
                Zm[i, j] -= h * unit
                FD[i, j] += unit * (L.loss_only(Zp, C) - L.loss_only(Zm, C)) / (2 * h)
    assert np.abs(FD - G).max() <= 1e-6 * np.abs(G).max()


def test_loss_is_gauge_invariant_and_zero_at_eigenbasis():
    rng = np.random.default_rng(5)
    n = 7
    C = herm(rng, n)
    w, V = np.linalg.eigh(C)
    assert L.loss_only(V, C) < 1e-28
    # column phases, a permutation and a positive upper-triangular right factor change nothing
    Rt = np.triu(rng.standard_normal((n, n)) * 0.1) + np.diag(1 + rng.random(n))
    Z = (V * np.exp(1j * rng.random(n)))[:, rng.permutation(n)] @ Rt
    assert L.loss_only(Z, C) < 1e-24


def test_mlp_backward_finite_differences():
    rng = np.random.default_rng(7)
    dims = [8, 5, 4, 3, 6]
    w = []
    for l in range(4):
        w += [rng.standard_normal((dims[l + 1], dims[l])) * 0.5, rng.standard_normal(dims[l + 1]) * 0.1]
    X = rng.standard_normal((3, 8))
    dout = rng.standard_normal((3, 6))
    hidden, out = L.mlp_forward(X, w)
    grads = L.mlp_backward(X, w, hidden, dout)
    h = 1e-6
    for k in range(8):
        fd = np.zeros_like(w[k])
        for idx in np.ndindex(w[k].shape):
            wp = [a.copy() for a in w]
            wm = [a.copy() for a in w]
            wp[k][idx] += h
            wm[k][idx] -= h
            fd[idx] = (np.sum(dout * L.mlp_forward(X, wp)[1]) - np.sum(dout * L.mlp_forward(X, wm)[1])) / (2 * h)
        assert np.abs(fd - grads[k]).max() <= 1e-7 * max(1.0, np.abs(grads[k]).max()), k


def test_round_robin_covers_every_pair_once():
    n = 54
    seen = set()
    for r in range(n - 1):
        pairs = L.round_robin(n, r)
        idx = [i for p in pairs for i in p]
        assert sorted(idx) == list(range(n))
        seen.update(pairs)
    assert len(seen) == n * (n - 1) // 2


@pytest.mark.parametrize("n", [6, 12])
def test_jacobi_matches_eigh(n):
    rng = np.random.default_rng(n)
    C = herm(rng, n)
    w, Q, sweeps = L.jacobi(C, tol=1e-13)
    assert np.abs(np.sort(w) - np.linalg.eigvalsh(C)).max() <= 1e-12 * np.abs(C).max()
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() < 1e-13
    assert np.abs(C @ Q - Q * w).max() <= 1e-10 * np.abs(C).max()
    assert 3 <= sweeps <= 12
    # a good starting basis needs fewer sweeps
    _, V = np.linalg.eigh(C)
    E = rng.standard_normal((n, n)) * 1e-3
    Q0, _ = np.linalg.qr(V @ (np.eye(n) + E - E.T))
    assert L.jacobi(C, Q0, tol=1e-13)[2] < sweeps


def test_features_definition():
    p = np.array([[0.25, 0.5, 0.0, 1.1, 1e-2]])
    x = L.features(p)[0]
    assert np.allclose(x, [0, 1, -1, 0, 1, 0, 0.5, 0.0], atol=1e-12)
