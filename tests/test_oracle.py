"""CPU tests pinning the oracle (oracle/geneig_oracle.c) before it is trusted.

Each oracle function is checked against something it does not itself
compute: the reference's zhegvd fixtures, LAPACK (zpotrf / zhegst / zheev /
zhegv), closed forms, invariances and brute force, and every metric is
checked by planting a known error.
"""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, REFERENCE_DATA

EPS = np.finfo(np.float64).eps


def nat(M):
    """storage (column-major per pair) <-> natural index order."""
    return np.swapaxes(M, -1, -2)


def fmt(m):
    return {k: ("%.2e" % v if isinstance(v, float) else v) for k, v in m.items()}


# ---- RNG / generators ----------------------------------------------------------
def test_splitmix64_known_answers(oracle):
    # SplitMix64 reference sequence from state 0 (Steele, Lea, Flood 2014;
    # the values every published implementation reproduces)
    L = oracle.lib()
    state, outs = 0, []
    for _ in range(3):
        outs.append(L.go_splitmix64(state))
        state = (state + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_normal_moments(oracle):
    L = oracle.lib()
    import ctypes as C

    re, im = C.c_double(), C.c_double()
    xs = []
    for i in range(20000):
        L.go_cnormal(9, 0, 3, i, C.byref(re), C.byref(im))
        xs += [re.value, im.value]
    x = np.array(xs)
    assert abs(x.mean()) < 0.02 and abs(x.var() - 1.0) < 0.03
    assert abs(np.mean(x ** 4) - 3.0) < 0.15


def test_generic_family_properties(oracle):
    A, B = oracle.gen_generic(16, 4, seed=42)
    An, Bn = nat(A), nat(B)
    assert np.abs(An - np.conj(nat(An))).max() == 0.0
    assert np.abs(Bn - np.conj(nat(Bn))).max() == 0.0
    assert np.linalg.eigvalsh(Bn).min() >= 1.0 - 1e-12     # G G^H / n + I >= I
    A2, B2 = oracle.gen_generic(16, 2, seed=42, first_pair=2)
    assert np.array_equal(A2, A[2:]) and np.array_equal(B2, B[2:])


def test_spinwave_family_properties(oracle):
    A, B = oracle.gen_spinwave(27, 6, seed=1, first_q=100)
    Bn = nat(B)
    assert np.abs(Bn - np.conj(nat(Bn))).max() == 0.0
    lam = np.linalg.eigvalsh(Bn)[:, 0]
    t = np.array([oracle.lib().go_spinwave_target(1, 100 + q) for q in range(6)])
    assert np.allclose(lam, t, rtol=1e-9, atol=1e-13)
    assert ((t >= 1e-3) & (t <= 1e-1)).all()
    eta = np.diag(np.r_[np.ones(27), -np.ones(27)])
    assert np.array_equal(nat(A)[0], eta.astype(complex))
    # BdG structure: H = [[Aq, Bq], [conj Bq, conj Aq]], Bq symmetric
    H = Bn[0]
    Aq, Bq = H[:27, :27], H[:27, 27:]
    assert np.allclose(H[27:, 27:], np.conj(Aq)) and np.allclose(H[27:, :27], np.conj(Bq))
    assert np.allclose(Bq, Bq.T)


# ---- solver restatement vs the reference fixtures ----------------------------------
def fixture_paths():
    ref = sorted(glob.glob(os.path.join(REFERENCE_DATA, "n*_b8_s42_*.bin")))
    return ref if ref else sorted(glob.glob(os.path.join(GOLDEN, "ref_n*_*.bin")))


@pytest.mark.parametrize("path", fixture_paths(), ids=os.path.basename)
def test_oracle_vs_reference_fixture(oracle, path):
    """The reference's zhegvd outputs (tests/lapack_reference.cpp:186-194) at the
    reference's own tolerances (tests/test_full_n54.cpp:46-48), tightened."""
    fx = oracle.geig_read(path)
    W, X, info = oracle.zhegv(fx["H"], fx["S"])
    assert (info == 0).all()
    m = oracle.compare(fx["H"], fx["S"], W, X, fx["W"], fx["U"])
    n = fx["n"]
    print(os.path.basename(path), fmt(m))
    assert m["eig_rel_scaled"] < 1e-13
    assert m["phase_err"] < 1e-12
    assert m["resid"] < 5 * n * EPS and m["orth"] < 5 * n * EPS


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "own_*.bin"))), ids=os.path.basename)
def test_oracle_vs_own_golden(oracle, path):
    fx = oracle.geig_read(path)
    W, X, info = oracle.zhegv(fx["H"], fx["S"])
    assert (info == 0).all()
    m = oracle.compare(fx["H"], fx["S"], W, X, fx["W"], fx["U"])
    print(os.path.basename(path), fmt(m))
    n = fx["n"]
    assert m["eig_rel_scaled"] < 1e-12
    assert m["phase_err"] < 1e-9
    assert m["resid"] < 50 * n * EPS and m["orth"] < 50 * n * EPS


def test_golden_subset_matches_reference(oracle):
    """tests/golden/ref_* are byte-identical prefixes of the reference's fixtures."""
    if not os.path.isdir(REFERENCE_DATA):
        pytest.skip("reference fixtures not mounted")
    for p in sorted(glob.glob(os.path.join(GOLDEN, "ref_n*_*.bin"))):
        n, mode = os.path.basename(p)[5:-4].split("_")
        g = oracle.geig_read(p)
        r = oracle.geig_read(os.path.join(REFERENCE_DATA, "n%s_b8_s42_%s.bin" % (n[1:] if n[0] == "n" else n, mode)),
                             batch=g["batch"])
        for k in ("H", "S", "W", "U"):
            assert np.array_equal(g[k], r[k]), (p, k)


# ---- stage functions vs LAPACK -----------------------------------------------------
@pytest.fixture(scope="module")
def lapack(oracle):
    if oracle.lapack_load() is None:
        pytest.skip("no LAPACK library")
    return oracle


def test_zpotrf_vs_lapack(lapack):
    """tests/test_chol_n54.cpp:141-165 rule."""
    O = lapack
    fx = O.geig_read(os.path.join(GOLDEN, "ref_n54_random.bin"))
    for b in range(fx["batch"]):
        L, info = O.zpotrf(fx["S"][b])
        Lr, li = O.lapack_zpotrf(fx["S"][b])
        assert info == 0 and li == 0
        assert np.abs(L - Lr).max() / np.linalg.norm(Lr) < 1e-12
        Ln = nat(L)
        assert np.abs(Ln @ Ln.conj().T - nat(fx["S"][b])).max() < 1e-12 * np.abs(fx["S"][b]).max()


def test_zpotrf_identity_and_failure(oracle):
    I = np.eye(7, dtype=complex)
    L, info = oracle.zpotrf(I)
    assert info == 0 and np.array_equal(L, I)
    Bad = I.copy()
    Bad[3, 3] = -2.0
    L, info = oracle.zpotrf(Bad)
    assert info == 4                       # zpotrf numbering: leading minor of order 4
    A = np.eye(7, dtype=complex)[None]
    W, X, info = oracle.zhegv(A, Bad[None])
    assert info[0] == 7 + 4                # zhegv numbering: n + j
    assert np.isnan(W).all()


def test_reduce_vs_lapack_hegst(lapack):
    """tests/test_reduce_n54.cpp:127-160 rule (lower triangle, 1e-11)."""
    O = lapack
    fx = O.geig_read(os.path.join(GOLDEN, "ref_n54_nearid.bin"))
    for b in range(fx["batch"]):
        L, _ = O.zpotrf(fx["S"][b])
        Cm = O.reduce(fx["H"][b], L)
        Mr, li = O.lapack_zhegst(fx["H"][b], O.lapack_zpotrf(fx["S"][b])[0])
        assert li == 0
        mask = np.triu(np.ones_like(Cm, dtype=bool))   # storage upper = matrix lower
        rel = np.abs(np.where(mask, Cm - Mr, 0)).max() / np.linalg.norm(np.where(mask, Mr, 0))
        assert rel < 1e-11
        # eig(C) = eig(A, B)
        assert np.allclose(np.linalg.eigvalsh(nat(Cm)), fx["W"][b], rtol=0, atol=1e-12 * np.abs(fx["W"][b]).max())


def test_heev_vs_lapack(lapack):
    """tests/test_reduce_eig_n54.cpp:144-202 rule (1e-10 eigenvalues and phases)."""
    O = lapack
    A, _ = O.gen_generic(54, 3, seed=17)
    for b in range(3):
        W, V, info = O.heev(A[b])
        wr, Vr, li = O.lapack_zheev(A[b])
        assert info == 0 and li == 0
        assert np.max(np.abs(W - wr) / np.maximum(np.abs(wr), 1)) < 1e-13
        inner = np.abs(np.sum(np.conj(V) * Vr, axis=1))
        assert np.max(1 - inner) < 1e-12


def test_heev_closed_forms(oracle):
    # diagonal input -> sorted diagonal, identity-permutation vectors
    d = np.array([3.0, -1.0, 2.0, 0.5, -4.0])
    W, V, info = oracle.heev(np.diag(d).astype(complex))
    assert info == 0 and np.array_equal(W, np.sort(d))
    # 2x2 closed form: [[a, b],[conj b, c]] -> (a+c)/2 +- sqrt(((a-c)/2)^2 + |b|^2)
    a, c, b = 1.3, -0.7, 0.4 - 0.9j
    M = np.array([[a, b], [np.conj(b), c]])
    W, V, info = oracle.heev(nat(M))
    r = np.sqrt(((a - c) / 2) ** 2 + abs(b) ** 2)
    assert np.allclose(W, [(a + c) / 2 - r, (a + c) / 2 + r], atol=1e-15)
    # complex tridiagonal with a non-real last off-diagonal (the 1x1 phase reflector)
    T = np.diag([1.0, 2.0, 3.0]).astype(complex)
    T[2, 1], T[1, 2] = 0.5j, -0.5j
    W, V, info = oracle.heev(nat(T))
    assert np.allclose(W, np.linalg.eigvalsh(T), atol=1e-15)


def test_zhegv_vs_lapack_and_invariants(lapack):
    O = lapack
    n = 24
    A, B = O.gen_generic(n, 4, seed=5)
    W, X, info = O.zhegv(A, B)
    Wr, Xr, ir = O.lapack_zhegv(A, B)
    m = O.compare(A, B, W, X, Wr, Xr)
    assert m["eig_rel_scaled"] < 1e-13 and m["phase_err"] < 1e-12
    # A = alpha B -> all eigenvalues alpha
    W2, X2, _ = O.zhegv(2.5 * B, B)
    assert np.allclose(W2, 2.5, atol=1e-13)
    # (sA, tB) scales eigenvalues by s/t
    W3, _, _ = O.zhegv(3.0 * A, 0.5 * B)
    assert np.allclose(W3, 6.0 * W, atol=1e-12 * np.abs(W).max() * 6)
    # unitary congruence: (Q^H A Q, Q^H B Q) has the same eigenvalues
    Q, _ = np.linalg.qr(nat(O.gen_generic(n, 1, seed=99)[0])[0] + 1j * np.eye(n))
    An, Bn = nat(A[0]), nat(B[0])
    W4, _, _ = O.zhegv(nat(Q.conj().T @ An @ Q)[None], nat(Q.conj().T @ Bn @ Q)[None])
    assert np.allclose(W4[0], W[0], atol=1e-12 * np.abs(W[0]).max())


def test_zhegv_brute_force_small(oracle):
    """n <= 8 against scipy.linalg.eigh(a, b) (an independent LAPACK path)."""
    import scipy.linalg as sl

    for n in (1, 2, 3, 5, 8):
        A, B = oracle.gen_generic(n, 3, seed=n)
        W, X, info = oracle.zhegv(A, B)
        for b in range(3):
            w = sl.eigh(nat(A[b]), nat(B[b]), eigvals_only=True)
            assert np.allclose(W[b], w, atol=1e-13 * max(1, np.abs(w).max()))


def test_spinwave_pm_pairing(oracle):
    """(eta, H) pairs with H positive definite: eigenvalues come in +- pairs."""
    A, B = oracle.gen_spinwave(27, 8, seed=3)
    W, X, info = oracle.zhegv(A, B)
    assert (info == 0).all()
    assert np.abs(W + W[:, ::-1]).max() < 1e-9 * np.abs(W).max()
    assert ((W < 0).sum(axis=1) == 27).all()


# ---- metrics: planted errors -----------------------------------------------------------
def test_metrics_detect_planted_errors(oracle):
    A, B = oracle.gen_generic(16, 1, seed=8)
    W, X, _ = oracle.zhegv(A, B)
    base = oracle.compare(A[0], B[0], W[0], X[0], W[0], X[0])
    assert base["eig_rel_scaled"] == 0.0 and base["phase_err"] < 1e-13
    # eigenvalue off by 1e-9 relative
    W2 = W[0].copy()
    W2[5] *= 1 + 1e-9
    m = oracle.compare(A[0], B[0], W2, X[0], W[0], X[0])
    assert 0.5e-9 < m["eig_rel_scaled"] * max(1, 1 / abs(W[0][5])) < 2e-9 or m["eig_rel"] > 0.5e-9
    assert m["resid"] > 1e-12
    # column phase change is invisible, a rotated column is not
    X2 = X[0].copy()
    X2[3] *= np.exp(0.7j)
    m = oracle.compare(A[0], B[0], W[0], X2, W[0], X[0])
    assert m["phase_err"] < 1e-13 and m["resid"] < 1e-14
    X3 = X[0].copy()
    X3[3] = np.cos(1e-3) * X[0][3] + np.sin(1e-3) * X[0][4]
    m = oracle.compare(A[0], B[0], W[0], X3, W[0], X[0])
    assert m["phase_err"] > 1e-7 and m["orth"] > 1e-5
    # dropping the conjugate in one column breaks orthogonality
    X4 = X[0].copy()
    X4[2] = np.conj(X4[2])
    m = oracle.compare(A[0], B[0], W[0], X4, W[0], X[0])
    assert m["resid"] > 1e-6


def test_geig_roundtrip(oracle, tmp_path):
    A, B = oracle.gen_generic(5, 3, seed=1)
    W, X, _ = oracle.zhegv(A, B)
    p = str(tmp_path / "x.bin")
    oracle.geig_write(p, A, B, W, X, seed=77)
    hdr = oracle.geig_header(p)
    assert list(hdr) == [0x47454947, 1, 5, 3, 77, 0]
    r = oracle.geig_read(p)
    assert all(np.array_equal(r[k], v) for k, v in (("H", A), ("S", B), ("W", W), ("U", X)))
    assert os.path.getsize(p) == 24 + 3 * (3 * 25 * 16 + 5 * 8)


def test_fnv1a64_known_answers(oracle):
    # FNV-1a 64 published test vectors (Fowler/Noll/Vo)
    assert oracle.fnv1a64(b"") == 0xcbf29ce484222325
    assert oracle.fnv1a64(b"a") == 0xaf63dc4c8601ec8c
    assert oracle.fnv1a64(b"foobar") == 0x85944171f73967e8


def test_rayleigh_ld_pinned(oracle):
    """The long-double Rayleigh quotient used as the sharper reference at large
    cond(B): exact on a diagonal pair, equal to the eigenvalues for exact
    eigenvectors, and second order in an eigenvector perturbation (a first-order
    or transposed formula fails one of these)."""
    rng = np.random.default_rng(5)
    n = 12
    a = rng.standard_normal(n)
    b = rng.uniform(0.5, 2.0, n)
    A = np.diag(a).astype(np.complex128)[None]
    B = np.diag(b).astype(np.complex128)[None]
    X = (np.eye(n) * (1.0 + 0.5j)).astype(np.complex128)[None]   # scaled unit vectors
    R = oracle.rayleigh_ld(A, B, X)
    assert np.allclose(R[0], a / b, rtol=2e-16, atol=0.0)
    A, B = oracle.gen_generic_cond(24, 3, seed=2, cond=1e2)
    Wo, Xo, io = oracle.zhegv(A, B)
    R = oracle.rayleigh_ld(A, B, Xo)
    assert oracle.eig_err_scaled(R, Wo) < 1e-14
    # perturb the eigenvectors by ~1e-6: the quotients move by O(1e-12), not O(1e-6)
    P = Xo + 1e-6 * (rng.standard_normal(Xo.shape) + 1j * rng.standard_normal(Xo.shape)) * np.abs(Xo).max()
    Rp = oracle.rayleigh_ld(A, B, P)
    d = oracle.eig_err_scaled(Rp, Wo)
    assert 1e-15 < d < 1e-9, d


@pytest.mark.parametrize("n,cond", [(16, 1e2), (54, 1e4), (54, 1e8), (96, 1e6)])
def test_gen_generic_cond_hits_the_requested_condition(oracle, n, cond):
    """B = Q diag(logspace(0, log10 cond)) Q^H with Q unitary: cond(B) is the
    requested value (to rounding of the construction), its eigenvalues span
    [1, cond], and A stays Hermitian."""
    A, B = oracle.gen_generic_cond(n, 4, seed=9, cond=cond)
    lam = np.linalg.eigvalsh(np.swapaxes(B, -1, -2))
    got = lam[:, -1] / lam[:, 0]
    assert np.all(np.abs(got / cond - 1.0) < 1e-6 * max(1.0, cond * 1e-8)), got
    assert np.allclose(lam[:, 0], 1.0, rtol=1e-6) and np.allclose(lam[:, -1], cond, rtol=1e-6)
    assert np.abs(A - np.conj(np.swapaxes(A, -1, -2))).max() == 0.0


def test_rayleigh_ld_batch_equals_python_form(oracle):
    """The C batched long-double Rayleigh quotients agree with rayleigh_ld's
    numpy long-double evaluation (the same sums) to a few ulps of FP64, and on
    a diagonal pair equal the exact diagonal ratios."""
    A, B = oracle.gen_generic_cond(24, 6, seed=13, cond=1e6)
    W, X, info = oracle.zhegv(A, B)
    R1 = oracle.rayleigh_ld(A, B, X)
    R2 = oracle.rayleigh_ld_batch(A, B, X, nthreads=2)
    assert np.abs(R1 - R2).max() <= 4 * np.finfo(np.float64).eps * np.abs(R1).max()
    n = 5
    d = np.array([1.0, -2.0, 3.5, 0.25, 7.0])
    b = np.array([2.0, 1.0, 4.0, 0.5, 3.0])
    Ad = np.diag(d).astype(np.complex128)[None]
    Bd = np.diag(b).astype(np.complex128)[None]
    Xd = np.eye(n, dtype=np.complex128)[None] * (1.0 + 2.0j)
    assert np.array_equal(oracle.rayleigh_ld_batch(Ad, Bd, Xd)[0], d / b)
