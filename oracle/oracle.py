"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() (to
check its result) and by bench.py's cpu_baseline leg.  The product library never
imports it.  Arrays are numpy, column-major per pair: a batch of n x n
matrices is an array of shape (batch, n, n) whose pair b, element (i, j) is
``M[b, j, i]`` (i.e. each pair stored as its transpose in C order), exactly
the device layout of the reference (src/nvrtc/kernel_source.hpp:200-205).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

LAPACK_CANDIDATES = [
    "/opt/conda/lib/libmkl_rt.so",
    "/opt/conda/lib/libmkl_rt.so.1",
]


def _scipy_openblas():
    try:
        import scipy

        d = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        for f in sorted(os.listdir(d)):
            if f.startswith("libscipy_openblas") and f.endswith(".so"):
                return os.path.join(d, f)
    except Exception:
        pass
    return None


class Metrics(C.Structure):
    _fields_ = [
        ("eig_rel_scaled", C.c_double),
        ("eig_rel", C.c_double),
        ("phase_err", C.c_double),
        ("phase_err_degen", C.c_double),
        ("degen_excluded", C.c_int),
        ("resid", C.c_double),
        ("orth", C.c_double),
        ("resid_raw", C.c_double),
        ("orth_raw", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Bond(C.Structure):
    _fields_ = [("shift", C.c_int), ("r", C.c_int * 3), ("J", C.c_double)]


_lib = None


def build():
    import subprocess

    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u64, i32, dbl = C.c_uint64, C.c_int, C.c_double
    sig = {
        "go_splitmix64": (u64, [u64]),
        "go_rng_key": (u64, [u64, C.c_uint32, u64, C.c_uint32]),
        "go_u01": (dbl, [u64, C.c_uint32, u64, C.c_uint32]),
        "go_cnormal": (None, [u64, C.c_uint32, u64, C.c_uint32, P, P]),
        "go_gen_generic": (None, [i32, i32, u64, u64, P, P]),
        "go_gen_generic_cond": (None, [i32, i32, u64, u64, dbl, P, P]),
        "go_spinwave_bonds": (None, [i32, i32, u64, P]),
        "go_spinwave_q": (None, [i32, i32, i32, u64, P]),
        "go_spinwave_H0": (None, [i32, i32, P, u64, dbl, P, P]),
        "go_spinwave_J": (dbl, [u64, i32, i32]),
        "go_residuals4": (None, [i32, P, P, P, P, P]),
        "go_spinwave_target": (dbl, [u64, u64]),
        "go_gen_spinwave": (i32, [i32, i32, u64, i32, i32, i32, u64, i32, dbl, P, P]),
        "go_zpotrf_lower": (i32, [i32, P, P]),
        "go_reduce": (None, [i32, P, P, P]),
        "go_heev": (i32, [i32, P, P, P]),
        "go_heevals": (i32, [i32, P, P]),
        "go_backtransform": (None, [i32, P, P, P]),
        "go_zhegv": (i32, [i32, P, P, P, P]),
        "go_zhegv_batch": (i32, [i32, i32, P, P, P, P, P, i32]),
        "go_lapack_load": (i32, [C.c_char_p]),
        "go_lapack_name": (C.c_char_p, []),
        "go_lapack_zhegv": (i32, [i32, P, P, P, P]),
        "go_lapack_zhegv_batch": (i32, [i32, i32, P, P, P, P, P, i32]),
        "go_lapack_zpotrf": (i32, [i32, P]),
        "go_lapack_zhegst": (i32, [i32, P, P]),
        "go_lapack_zheev": (i32, [i32, P, P]),
        "go_residuals": (None, [i32, P, P, P, P, P, P]),
        "go_compare": (None, [i32, P, P, P, P, P, P, dbl, P]),
        "go_compare_batch": (None, [i32, i32, P, P, P, P, P, P, dbl, i32, P]),
        "go_geig_read_header": (i32, [C.c_char_p, P]),
        "go_geig_read": (i32, [C.c_char_p, i32, i32, P, P, P, P]),
        "go_geig_write": (i32, [C.c_char_p, i32, i32, i32, P, P, P, P]),
        "go_fnv1a64": (u64, [P, u64, u64]),
        "go_rayleigh_ld_batch": (None, [i32, i32, P, P, P, P, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def to_colmajor(M):
    """(batch, n, n) natural-index array -> storage array (pair stored column-major)."""
    return np.ascontiguousarray(np.swapaxes(M, -1, -2))


from_colmajor = to_colmajor


# ---- generators -----------------------------------------------------------
def gen_generic(n, batch, seed, first_pair=0):
    A = np.empty((batch, n, n), np.complex128)
    B = np.empty((batch, n, n), np.complex128)
    lib().go_gen_generic(n, batch, seed, first_pair, _p(A), _p(B))
    return A, B


def gen_generic_cond(n, batch, seed, cond, first_pair=0):
    A = np.empty((batch, n, n), np.complex128)
    B = np.empty((batch, n, n), np.complex128)
    lib().go_gen_generic_cond(n, batch, seed, first_pair, float(cond), _p(A), _p(B))
    return A, B


def gen_spinwave(nsub, batch, seed, grid=(64, 32, 32), first_q=0, nbonds=12, jscale=1.0):
    n = 2 * nsub
    A = np.empty((batch, n, n), np.complex128)
    B = np.empty((batch, n, n), np.complex128)
    bad = lib().go_gen_spinwave(nsub, nbonds, seed, grid[0], grid[1], grid[2],
                                first_q, batch, float(jscale), _p(A), _p(B))
    if bad:
        raise RuntimeError("spin-wave generator: eigenvalue shift failed for %d pairs" % bad)
    return A, B


def spinwave_bonds(nsub, seed, nbonds=12):
    arr = (Bond * nbonds)()
    lib().go_spinwave_bonds(nsub, nbonds, seed, C.cast(arr, C.c_void_p))
    return [(b.shift, tuple(b.r), b.J) for b in arr]


def spinwave_H0(nsub, seed, q, nbonds=12, jscale=1.0):
    arr = (Bond * nbonds)()
    lib().go_spinwave_bonds(nsub, nbonds, seed, C.cast(arr, C.c_void_p))
    qa = np.asarray(q, np.float64)
    H = np.empty((2 * nsub, 2 * nsub), np.complex128)
    lib().go_spinwave_H0(nsub, nbonds, C.cast(arr, C.c_void_p), seed, float(jscale), _p(qa), _p(H))
    return H


# ---- solver ----------------------------------------------------------------
def zhegv(A, B, nthreads=0):
    """Oracle solve of a batch; returns (W, X, info)."""
    A = _c128(A)
    B = _c128(B)
    batch, n, _ = A.shape
    W = np.empty((batch, n), np.float64)
    X = np.empty_like(A)
    info = np.empty(batch, np.int32)
    lib().go_zhegv_batch(n, batch, _p(A), _p(B), _p(W), _p(X), _p(info), nthreads)
    return W, X, info


def zpotrf(B):
    B = _c128(B)
    n = B.shape[-1]
    L = np.empty_like(B)
    info = lib().go_zpotrf_lower(n, _p(B), _p(L))
    return L, info


def reduce(A, L):
    A = _c128(A)
    L = _c128(L)
    Cm = np.empty_like(A)
    lib().go_reduce(A.shape[-1], _p(A), _p(L), _p(Cm))
    return Cm


def heev(Cm):
    Cm = _c128(Cm)
    n = Cm.shape[-1]
    W = np.empty(n)
    V = np.empty_like(Cm)
    info = lib().go_heev(n, _p(Cm), _p(W), _p(V))
    return W, V, info


def backtransform(L, V):
    L = _c128(L)
    V = _c128(V)
    X = np.empty_like(V)
    lib().go_backtransform(V.shape[-1], _p(L), _p(V), _p(X))
    return X


# ---- LAPACK -----------------------------------------------------------------
def lapack_load():
    for p in LAPACK_CANDIDATES + [_scipy_openblas()]:
        if p and os.path.exists(p) and lib().go_lapack_load(p.encode()) == 0:
            return p
    return None


def lapack_zhegv(A, B, nthreads=0):
    if lapack_load() is None:
        raise RuntimeError("no LAPACK library with zhegv found")
    A = _c128(A)
    B = _c128(B)
    batch, n, _ = A.shape
    W = np.empty((batch, n), np.float64)
    X = np.empty_like(A)
    info = np.empty(batch, np.int32)
    lib().go_lapack_zhegv_batch(n, batch, _p(A), _p(B), _p(W), _p(X), _p(info), nthreads)
    return W, X, info


def lapack_zpotrf(B):
    lapack_load()
    a = _c128(B).copy()
    info = lib().go_lapack_zpotrf(a.shape[-1], _p(a))
    return a, info


def lapack_zhegst(A, L):
    lapack_load()
    a = _c128(A).copy()
    info = lib().go_lapack_zhegst(a.shape[-1], _p(a), _p(_c128(L)))
    return a, info


def lapack_zheev(Cm):
    lapack_load()
    a = _c128(Cm).copy()
    n = a.shape[-1]
    w = np.empty(n)
    info = lib().go_lapack_zheev(n, _p(a), _p(w))
    return w, a, info


# ---- metrics -----------------------------------------------------------------
def compare(A, B, W, X, Wr, Xr=None, degen_gap=1e-6, nthreads=0):
    A, B, X = _c128(A), _c128(B), _c128(X)
    W = np.ascontiguousarray(W, np.float64)
    Wr = np.ascontiguousarray(Wr, np.float64)
    Xr = None if Xr is None else _c128(Xr)
    m = Metrics()
    if A.ndim == 2:
        lib().go_compare(A.shape[-1], _p(A), _p(B), _p(W), _p(X), _p(Wr), _p(Xr),
                         degen_gap, C.byref(m))
    else:
        lib().go_compare_batch(A.shape[-1], A.shape[0], _p(A), _p(B), _p(W), _p(X),
                               _p(Wr), _p(Xr), degen_gap, nthreads, C.byref(m))
    return m.as_dict()


def rayleigh_ld(A, B, X):
    """Rayleigh quotients x_j^H A x_j / x_j^H B x_j of every column of X, per
    pair, evaluated in long double (64-bit significand) and returned as float64
    (batch x n).  A checker with no reference counterpart: the quotient is
    stationary at an eigenvector, so its error is second order in the error of
    x_j, and with the oracle's eigenvectors it gives eigenvalues much closer to
    exact than any FP64 eigenvalue when B is ill-conditioned (there the FP64
    eigenvalues, the oracle's included, carry errors ~ eps cond(B))."""
    A, B, X = _c128(A), _c128(B), _c128(X)
    out = np.empty(X.shape[:2], np.float64)
    for p in range(A.shape[0]):
        Am = A[p].T.astype(np.clongdouble)        # pair p stored as its transpose
        Bm = B[p].T.astype(np.clongdouble)
        Xm = X[p].T.astype(np.clongdouble)        # columns = eigenvectors
        num = (Xm.conj() * (Am @ Xm)).sum(axis=0).real
        den = (Xm.conj() * (Bm @ Xm)).sum(axis=0).real
        out[p] = (num / den).astype(np.float64)
    return out


def rayleigh_ld_batch(A, B, X, nthreads=0):
    """rayleigh_ld in C (long double, one pair per OpenMP thread): the same
    quotients for many pairs at once (batch x n, float64)."""
    A, B, X = _c128(A), _c128(B), _c128(X)
    out = np.empty(X.shape[:2], np.float64)
    lib().go_rayleigh_ld_batch(X.shape[-1], X.shape[0], _p(A), _p(B), _p(X), _p(out), int(nthreads))
    return out


def eig_err_scaled(W, Wr):
    """max_i |W_i - Wr_i| / max(|Wr_i|, 1) over a batch (the oracle's eig_rel_scaled)."""
    W, Wr = np.asarray(W, np.float64), np.asarray(Wr, np.float64)
    return float((np.abs(W - Wr) / np.maximum(np.abs(Wr), 1.0)).max())


def residuals(A, B, W, X):
    """(resid, orth, resid_raw, orth_raw) of one pair."""
    A, B, X = _c128(A), _c128(B), _c128(X)
    W = np.ascontiguousarray(W, np.float64)
    o = np.zeros(4)
    lib().go_residuals4(A.shape[-1], _p(A), _p(B), _p(W), _p(X), _p(o))
    return tuple(float(v) for v in o)


# ---- GEIG v1 -----------------------------------------------------------------
def geig_header(path):
    hdr = np.zeros(6, np.int32)
    rc = lib().go_geig_read_header(path.encode(), _p(hdr))
    if rc:
        raise IOError("GEIG header read failed (%d): %s" % (rc, path))
    return hdr


def geig_read(path, batch=None):
    hdr = geig_header(path)
    n, b = int(hdr[2]), int(hdr[3])
    if batch is None:
        batch = b
    H = np.empty((batch, n, n), np.complex128)
    S = np.empty_like(H)
    U = np.empty_like(H)
    W = np.empty((batch, n), np.float64)
    rc = lib().go_geig_read(path.encode(), n, batch, _p(H), _p(S), _p(W), _p(U))
    if rc:
        raise IOError("GEIG read failed (%d): %s" % (rc, path))
    return dict(n=n, batch=batch, seed=int(hdr[4]), H=H, S=S, W=W, U=U)


def geig_write(path, H, S, W, U, seed=0):
    H, S, U = _c128(H), _c128(S), _c128(U)
    W = np.ascontiguousarray(W, np.float64)
    rc = lib().go_geig_write(path.encode(), H.shape[-1], H.shape[0], seed,
                             _p(H), _p(S), _p(W), _p(U))
    if rc:
        raise IOError("GEIG write failed (%d): %s" % (rc, path))


def fnv1a64(data: bytes, seed=0):
    buf = C.create_string_buffer(data, len(data))
    return lib().go_fnv1a64(C.cast(buf, C.c_void_p), len(data), seed)
