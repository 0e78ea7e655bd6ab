"""GEIG v1 fixture files and the reference's fixture generator.

GEIG v1 (tests/lapack_reference.cpp:252-275, read back by tests/dump_pair.cpp):
six little-endian int32 (magic 0x47454947 "GEIG", version 1, N, B, seed,
reserved 0), then H and S (B column-major N x N complex128 each), W (B x N
float64) and U (B x N x N complex128).

``reference_pairs`` regenerates the (H, S) batch that lapack_reference writes
for a seed: a 64-bit Mersenne Twister seeded with the seed
(tests/lapack_reference.cpp:162), uniform doubles as libstdc++'s
generate_canonical draws them from one 64-bit output, standard normals by
the Marsaglia polar method (one accepted point gives two deviates; the
second is kept for the next draw, and each matrix starts a fresh
distribution), and the three matrix recipes of tests/lapack_reference.cpp:84-150.
This is host-side input tooling, not the solve path; the solve goes through
the GPU library (``geig_tool.py``).
"""
import math
import struct

import numpy as np

MAGIC, VERSION = 0x47454947, 1
_HDR = struct.Struct("<6i")


def read(path):
    """Return dict(n, batch, seed, H, S, W, U) of a GEIG v1 file (arrays B x N x N, B x N)."""
    with open(path, "rb") as f:
        head = f.read(_HDR.size)
        if len(head) != _HDR.size:
            raise IOError("GEIG header read failed: %s" % path)
        magic, version, n, b, seed, _ = _HDR.unpack(head)
        if magic != MAGIC:
            raise IOError("bad magic 0x%08x (expected 0x%08x): %s" % (magic & 0xFFFFFFFF, MAGIC, path))
        if version != VERSION:
            raise IOError("unsupported version %d (expected %d): %s" % (version, VERSION, path))
        if n <= 0 or b <= 0:
            raise IOError("bad sizes N=%d B=%d: %s" % (n, b, path))
        mats = b * n * n

        def block(count, dtype, what):
            a = np.fromfile(f, dtype=dtype, count=count)
            if a.size != count:
                raise IOError("short read in %s: %s" % (what, path))
            return a

        H = block(mats, np.complex128, "H")
        S = block(mats, np.complex128, "S")
        W = block(b * n, np.float64, "W")
        U = block(mats, np.complex128, "U")
    # column-major n x n per pair: element (i, j) at i + j n, so the stored
    # block of pair p reshaped (n, n) is its transpose
    t = lambda a: a.reshape(b, n, n).transpose(0, 2, 1)
    return {"n": n, "batch": b, "seed": seed, "H": t(H), "S": t(S), "W": W.reshape(b, n), "U": t(U)}


def write(path, H, S, W, U, seed=0):
    """Write a GEIG v1 file from B x N x N (row, column) arrays and B x N eigenvalues."""
    H, S, U = (np.asarray(a, dtype=np.complex128) for a in (H, S, U))
    W = np.asarray(W, dtype=np.float64)
    b, n, _ = H.shape
    col = lambda a: np.ascontiguousarray(a.transpose(0, 2, 1))
    with open(path, "wb") as f:
        f.write(_HDR.pack(MAGIC, VERSION, n, b, int(seed), 0))
        for a in (col(H), col(S), np.ascontiguousarray(W), col(U)):
            a.tofile(f)


# ---- the reference's generator ---------------------------------------------------
class MT19937_64:
    """64-bit Mersenne Twister (Matsumoto & Nishimura; the std::mt19937_64 parameters)."""

    _W, _N, _M = 64, 312, 156
    _A = np.uint64(0xB5026F5AA96619E9)
    _UPPER, _LOWER = np.uint64(0xFFFFFFFF80000000), np.uint64(0x7FFFFFFF)

    def __init__(self, seed):
        mt = [0] * self._N
        mt[0] = seed & 0xFFFFFFFFFFFFFFFF
        for i in range(1, self._N):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & 0xFFFFFFFFFFFFFFFF
        self.mt = np.array(mt, dtype=np.uint64)
        self.buf = np.empty(0, dtype=np.uint64)
        self.pos = 0

    def _twist(self):
        mt, n, m = self.mt, self._N, self._M

        def step(lo, hi):
            x = (mt[lo:hi] & self._UPPER) | (mt[lo + 1:hi + 1] & self._LOWER)
            xa = x >> np.uint64(1)
            xa ^= np.where((x & np.uint64(1)) != 0, self._A, np.uint64(0))
            return xa

        # i < n - m reads mt[i + m] before it is rewritten; later i read the
        # rewritten mt[i + m - n]; i = n - 1 pairs with the new mt[0]
        mt[: n - m] = mt[m:n] ^ step(0, n - m)
        mt[n - m: n - 1] = mt[: m - 1] ^ step(n - m, n - 1)
        x = (mt[n - 1] & self._UPPER) | (mt[0] & self._LOWER)
        xa = (x >> np.uint64(1)) ^ (self._A if int(x) & 1 else np.uint64(0))
        mt[n - 1] = mt[m - 1] ^ xa
        y = mt.copy()
        y ^= (y >> np.uint64(29)) & np.uint64(0x5555555555555555)
        y ^= (y << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000)
        y ^= (y << np.uint64(37)) & np.uint64(0xFFF7EEE000000000)
        y ^= y >> np.uint64(43)
        self.buf, self.pos = y, 0

    def next_u64(self):
        if self.pos >= self.buf.size:
            self._twist()
        v = int(self.buf[self.pos])
        self.pos += 1
        return v

    def canonical(self):
        """generate_canonical<double, 53> from one 64-bit output: double(x) / 2^64, kept below 1."""
        u = float(self.next_u64()) / 18446744073709551616.0
        return u if u < 1.0 else math.nextafter(1.0, 0.0)


class PolarNormal:
    """Standard normal deviates by the Marsaglia polar method, two per accepted point."""

    def __init__(self, rng):
        self.rng, self.saved = rng, None

    def __call__(self):
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v
        while True:
            x = 2.0 * self.rng.canonical() - 1.0
            y = 2.0 * self.rng.canonical() - 1.0
            r2 = x * x + y * y
            if 0.0 < r2 <= 1.0:
                break
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        self.saved = x * mult
        return y * mult


def _gauss_matrix(n, rng, imag_first):
    """n x n complex normal entries, drawn column by column (column-major order)."""
    g = PolarNormal(rng)
    A = np.empty((n, n), dtype=np.complex128)
    for j in range(n):
        for i in range(n):
            if imag_first:
                im = g()
                re = g()
            else:
                re = g()
                im = g()
            A[i, j] = complex(re, im)
    return A


def _cmul_conj_acc(acc, a, b):
    """acc + a * conj(b) with the component order of a straightforward complex product."""
    c, d = b.real, -b.imag
    re = a.real * c - a.imag * d
    im = a.real * d + a.imag * c
    return (acc.real + re) + 1j * (acc.imag + im)


def reference_pairs(n, batch, seed, s_mode="random", imag_first=True):
    """(H, S), each batch x n x n, as tests/lapack_reference.cpp:161-172 generates them.

    imag_first: which part of cdouble(g(rng), g(rng)) is drawn first (C++ leaves
    the argument order unspecified; the compiler that made the reference's
    fixtures evaluated right to left)."""
    if s_mode not in ("random", "near-identity"):
        raise ValueError("s_mode must be random or near-identity")
    rng = MT19937_64(seed)
    H = np.empty((batch, n, n), dtype=np.complex128)
    S = np.empty_like(H)
    for b in range(batch):
        A = _gauss_matrix(n, rng, imag_first)
        H[b] = A + A.conj().T
        G = _gauss_matrix(n, rng, imag_first)
        if s_mode == "random":
            acc = np.zeros((n, n), dtype=np.complex128)
            for k in range(n):
                acc = _cmul_conj_acc(acc, G[:, k][:, None], G[:, k][None, :])
            acc[np.arange(n), np.arange(n)] += float(n)
            S[b] = acc
        else:
            M = 0.1 * (G + G.conj().T)
            M[np.arange(n), np.arange(n)] += 1.0 + 0.5 * n
            S[b] = M
    return H, S


def dump(path):
    """tests/dump_pair.cpp's report of a GEIG v1 file, as a list of lines."""
    d = read(path)
    n, b = d["n"], d["batch"]
    out = ["file:    %s" % path, "magic:   0x%08x" % MAGIC, "version: %d" % VERSION, "N:       %d" % n,
           "B:       %d" % b, "seed:    %d" % d["seed"]]
    herm = float(np.abs(d["H"] - d["H"].conj().transpose(0, 2, 1)).max())
    for p in range(b):
        w = d["W"][p]
        out.append("batch %d: eigenvalues [%.6e, %.6e]   sorted_ascending=%s"
                   % (p, w.min(), w.max(), "yes" if np.all(np.diff(w) >= 0) else "NO"))
    out.append("max |Im W|:                      %.3e (LAPACK W is real-typed; expect 0)" % 0.0)
    out.append("max Hermiticity floor on H:      %.3e (expect ~0)" % herm)
    return out
