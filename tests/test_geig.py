"""GEIG v1 fixture tooling (the reference's tests/lapack_reference.cpp and
tests/dump_pair.cpp): file I/O, the reference's input generator pinned byte
for byte against the reference's own fixtures, and dump_pair's report.
The GPU solve of `geig_tool.py make/check` is in test_gpu_geig.py."""
import glob
import importlib.util
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

PKG = os.path.join(ROOT, "sunny-batched-geneig_amd")


@pytest.fixture(scope="module")
def G():
    spec = importlib.util.spec_from_file_location("geig_under_test", os.path.join(PKG, "geig.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_mt19937_64_known_answer(G):
    """The C++ standard's check value ([rand.predef]): the 10000th output of a
    default-seeded (5489) mt19937_64 is 9981545732273789042."""
    r = G.MT19937_64(5489)
    for _ in range(9999):
        r.next_u64()
    assert r.next_u64() == 9981545732273789042


def test_canonical_and_polar_normal(G):
    """Uniforms in [0, 1) with 2^-64 resolution; polar normals have the right moments."""
    r = G.MT19937_64(7)
    u = np.array([r.canonical() for _ in range(20000)])
    assert (u >= 0).all() and (u < 1).all() and abs(u.mean() - 0.5) < 0.01
    g = G.PolarNormal(G.MT19937_64(11))
    z = np.array([g() for _ in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "ref_*.bin"))),
                         ids=lambda p: os.path.basename(p))
def test_reference_generator_matches_reference_fixtures(G, path):
    """lapack_reference's inputs for the fixture's seed, regenerated here, equal
    the reference's own H and S bit for bit (the committed files are the first
    pairs of the reference's seed-42 fixtures)."""
    d = G.read(path)
    mode = "random" if "random" in path else "near-identity"
    H, S = G.reference_pairs(d["n"], d["batch"], d["seed"], mode)
    assert H.tobytes() == d["H"].tobytes() and S.tobytes() == d["S"].tobytes()


def test_argument_order_matters(G):
    """Drawing the real part first gives different matrices: the test above pins
    the order, not just the distribution."""
    d = G.read(os.path.join(GOLDEN, "ref_n16_random.bin"))
    H, _ = G.reference_pairs(16, 1, d["seed"], "random", imag_first=False)
    assert not np.array_equal(H[0], d["H"][0])


def test_write_read_round_trip(G, tmp_path):
    rng = np.random.default_rng(3)
    b, n = 3, 5
    H = rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))
    S = rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))
    U = rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))
    W = rng.standard_normal((b, n))
    p = str(tmp_path / "x.bin")
    G.write(p, H, S, W, U, seed=99)
    d = G.read(p)
    assert (d["n"], d["batch"], d["seed"]) == (n, b, 99)
    for k, v in (("H", H), ("S", S), ("U", U), ("W", W)):
        assert np.array_equal(d[k], v)
    assert os.path.getsize(p) == 24 + 16 * 3 * b * n * n + 8 * b * n
    # column-major on disk: the first stored H entries are column 0 of pair 0
    with open(p, "rb") as f:
        f.seek(24)
        first = np.frombuffer(f.read(16 * n), dtype=np.complex128)
    assert np.array_equal(first, H[0][:, 0])


def test_read_rejects_bad_files(G, tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(struct.pack("<6i", 0x12345678, 1, 4, 1, 0, 0))
    with pytest.raises(IOError, match="magic"):
        G.read(str(p))
    p.write_bytes(struct.pack("<6i", G.MAGIC, 2, 4, 1, 0, 0))
    with pytest.raises(IOError, match="version"):
        G.read(str(p))
    p.write_bytes(struct.pack("<6i", G.MAGIC, 1, 4, 1, 0, 0) + b"\0" * 100)
    with pytest.raises(IOError, match="short read"):
        G.read(str(p))


def test_dump_report(G):
    """dump_pair's fields (tests/dump_pair.cpp:73-113) on a reference fixture."""
    path = os.path.join(GOLDEN, "ref_n54_random.bin")
    lines = G.dump(path)
    assert lines[1] == "magic:   0x47454947" and lines[3] == "N:       54" and lines[5] == "seed:    42"
    assert all("sorted_ascending=yes" in l for l in lines if l.startswith("batch "))
    assert lines[-1].startswith("max Hermiticity floor on H:      0.000e+00")


REF_DATA = "/root/reference/tests/data"
WU_FILES = sorted(glob.glob(os.path.join(GOLDEN, "ref_wu_n*_*.npz")))


@pytest.mark.parametrize("path", WU_FILES, ids=lambda p: os.path.basename(p))
def test_reference_wu_fixtures(G, path):
    """ref_wu_*.npz hold W and U of all 8 pairs of a reference fixture; with the
    inputs regenerated from the seed they replace the whole file.  Where the
    reference's files are present, the regenerated H, S of all 8 pairs equal
    the reference's bit for bit and W, U equal its stored outputs."""
    d = np.load(path)
    W, U = d["W"], d["U"]
    name = os.path.basename(path)[len("ref_wu_"):-len(".npz")]        # n{N}_{mode}
    n = int(name.split("_")[0][1:])
    assert W.shape == (8, n) and U.shape == (8, n, n) and W.dtype == np.float64 and U.dtype == np.complex128
    assert (np.diff(W, axis=1) >= 0).all()                              # ascending, as zhegvd returns them
    src = os.path.join(REF_DATA, "%s_b8_s42_%s.bin" % (name.split("_")[0], name.split("_")[1]))
    if not os.path.exists(src):
        pytest.skip("the reference's fixtures are not on this host")
    ref = G.read(src)
    mode = "random" if name.endswith("random") else "near-identity"
    H, S = G.reference_pairs(n, 8, 42, mode)
    assert H.tobytes() == ref["H"].tobytes() and S.tobytes() == ref["S"].tobytes()
    # geig.py returns matrices in index order (i, j); the npz keeps the
    # column-major storage order of the file and the device, [b, j, i]
    assert W.tobytes() == ref["W"].tobytes()
    assert U.tobytes() == np.ascontiguousarray(np.swapaxes(ref["U"], -1, -2)).tobytes()
