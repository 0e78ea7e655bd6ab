"""The C++ interface include/geneig_solver.hpp, compiled and driven by a C++
caller written against the reference's spellings (tests/cpp/solver_hpp_driver.cpp;
reference surface src/nvrtc/nvrtc_solver.hpp:27-217, its tests
tests/test_nvrtc_n54.cpp and tests/test_nvrtc_multiN.cpp).

CPU: the header and driver compile with hipcc against libgeneig.so, and the
device-free checks pass (aliases, cache_dir / cubin_path, a constructor that
cannot reach a device throws std::runtime_error carrying the library's message).
GPU: the driver solves reference fixtures through NvrtcGeneigSolver::launch and
launch_chunked(make_streams(2)) (bit-identical), exercises the accessors and the
module cache, and its eigenpairs meet the fixtures' LAPACK solutions."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
DRIVER = os.path.join(CPP, "solver_hpp_driver")
GOLDEN = os.path.join(ROOT, "tests", "golden")
PKG = os.path.join(ROOT, "sunny-batched-geneig_amd")


def _checks(out):
    lines = [l for l in out.splitlines() if l.startswith("check ")]
    failed = [l for l in lines if " FAIL" in l]
    return lines, failed


def _build(out_path):
    subprocess.check_call(["make", "-s", "-B", "-C", CPP, "OUT=%s" % out_path])


def _env(tmp_path):
    # a driver built outside tests/cpp finds libgeneig.so through the loader path
    ld = os.environ.get("LD_LIBRARY_PATH", "")
    return dict(os.environ, SUNNY_GENEIG_CACHE_DIR=str(tmp_path / "cache"),
                LD_LIBRARY_PATH=PKG + (":" + ld if ld else ""))


def test_solver_hpp_compiles_and_device_free_checks(geneig, tmp_path):
    """hipcc compiles the header and the reference-spelled caller; the
    device-free checks pass (on a machine without a GPU the constructor's
    hipSetDevice fails and must surface as std::runtime_error)."""
    geneig.lib()                                   # libgeneig.so is built
    exe = str(tmp_path / "driver")
    _build(exe)
    env = _env(tmp_path)
    r = subprocess.run([exe, "info"], capture_output=True, text=True, env=env, timeout=120)
    print(r.stdout)
    lines, failed = _checks(r.stdout)
    assert r.returncode == 0 and not failed and len(lines) == 4, (r.stdout, r.stderr)
    assert str(tmp_path / "cache") in r.stdout    # cache_dir honours SUNNY_GENEIG_CACHE_DIR


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ref_n32_random.bin", "ref_n32_nearid.bin", "ref_n96_random.bin"])
def test_solver_hpp_solves_reference_fixture(geneig, oracle, tmp_path, name):
    from test_gpu_parity import check_pairs

    exe = DRIVER if os.path.exists(DRIVER) else str(tmp_path / "driver")
    if exe != DRIVER:
        _build(exe)
    fx = oracle.geig_read(os.path.join(GOLDEN, name))
    n, batch = fx["n"], fx["batch"]
    H = np.ascontiguousarray(fx["H"], dtype=np.complex128)
    S = np.ascontiguousarray(fx["S"], dtype=np.complex128)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        f.write(H.tobytes())
        f.write(S.tobytes())
    env = _env(tmp_path)
    print("C++ driver: %s n=%d batch=%d" % (name, n, batch), flush=True)
    r = subprocess.run([exe, "solve", str(n), str(batch), str(src), str(dst)], capture_output=True, text=True,
                       env=env, timeout=300)
    print(r.stdout, r.stderr)
    lines, failed = _checks(r.stdout)
    assert r.returncode == 0 and not failed, (r.stdout, r.stderr)
    names = {l.split()[1] for l in lines}
    for want in ("launch", "launch_chunked_bitwise", "block_dim", "second_solver_in_process", "cubin_path_on_disk",
                 "ctor_n129_throws", "make_streams", "destroy_streams"):
        assert want in names, want
    raw = open(dst, "rb").read()
    nw, nu = 8 * n * batch, 16 * n * n * batch
    assert len(raw) == nw + nu + 4 * batch
    W = np.frombuffer(raw[:nw], dtype=np.float64).reshape(batch, n)
    U = np.frombuffer(raw[nw:nw + nu], dtype=np.complex128).reshape(batch, n, n)
    info = np.frombuffer(raw[nw + nu:], dtype=np.int32)
    assert (info == 0).all()
    check_pairs(oracle, H, S, W, U, fx["W"], fx["U"], n, "C++ " + name)
