"""GPU tests of the runtime-n solver contract (per-n hipRTC specialization,
module cache, autotuner, pairs-per-workgroup, chunked launches), mirroring the
reference's tests:
  tests/test_cache_n54.cpp          cold -> disk -> in-process -> no-cache, bit-identical outputs
  tests/test_cache_invalidation.cpp corrupt code object -> recompile and rewrite
  tests/test_autotune_n54.cpp       sweep runs once, then the cache is hit
  tests/test_bpb_sweep_n54.cpp      explicit pairs-per-workgroup sweep, same results
  tests/test_nvrtc_multiN.cpp       every fixture size through the runtime path
and NvrtcGeneigSolver::launch_chunked (src/nvrtc/nvrtc_solver.cpp:1078-1107),
plus its host-buffer variant with copy overlap (no reference counterpart).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def to_dev(a, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def run(geneig, solver, A, B, device):
    import torch

    n = A.shape[-1]
    W, U, info = geneig.empty_outputs(n, A.shape[0], device)
    solver.launch(A if not isinstance(A, np.ndarray) else to_dev(A, device),
                  B if not isinstance(B, np.ndarray) else to_dev(B, device), W, U, info)
    torch.cuda.synchronize()
    return W.cpu().numpy(), U.cpu().numpy(), info.cpu().numpy()


@pytest.fixture
def cache_dir(tmp_path, monkeypatch):
    d = tmp_path / "hsaco"
    monkeypatch.setenv("SUNNY_GENEIG_CACHE_DIR", str(d))
    return d


def test_cache_layers_bit_identical(geneig, device, cache_dir):
    """compile -> disk -> in-process, outputs bit-identical (test_cache_n54.cpp:201-350)."""
    A, B = geneig.gen_generic(40, 96, seed=17)
    s1 = geneig.GeneigSolver(40)
    assert s1.cache_layer_used() == "compile"
    files = os.listdir(cache_dir)
    assert any(f.endswith(".hsaco") for f in files), files
    r1 = run(geneig, s1, A, B, device)
    s2 = geneig.GeneigSolver(40)              # module still alive in-process
    assert s2.cache_layer_used() == "in-process"
    del s1, s2
    import gc

    gc.collect()
    s3 = geneig.GeneigSolver(40)              # weak entry expired -> disk
    assert s3.cache_layer_used() == "disk"
    r3 = run(geneig, s3, A, B, device)
    s4 = geneig.GeneigSolver(40, cache_mode=geneig.CACHE_NOCACHE)
    assert s4.cache_layer_used() == "compile"
    r4 = run(geneig, s4, A, B, device)
    for r in (r3, r4):
        assert r1[0].tobytes() == r[0].tobytes() and r1[1].tobytes() == r[1].tobytes()
    print("acquire ms: disk %.1f" % s3.acquire_module_ms(), flush=True)


def test_cache_invalidation_recompiles(geneig, device, cache_dir):
    """Corrupt ELF magic -> discarded, recompiled, rewritten (test_cache_invalidation.cpp:238-312)."""
    s = geneig.GeneigSolver(24)
    del s
    import gc

    gc.collect()
    paths = [os.path.join(cache_dir, f) for f in os.listdir(cache_dir) if f.endswith(".hsaco")]
    assert paths
    for path in paths:
        with open(path, "r+b") as f:
            f.write(b"XXXX")
    s = geneig.GeneigSolver(24)
    assert s.cache_layer_used() == "compile"
    for path in paths:
        with open(path, "rb") as f:
            assert f.read(4) == b"\x7fELF"
    # a truncated file is rejected too
    del s
    gc.collect()
    for path in paths:
        with open(path, "r+b") as f:
            f.truncate(100)
    s = geneig.GeneigSolver(24)
    assert s.cache_layer_used() == "compile"
    A, B = geneig.gen_generic(24, 8, seed=2)
    W, U, info = run(geneig, s, A, B, device)
    assert (info == 0).all()


def test_autotune_then_cache_hit(geneig, oracle, device, cache_dir):
    """Sweep on first construction, winner reused after (test_autotune_n54.cpp:308-346).
    The sweep picks threads per pair (64: one wave, swept over pairs per
    workgroup; 256: four waves, one pair per workgroup), then, for one wave
    per pair, K1's row-chunk width (the tile shape of its register walks) at
    the winning pairs per workgroup; all are persisted and reused."""
    s1 = geneig.GeneigSolver(32, tuning_mode=geneig.TUNING_AUTOTUNE)
    assert s1.was_autotuned() and s1.autotune_sweep_ran()
    bdx = s1.autotuned_block_dim_x()
    assert bdx in (64, 256) and s1.block_dim()[0] == bdx
    assert s1.batches_per_block() in ((1, 2, 4, 8) if bdx == 64 else (1,))
    assert s1.k1_chunk_rows() in ((4, 8, 16) if bdx == 64 else (0,))
    s2 = geneig.GeneigSolver(32, tuning_mode=geneig.TUNING_AUTOTUNE)
    assert s2.was_autotuned() and not s2.autotune_sweep_ran()
    assert s2.batches_per_block() == s1.batches_per_block()
    assert s2.k1_chunk_rows() == s1.k1_chunk_rows()
    assert s2.block_dim() == s1.block_dim()
    tune = [f for f in os.listdir(cache_dir) if f.startswith("autotune_32_")]
    assert len(tune) == 1
    fields = open(os.path.join(cache_dir, tune[0])).read().split()
    print("autotune record (bpb variant pairs/s k1_rows block_dim_x):", fields, flush=True)
    assert len(fields) == 5 and int(fields[0]) == s1.batches_per_block() and int(fields[4]) == bdx
    assert int(fields[3]) in (0, 4, 16)
    if bdx == 64:
        assert (int(fields[3]) or 8) == s1.k1_chunk_rows()
    A, B = oracle.gen_generic(32, 16, seed=3)
    W, X, info = run(geneig, s2, A, B, device)
    Wo, Xo, _ = oracle.zhegv(A, B)
    m = oracle.compare(A, B, W, X, Wo, Xo)
    assert m["eig_rel_scaled"] < 1e-12 and m["phase_err"] < 1e-10


@pytest.mark.parametrize("n,bpb", [(54, 1), (54, 2), (24, 4), (24, 8)])
def test_bpb_sweep_same_results(geneig, oracle, device, cache_dir, n, bpb):
    """Explicit pairs-per-workgroup, including a partial last workgroup
    (the reference required padding; test_bpb_sweep_n54.cpp:234-338)."""
    A, B = oracle.gen_generic(n, 13, seed=21)
    s = geneig.GeneigSolver(n, batches_per_block_request=bpb)
    assert s.batches_per_block() == bpb and s.block_dim() == (64, bpb, 1)
    W, X, info = run(geneig, s, A, B, device)
    assert (info == 0).all()
    Wo, Xo, _ = oracle.zhegv(A, B)
    m = oracle.compare(A, B, W, X, Wo, Xo)
    assert m["eig_rel_scaled"] < 1e-12 and m["phase_err"] < 1e-10


def test_bpb_that_does_not_fit_raises(geneig, device, cache_dir):
    """An explicit pairs-per-workgroup beyond the LDS budget is refused, as the
    reference refuses an explicit BPB that does not fit (nvrtc_solver.cpp:1021-1029)."""
    with pytest.raises(geneig.GeneigError, match="LDS"):
        geneig.GeneigSolver(54, batches_per_block_request=8)   # 8 x 26.3 KB > 160 KB


def test_runtime_matches_static_n54(geneig, oracle, device, cache_dir):
    """The hipRTC module (compiled for the device's full target) and the hipcc
    build of the same source agree on the eigenvalues to rounding and both
    meet the FP64 bar against the oracle."""
    import torch

    from test_gpu_parity import check_pairs

    A, B = geneig.gen_spinwave(27, 64, seed=4)
    W1, U1, i1 = geneig.empty_outputs(54, 64, device)
    geneig.full_n54_batched(A, B, W1, U1, i1)
    W2, U2, i2 = run(geneig, geneig.GeneigSolver(54, batches_per_block_request=1), A, B, device)
    torch.cuda.synchronize()
    W1h, U1h = W1.cpu().numpy(), U1.cpu().numpy()
    dw = np.abs(W1h - W2).max() / np.abs(W1h).max()
    print("static vs hipRTC FP64: max rel diff W %.2e" % dw, flush=True)
    assert dw <= 1e-13
    Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
    Wo, Xo, _ = oracle.zhegv(Ah, Bh)
    check_pairs(oracle, Ah, Bh, W1h, U1h, Wo, Xo, 54, "static")
    check_pairs(oracle, Ah, Bh, W2, U2, Wo, Xo, 54, "hipRTC")


def test_launch_chunked_streams(geneig, oracle, device, cache_dir):
    """Chunks round-robined over streams give the same outputs as one launch."""
    import torch

    A, B = geneig.gen_generic(16, 100, seed=8)
    s = geneig.GeneigSolver(16)
    W0, U0, i0 = run(geneig, s, A, B, device)
    W, U, info = geneig.empty_outputs(16, 100, device)
    streams = geneig.make_streams(3)
    torch.cuda.synchronize()
    s.launch_chunked(A, B, W, U, info, total=100, chunk_size=17, streams=streams)
    torch.cuda.synchronize()
    assert np.array_equal(W.cpu().numpy(), W0) and np.array_equal(U.cpu().numpy(), U0)


@pytest.mark.parametrize("pinned", [True, False])
def test_launch_chunked_host_matches_device(geneig, oracle, device, cache_dir, pinned):
    """Host-resident pairs through launch_chunked_host (copy in, solve, copy out per
    chunk on its stream) give the device launch's outputs bit for bit; pageable
    host buffers are slower but equal."""
    import torch

    n, total = 32, 300
    A, B = geneig.gen_generic(n, total, seed=4)
    s = geneig.GeneigSolver(n)
    W0, U0, i0 = run(geneig, s, A, B, device)
    Ah = A.cpu().pin_memory() if pinned else A.cpu()
    Bh = B.cpu().pin_memory() if pinned else B.cpu()
    W = torch.full((total, n), float("nan"), dtype=torch.float64, pin_memory=pinned)
    U = torch.full((total, n, n), float("nan"), dtype=torch.complex128, pin_memory=pinned)
    info = torch.full((total,), -1, dtype=torch.int32, pin_memory=pinned)
    streams = geneig.make_streams(3)
    torch.cuda.synchronize()
    for rep in range(2):   # the second call reuses the per-stream staging slots
        s.launch_chunked_host(Ah, Bh, W, U, info, total=total, chunk_size=41, streams=streams)
        for st in streams:
            st.synchronize()
        assert np.array_equal(W.numpy(), W0) and np.array_equal(U.numpy(), U0), rep
        assert np.array_equal(info.numpy(), i0) and (info.numpy() == 0).all()
    with pytest.raises(geneig.GeneigError, match="host tensor"):
        s.launch_chunked_host(A, B, W, U, info, total=total, chunk_size=41, streams=streams)


def test_workspace_chunking_small_budget(geneig, oracle, device):
    """A batch larger than one workspace chunk goes through several K1-K2-K3 rounds."""
    import subprocess
    import sys

    code = r'''
import os, sys, numpy as np, torch
sys.path.insert(0, %r); sys.path.insert(0, %r)
import conftest, oracle as O
g = conftest.load_package()
A, B = g.gen_generic(54, 700, seed=12)
W, U, info = g.empty_outputs(54, 700, "cuda")
g.full_n54_batched(A, B, W, U, info); torch.cuda.synchronize()
Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
Wo, Xo, io = O.zhegv(Ah, Bh)
m = O.compare(Ah, Bh, W.cpu().numpy(), U.cpu().numpy(), Wo, Xo)
assert int(info.abs().max()) == 0 and m["eig_rel_scaled"] < 1e-12 and m["phase_err"] < 1e-10, m
print("ok", m["eig_rel_scaled"])
''' % (os.path.join(os.path.dirname(GOLDEN)), os.path.join(os.path.dirname(os.path.dirname(GOLDEN)), "oracle"))
    env = dict(os.environ, GENEIG_WORKSPACE_MB="16")   # 16 MiB -> chunks of 141 pairs at n = 54
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr[-2000:], flush=True)
    assert out.returncode == 0 and "ok" in out.stdout


def test_largest_size_solves_and_larger_sizes_are_refused(geneig, oracle, device, cache_dir):
    """The one-wave kernels keep the pair's n x LD working matrix in LDS above
    n = 64 (16 n LD bytes in FP64, LD = n + 1 for even n, plus the two
    broadcast vectors): n = 98 needs 159,840 B and fits the 160 KB a
    workgroup may use, n = 100 needs 166,208 B and does not, so a solver
    forced to one wave per pair is refused there with GENEIG_E_FIT (code 4).
    The four-wave kernels (the default above n = 64) hold the matrix in
    registers and reach the reference's n = 128 (tests/test_nvrtc_multiN.cpp:
    218-271).  The precision form's refinement keeps X (n x LD complex) in LDS:
    n = 100 still fits (163,084 B), n = 128 is refused."""
    from test_gpu_parity import check_pairs

    n = 98
    A, B = geneig.gen_generic(n, 16, seed=23)
    Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
    Wo, Xo, io = oracle.zhegv(Ah, Bh)
    assert (io == 0).all()
    for tag, s in (("fp64 one wave", geneig.GeneigSolver(n, force_block_dim_x=64)),
                   ("precision", geneig.GeneigMixedSolver(n))):
        print("n=%d %s: %d B LDS per workgroup" % (n, tag, s.shared_mem_bytes()), flush=True)
        W, X, info = run(geneig, s, A, B, device)
        assert (info == 0).all()
        check_pairs(oracle, Ah, Bh, W, X, Wo, Xo, n, "n=%d %s" % (n, tag))
    for big in (100, 128):
        Ah, Bh = oracle.gen_generic(big, 4, seed=big)
        Wo, Xo, io = oracle.zhegv(Ah, Bh)
        s = geneig.GeneigSolver(big)
        assert s.block_dim() == (256, 1, 1)
        W, X, info = run(geneig, s, Ah, Bh, device)
        assert (info == 0).all()
        check_pairs(oracle, Ah, Bh, W, X, Wo, Xo, big, "n=%d fp64 four waves" % big)
        refused = [lambda m: geneig.GeneigSolver(m, force_block_dim_x=64)]
        if big == 100:      # the refinement's X (100 x 101 complex) still fits: 163,084 B
            sm = geneig.GeneigMixedSolver(big)
            print("n=%d precision: %d B LDS per workgroup" % (big, sm.shared_mem_bytes()), flush=True)
            W, X, info = run(geneig, sm, Ah, Bh, device)
            assert (info == 0).all()
            check_pairs(oracle, Ah, Bh, W, X, Wo, Xo, big, "n=%d precision" % big)
        else:
            refused.append(geneig.GeneigMixedSolver)
        for make in refused:
            with pytest.raises(geneig.GeneigError) as e:
                make(big)
            print("n=%d refused: %s" % (big, e.value), flush=True)
            assert "(code 4)" in str(e.value) and "LDS" in str(e.value)
