"""The reference's own fixtures and synthetic batches, restated: every pair of
its twelve seed-42 fixtures (inputs regenerated bit for bit by geig.py,
outputs the stored zhegvd W, U), the 1,024-pair n = 54 batch of
tests/test_full_batched_n54.cpp:366-367 (seed 0xC0FFEE0042), its missing
n54_b32_s123_random fixture (tests/test_full_batched_n54.cpp:344, inputs
regenerated, checked in batches of 1, 8 and 32 as there) and the 1,024-pair
n = 32 batch of tests/test_blockdim_n32.cpp:197 (seed 0xBEEFC0DE32), each
against LAPACK zhegv on every pair.  The reference's bars are 1e-10 on the
eigenvalues and eigenvector phases (tests/test_full_batched_n54.cpp:44-46);
these tests hold the tighter bars of the rest of the suite."""
import glob
import importlib.util
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from test_gpu_parity import check_pairs, solve_batched

pytestmark = pytest.mark.gpu

WU_FILES = sorted(glob.glob(os.path.join(GOLDEN, "ref_wu_n*_*.npz")))


@pytest.fixture(scope="module")
def G():
    spec = importlib.util.spec_from_file_location("geig_for_gpu_tests",
                                                  os.path.join(ROOT, "sunny-batched-geneig_amd", "geig.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def pairs(G, n, batch, seed, mode):
    """geig.reference_pairs in the device's column-major storage ([b, j, i])."""
    H, S = G.reference_pairs(n, batch, seed, mode)
    return (np.ascontiguousarray(np.swapaxes(H, -1, -2)), np.ascontiguousarray(np.swapaxes(S, -1, -2)))


def lapack(oracle, A, B):
    """LAPACK zhegv (what the reference's harness calls), or the oracle where no LAPACK loads."""
    if oracle.lapack_load() is not None:
        return oracle.lapack_zhegv(A, B)
    return oracle.zhegv(A, B)


@pytest.mark.parametrize("path", WU_FILES, ids=lambda p: os.path.basename(p))
def test_every_pair_of_reference_fixtures(geneig, oracle, device, G, path):
    d = np.load(path)
    name = os.path.basename(path)[len("ref_wu_"):-len(".npz")]
    n = int(name.split("_")[0][1:])
    H, S = pairs(G, n, 8, 42, "random" if name.endswith("random") else "near-identity")
    solver = None if n == 54 else geneig.GeneigSolver(n, cache_mode=geneig.CACHE_NODISK)
    W, X, info = solve_batched(geneig, n, H, S, device, solver=solver)
    assert (info == 0).all()
    check_pairs(oracle, H, S, W, X, d["W"], d["U"], n, "all 8 pairs of %s" % name)


def test_reference_large_batch_n54(geneig, oracle, device, G):
    """tests/test_full_batched_n54.cpp:222-236,367: 1,024 pairs from seed
    0xC0FFEE0042 through the batched launch, every pair against zhegv; the
    precision form on the same batch too."""
    import torch

    H, S = pairs(G, 54, 1024, 0xC0FFEE0042, "random")
    Wr, Xr, ir = lapack(oracle, H, S)
    assert (ir == 0).all()
    W, X, info = solve_batched(geneig, 54, H, S, device)
    assert (info == 0).all()
    check_pairs(oracle, H, S, W, X, Wr, Xr, 54, "reference 1024-pair batch, plain")
    dH, dS = torch.from_numpy(H).to(device), torch.from_numpy(S).to(device)
    Wm, Um, im = geneig.empty_outputs(54, 1024, device)
    fb = geneig.mixed_n54_batched(dH, dS, Wm, Um, im)
    torch.cuda.synchronize()
    assert int(im.abs().max()) == 0 and fb == 0
    check_pairs(oracle, H, S, Wm.cpu().numpy(), Um.cpu().numpy(), Wr, Xr, 54, "reference 1024-pair batch, precision")


def test_reference_missing_fixture_n54_b32_s123(geneig, oracle, device, G):
    """tests/test_full_batched_n54.cpp:344-360 runs tests/data/n54_b32_s123_random.bin
    (not shipped) in batches of 1, 8 and 32: its inputs regenerated from seed
    123, its outputs from zhegv here."""
    H, S = pairs(G, 54, 32, 123, "random")
    Wr, Xr, ir = lapack(oracle, H, S)
    for b in (1, 8, 32):
        W, X, info = solve_batched(geneig, 54, H[:b], S[:b], device)
        assert (info == 0).all()
        check_pairs(oracle, H[:b], S[:b], W, X, Wr[:b], Xr[:b], 54, "n54_b32_s123 B=%d" % b)


@pytest.mark.parametrize("bpb", [1, 2])
def test_reference_blockdim_n32(geneig, oracle, device, G, bpb):
    """tests/test_blockdim_n32.cpp:190-197: 1,024 pairs, n = 32, seed 0xBEEFC0DE32,
    two pairs per block in the reference's sweep; here one and two pairs per
    workgroup, every pair against zhegv (the reference checks 8)."""
    H, S = pairs(G, 32, 1024, 0xBEEFC0DE32, "random")
    Wr, Xr, ir = lapack(oracle, H, S)
    s = geneig.GeneigSolver(32, batches_per_block_request=bpb, cache_mode=geneig.CACHE_NODISK)
    assert s.batches_per_block() == bpb
    W, X, info = solve_batched(geneig, 32, H, S, device, solver=s)
    assert (info == 0).all()
    check_pairs(oracle, H, S, W, X, Wr, Xr, 32, "reference n=32 batch, %d pair(s) per workgroup" % bpb)
