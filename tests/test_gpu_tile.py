"""GPU tests of the four-wave K1 / K3 (one pair per workgroup of 256 threads,
the matrix spread over the workgroup's registers; threads per pair 256, the
reference's block_dim_x knob, src/nvrtc/nvrtc_solver.hpp:81-217).

Each case runs through the C ABI (runtime solver, hipRTC module compiled with
the four-wave kernels) and is checked against the oracle's zhegv restatement
at the FP64 bars (eigenvalue <= 1e-12 scaled by max(|lambda|, 1), residual and
B-orthogonality <= 50 n eps), against the reference's fixtures, and against
the one-wave path where both exist.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from test_gpu_parity import REF_FILES, check_pairs, to_dev, to_host

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
TILE = 256


def solve(geneig, s, A, B, device, inplace=False):
    import torch

    n = A.shape[-1]
    dA, dB = to_dev(A, device), to_dev(B, device)
    W, U, info = geneig.empty_outputs(n, A.shape[0], device)
    info.fill_(-1)                 # every pair's info is written by the solve
    if inplace:
        s.launch_inplace(dA, dB, W, info)
        U = dA
    else:
        s.launch(dA, dB, W, U, info)
    torch.cuda.synchronize()
    return to_host(W), to_host(U), to_host(info)


@pytest.mark.parametrize("n", [1, 2, 3, 8, 16, 17, 33, 54, 64, 65, 96, 100, 128])
def test_tile_fp64_vs_oracle(geneig, oracle, device, n):
    batch = 6 if n >= 96 else 12
    print("four-wave FP64 solver n=%d, %d generic pairs" % (n, batch), flush=True)
    s = geneig.GeneigSolver(n, force_block_dim_x=TILE, cache_mode=geneig.CACHE_NODISK)
    assert s.block_dim() == (TILE, 1, 1) and s.batches_per_block() == 1
    A, B = oracle.gen_generic(n, batch, seed=100 + n)
    Wo, Xo, io = oracle.zhegv(A, B)
    assert (io == 0).all()
    W, X, info = solve(geneig, s, A, B, device)
    assert (info == 0).all(), info
    check_pairs(oracle, A, B, W, X, Wo, Xo, n, "tile n=%d" % n)


@pytest.mark.parametrize("path", REF_FILES, ids=os.path.basename)
def test_tile_vs_reference_fixture(geneig, oracle, device, path):
    """The reference's zhegvd fixtures (n = 16 ... 96, random and
    near-identity S) through the four-wave kernels."""
    fx = oracle.geig_read(path)
    n = fx["n"]
    print("four-wave solver on %s" % os.path.basename(path), flush=True)
    s = geneig.GeneigSolver(n, force_block_dim_x=TILE, cache_mode=geneig.CACHE_NODISK)
    W, X, info = solve(geneig, s, fx["H"], fx["S"], device)
    assert (info == 0).all()
    check_pairs(oracle, fx["H"], fx["S"], W, X, fx["W"], fx["U"], n, "tile " + os.path.basename(path))


@pytest.mark.parametrize("n", [40, 54, 96])
def test_tile_inplace_matches_out_of_place(geneig, oracle, device, n):
    """zhegvd storage (eigenvectors over H, the factor over S's lower
    triangle): the same kernels read L from S instead of U, so W and X equal
    the out-of-place solve bit for bit."""
    print("four-wave in-place n=%d" % n, flush=True)
    s = geneig.GeneigSolver(n, force_block_dim_x=TILE, cache_mode=geneig.CACHE_NODISK)
    A, B = oracle.gen_generic(n, 9, seed=7 * n)
    W1, X1, i1 = solve(geneig, s, A, B, device)
    W2, X2, i2 = solve(geneig, s, A, B, device, inplace=True)
    assert (i1 == 0).all() and (i2 == 0).all()
    assert W1.tobytes() == W2.tobytes() and X1.tobytes() == X2.tobytes()


def test_tile_not_positive_definite(geneig, oracle, device):
    """A B whose leading minor of order j is not positive definite: info =
    n + j (zhegv numbering) and NaN in W and X for that pair only."""
    n = 70
    A, B = oracle.gen_generic(n, 4, seed=5)
    B[2, 9, 9] = -1.0          # leading minor of order 10 fails
    print("four-wave n=%d with a non-PD B in pair 2" % n, flush=True)
    s = geneig.GeneigSolver(n, force_block_dim_x=TILE, cache_mode=geneig.CACHE_NODISK)
    W, X, info = solve(geneig, s, A, B, device)
    assert list(info) == [0, 0, n + 10, 0], info
    assert np.isnan(W[2]).all() and np.isnan(X[2]).all()
    assert np.isfinite(W[[0, 1, 3]]).all()


@pytest.mark.parametrize("n", [32, 54])
def test_tile_matches_one_wave_path(geneig, oracle, device, n):
    """The one-wave and four-wave kernels run the same algorithms in other
    summation orders: eigenvalues within a few ulps of each other, vectors
    equal up to phase."""
    A, B = oracle.gen_generic(n, 16, seed=3 * n)
    print("one-wave vs four-wave n=%d" % n, flush=True)
    s1 = geneig.GeneigSolver(n, force_block_dim_x=64, cache_mode=geneig.CACHE_NODISK)
    s4 = geneig.GeneigSolver(n, force_block_dim_x=TILE, cache_mode=geneig.CACHE_NODISK)
    W1, X1, _ = solve(geneig, s1, A, B, device)
    W4, X4, _ = solve(geneig, s4, A, B, device)
    m = oracle.compare(A, B, W4, X4, W1, X1)
    print("four-wave vs one-wave: %s" % m, flush=True)
    assert m["eig_rel_scaled"] <= 1e-13 and m["phase_err"] <= 1e-10


@pytest.mark.parametrize("n,kind", [(54, "spinwave"), (54, "generic"), (96, "generic")])
def test_tile_precision_form(geneig, oracle, device, n, kind):
    """The precision form with its Cholesky, complex64 K1 / K3 and FP64
    fallback K1 / K3 on four waves per pair, at the FP64 bars."""
    import torch

    if kind == "spinwave":
        A, B = geneig.gen_spinwave(n // 2, 64, seed=11)
        Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
    else:
        Ah, Bh = oracle.gen_generic(n, 16, seed=41 + n)
        A, B = to_dev(Ah, device), to_dev(Bh, device)
    print("four-wave precision form n=%d %s" % (n, kind), flush=True)
    s = geneig.GeneigMixedSolver(n, cache_mode=geneig.CACHE_NODISK, block_dim_x=TILE)
    assert s.block_dim()[0] == TILE
    W, U, info = geneig.empty_outputs(n, A.shape[0], device)
    info.fill_(-1)
    fb = s.launch_mixed(A, B, W, U, info)
    torch.cuda.synchronize()
    Wo, Xo, io = oracle.zhegv(Ah, Bh)
    assert (to_host(info) == 0).all()
    print("fallback pairs: %d" % fb, flush=True)
    check_pairs(oracle, Ah, Bh, to_host(W), to_host(U), Wo, Xo, n, "tile precision n=%d %s" % (n, kind))


@pytest.mark.parametrize("n", [54, 96])
def test_tile_precision_fallback_equals_fp64(geneig, oracle, device, n):
    """One refinement step at an unreachable tolerance hands every pair to
    the FP64 fallback, whose four-wave kernels walk the device list: the
    outputs then equal the four-wave FP64 solver's bit for bit."""
    import torch

    Ah, Bh = oracle.gen_generic(n, 20, seed=9 + n)
    print("four-wave precision form n=%d, every pair to the fallback" % n, flush=True)
    sm = geneig.GeneigMixedSolver(n, cache_mode=geneig.CACHE_NODISK, block_dim_x=TILE)
    A, B = to_dev(Ah, device), to_dev(Bh, device)
    W, U, info = geneig.empty_outputs(n, 20, device)
    fb = sm.launch_mixed(A, B, W, U, info, max_iter=1, tol=1e-300)
    torch.cuda.synchronize()
    assert fb == 20
    s = geneig.GeneigSolver(n, force_block_dim_x=TILE, cache_mode=geneig.CACHE_NODISK)
    W2, X2, i2 = solve(geneig, s, Ah, Bh, device)
    assert to_host(W).tobytes() == W2.tobytes() and to_host(U).tobytes() == X2.tobytes()
