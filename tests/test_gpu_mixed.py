"""GPU parity tests of the precision form (geneig_mixed_n54_batched_launch):
complex64 bulk, FP64 Cholesky, FP64 refinement on the matrix cores, FP64
re-solve of the pairs the refinement hands back.  Its outputs must meet the
bar of the FP64 path against the oracle (restated zhegv): eigenvalue error
<= 1e-12 relative to the largest |lambda|, scale-invariant residual and
B-orthogonality <= 50 n eps, eigenvectors equal up to phase.
"""
import numpy as np
import pytest

from test_gpu_parity import check_pairs, to_dev, to_host

pytestmark = pytest.mark.gpu


def mixed(geneig, A, B, device, **kw):
    import torch

    batch = A.shape[0]
    dA = A if isinstance(A, torch.Tensor) else to_dev(A, device)
    dB = B if isinstance(B, torch.Tensor) else to_dev(B, device)
    W, U, info = geneig.empty_outputs(54, batch, device)
    iters = torch.full((batch,), -1, dtype=torch.int32, device=device)
    fb = geneig.mixed_n54_batched(dA, dB, W, U, info, iters=iters, **kw)
    torch.cuda.synchronize()
    return to_host(W), to_host(U), to_host(info), to_host(iters), fb


def hist(iters):
    v, c = np.unique(iters, return_counts=True)
    return {int(a): int(b) for a, b in zip(v, c)}


def test_mixed_spinwave_vs_oracle(geneig, oracle, device):
    A, B = geneig.gen_spinwave(27, 1024, seed=1, first_q=5000)
    W, X, info, iters, fb = mixed(geneig, A, B, device)
    Ah, Bh = to_host(A), to_host(B)
    print("spin-wave n=54: fallback %d of 1024, refinement steps %s" % (fb, hist(iters)), flush=True)
    assert (info == 0).all()
    assert fb <= 1024 // 20
    Wo, Xo, io = oracle.zhegv(Ah, Bh)
    check_pairs(oracle, Ah, Bh, W, X, Wo, Xo, 54, "mixed spin-wave 1024")


def test_mixed_generic_vs_oracle(geneig, oracle, device):
    A, B = oracle.gen_generic(54, 256, seed=77)
    W, X, info, iters, fb = mixed(geneig, A, B, device)
    print("generic n=54: fallback %d of 256, refinement steps %s" % (fb, hist(iters)), flush=True)
    assert (info == 0).all() and fb == 0
    Wo, Xo, io = oracle.zhegv(A, B)
    check_pairs(oracle, A, B, W, X, Wo, Xo, 54, "mixed generic 256")


def test_mixed_fallback_matches_fp64_bitwise(geneig, oracle, device):
    """With a tolerance no refinement can meet every pair is re-solved by the
    FP64 pipeline over the pair list: outputs bit-identical to the FP64 path."""
    import torch

    A, B = oracle.gen_generic(54, 37, seed=5)
    W, X, info, iters, fb = mixed(geneig, A, B, device, max_iter=1, tol=1e-300)
    assert fb == 37 and (iters == 1).all()
    dA, dB = to_dev(A, device), to_dev(B, device)
    W2, U2, i2 = geneig.empty_outputs(54, 37, device)
    geneig.full_n54_batched(dA, dB, W2, U2, i2)
    torch.cuda.synchronize()
    assert W.tobytes() == to_host(W2).tobytes()
    assert X.tobytes() == to_host(U2).tobytes()
    assert (info == to_host(i2)).all()


def test_mixed_not_positive_definite(geneig, oracle, device):
    """B with a negative pivot: info = n + j and NaN outputs, as the FP64 path."""
    A, B = oracle.gen_generic(54, 3, seed=11)
    B = B.copy()
    B[1, 5, 5] = -1.0
    W, X, info, iters, fb = mixed(geneig, A, B, device)
    Wo, Xo, io = oracle.zhegv(A, B)
    assert info[0] == 0 and info[2] == 0 and info[1] == io[1] > 54
    assert np.isnan(W[1]).all() and np.isnan(X[1]).all()
    assert fb == 0
    keep = [0, 2]
    check_pairs(oracle, A[keep], B[keep], W[keep], X[keep], Wo[keep], Xo[keep], 54, "mixed non-PD neighbours")


def compare_to_fp64_path(oracle, A, B, n, cond, W, X, Wp, Xp, tag):
    """The precision form's outputs (W, X) and the FP64 path's (Wp, Xp) on the
    same pairs against the oracle: every metric meets the FP64 bar or stays
    within twice the FP64 path's figure; at cond(B) >= 1e6 the eigenvalues are
    checked against long-double Rayleigh quotients of the oracle's
    eigenvectors instead (the oracle's own FP64 eigenvalues are off by
    ~ eps cond(B) there), with the first-order bound eps cond(B)."""
    from test_gpu_parity import EPS, TOL_EIG, TOL_PHASE

    Wo, Xo, io = oracle.zhegv(A, B)
    assert (io == 0).all()
    mp = oracle.compare(A, B, Wp, Xp, Wo, Xo)
    mm = oracle.compare(A, B, W, X, Wo, Xo)
    bars = {"eig_rel_scaled": TOL_EIG, "phase_err": TOL_PHASE, "resid": 50 * n * EPS, "orth": 50 * n * EPS}
    for k, bar in bars.items():
        print("  %s %-15s mixed %.2e  fp64 path %.2e  bar %.1e" % (tag, k, mm[k], mp[k], bar), flush=True)
        if k == "eig_rel_scaled" and cond >= 1e6:
            continue
        assert mm[k] <= max(bar, 2.0 * mp[k]), (tag, k)
    if cond >= 1e6:
        R = oracle.rayleigh_ld(A, B, Xo)
        em, ep, eo = (oracle.eig_err_scaled(w, R) for w in (W, Wp, Wo))
        print("  %s eig vs long-double Rayleigh quotients: mixed %.2e  fp64 path %.2e  oracle %.2e  bar %.1e"
              % (tag, em, ep, eo, EPS * cond), flush=True)
        assert em <= EPS * cond and ep <= EPS * cond, tag


@pytest.mark.parametrize("cond", [1e2, 1e4, 1e6, 1e8])
def test_mixed_condition_sweep(geneig, oracle, device, cond):
    """cond(B) from 1e2 to 1e8: the refinement converges or the pair is re-solved
    in FP64.  Every metric against the oracle meets the FP64 bar or, where
    cond(B) puts the FP64 path itself above that bar, stays within twice the
    FP64 path's figure on the same pairs.  The eigenvalues at cond(B) >= 1e6
    are the exception: there the oracle's own FP64 eigenvalues are off by
    ~ eps cond(B) (2-3e-12 at 1e6), as large as the errors measured, so both
    forms are checked against long-double Rayleigh quotients of the oracle's
    eigenvectors (second order in their error) with the first-order
    perturbation bound eps cond(B) of a generalized eigenvalue under a
    relative perturbation of B."""
    import torch

    A, B = oracle.gen_generic_cond(54, 64, seed=3, cond=cond)
    W, X, info, iters, fb = mixed(geneig, A, B, device)
    assert (info == 0).all()
    W2, U2, i2 = geneig.empty_outputs(54, 64, device)
    geneig.full_n54_batched(to_dev(A, device), to_dev(B, device), W2, U2, i2)
    torch.cuda.synchronize()
    print("cond %.0e: fallback %d of 64, refinement steps %s" % (cond, fb, hist(iters)), flush=True)
    compare_to_fp64_path(oracle, A, B, 54, cond, W, X, to_host(W2), to_host(U2), "generic n=54 cond %.0e" % cond)
    if cond <= 1e4:
        assert fb == 0


def spinwave_with_cond(geneig, batch, seed, cond, device):
    """Spin-wave pairs (eta, H) whose H is shifted so that cond(H) = cond:
    H + (lambda_max / cond - lambda_min) I, the extremes from numpy's eigvalsh."""
    import torch

    A, B = geneig.gen_spinwave(27, batch, seed=seed)
    Ah, Bh = to_host(A), to_host(B)
    lam = np.linalg.eigvalsh(np.swapaxes(Bh, -1, -2))
    shift = lam[:, -1] / cond - lam[:, 0]
    Bh = Bh + shift[:, None, None] * np.eye(54)[None]
    lam2 = np.linalg.eigvalsh(np.swapaxes(Bh, -1, -2))
    got = lam2[:, -1] / lam2[:, 0]
    assert np.all(np.abs(got / cond - 1.0) < 1e-2), got
    return Ah, Bh


@pytest.mark.parametrize("cond", [1e6, 1e8])
def test_mixed_spinwave_condition_sweep(geneig, oracle, device, cond):
    """The n = 54 spin-wave family driven past its natural cond(B) (<~ 1e4) to
    1e6 and 1e8 by a smaller diagonal shift (lambda_min(H) = lambda_max / cond):
    the precision form against the oracle and the FP64 path, same rules as the
    generic sweep."""
    import torch

    A, B = spinwave_with_cond(geneig, 48, 21, cond, device)
    W, X, info, iters, fb = mixed(geneig, A, B, device)
    assert (info == 0).all()
    W2, U2, i2 = geneig.empty_outputs(54, 48, device)
    geneig.full_n54_batched(to_dev(A, device), to_dev(B, device), W2, U2, i2)
    torch.cuda.synchronize()
    assert int(i2.abs().max()) == 0
    print("spin-wave cond %.0e: fallback %d of 48, refinement steps %s" % (cond, fb, hist(iters)), flush=True)
    compare_to_fp64_path(oracle, A, B, 54, cond, W, X, to_host(W2), to_host(U2), "spin-wave cond %.0e" % cond)


@pytest.mark.parametrize("batch", [0, 1, 7])
def test_mixed_batch_edges(geneig, oracle, device, batch):
    import torch

    A, B = oracle.gen_generic(54, max(batch, 1), seed=3)
    W, U, info = geneig.empty_outputs(54, max(batch, 1), device)
    W.fill_(-7.0)
    fb = geneig.mixed_n54_batched(to_dev(A, device), to_dev(B, device), W, U, info, batch=batch)
    torch.cuda.synchronize()
    if batch == 0:
        assert fb == 0 and (to_host(W) == -7.0).all()
        return
    Wo, Xo, io = oracle.zhegv(A, B)
    check_pairs(oracle, A, B, to_host(W), to_host(U), Wo, Xo, 54, "mixed batch=%d" % batch)


def test_runtime_precision_n54_matches_static(geneig, oracle, device):
    """The hipRTC precision module for n = 54 (compiled for the device's full
    target, e.g. gfx950:sramecc+:xnack-) and the hipcc build of the same
    source agree on the eigenvalues to rounding, and both meet the FP64 bar
    (eigenvectors may differ inside near-degenerate clusters)."""
    import torch

    A, B = geneig.gen_spinwave(27, 96, seed=12)
    W1, U1, i1 = geneig.empty_outputs(54, 96, device)
    geneig.mixed_n54_batched(A, B, W1, U1, i1)
    s = geneig.GeneigMixedSolver(54, cache_mode=geneig.CACHE_NODISK)
    W2, U2, i2 = geneig.empty_outputs(54, 96, device)
    fb = s.launch_mixed(A, B, W2, U2, i2)
    torch.cuda.synchronize()
    assert fb == 0 and s.variant() == 1
    dw = (W1 - W2).abs().max().item() / W1.abs().max().item()
    print("static vs hipRTC precision form: max rel diff W %.2e" % dw, flush=True)
    assert dw <= 1e-12
    Ah, Bh = to_host(A), to_host(B)
    Wo, Xo, _ = oracle.zhegv(Ah, Bh)
    for tag, W, U in (("static", W1, U1), ("hipRTC", W2, U2)):
        check_pairs(oracle, Ah, Bh, to_host(W), to_host(U), Wo, Xo, 54, "precision n=54 %s" % tag)


@pytest.mark.parametrize("n,cond", [(96, 0), (96, 1e2), (96, 1e4), (96, 1e6), (96, 1e8), (16, 0), (33, 0),
                                    (64, 1e2)])
def test_runtime_precision_vs_oracle(geneig, oracle, device, n, cond):
    """Precision form through the runtime solver at other n, generic pairs and
    controlled cond(B); same rules as the n = 54 sweep."""
    import torch

    if cond:
        A, B = oracle.gen_generic_cond(n, 48, seed=5, cond=cond)
    else:
        A, B = oracle.gen_generic(n, 48, seed=5)
    s = geneig.GeneigMixedSolver(n)
    W, U, info = geneig.empty_outputs(n, 48, device)
    iters = torch.zeros(48, dtype=torch.int32, device=device)
    fb = s.launch_mixed(to_dev(A, device), to_dev(B, device), W, U, info, iters=iters)
    torch.cuda.synchronize()
    assert (to_host(info) == 0).all()
    W2, U2, i2 = geneig.empty_outputs(n, 48, device)
    geneig.GeneigSolver(n).launch(to_dev(A, device), to_dev(B, device), W2, U2, i2)
    torch.cuda.synchronize()
    print("n=%d cond %.0e: fallback %d of 48, steps %s" % (n, cond, fb, hist(to_host(iters))), flush=True)
    compare_to_fp64_path(oracle, A, B, n, cond, to_host(W), to_host(U), to_host(W2), to_host(U2),
                         "n=%d cond %.0e" % (n, cond))
    if cond <= 1e4:
        assert fb == 0


def test_mixed_concurrent_parts_match_single_launches(geneig, oracle, device):
    """The precision form of a one-chunk batch runs as three concurrent parts
    (Cholesky, complex64 stages and refinement per part on internal streams,
    fallback list shared); 30,000 pairs must equal three separate launches of
    10,000 (each too small to split) bit for bit, step counts included."""
    import torch

    total, part = 30000, 10000
    A, B = geneig.gen_spinwave(27, total, seed=10)
    W, U, info = geneig.empty_outputs(54, total, device)
    iters = torch.full((total,), -1, dtype=torch.int32, device=device)
    fb = geneig.mixed_n54_batched(A, B, W, U, info, iters=iters)
    W2, U2, i2 = geneig.empty_outputs(54, total, device)
    it2 = torch.full((total,), -1, dtype=torch.int32, device=device)
    fb2 = 0
    for off in range(0, total, part):
        sl = slice(off, off + part)
        fb2 += geneig.mixed_n54_batched(A[sl], B[sl], W2[sl], U2[sl], i2[sl], iters=it2[sl])
    torch.cuda.synchronize()
    assert fb == fb2
    assert torch.equal(info, i2) and torch.equal(iters, it2) and int(iters.min()) >= 1
    assert torch.equal(W, W2) and torch.equal(U, U2)


def test_mixed_fallback_parts_all_pairs(geneig, oracle, device):
    """24,576 generic pairs run as three concurrent parts; with a tolerance no
    refinement can meet, every pair is listed by three refinement dispatches
    at once (atomics on one counter) and re-solved by the FP64 pipeline over
    the list after the join, in the solver area the parts used.  The list's
    length is read on the device, and the fixed grids walk it in several
    trips.  Outputs equal the FP64 path bit for bit."""
    import torch

    total = 24576
    assert geneig.n54_workspace_bytes(total, 1) > 0
    A, B = geneig.gen_generic(54, total, seed=21)
    W, U, info = geneig.empty_outputs(54, total, device)
    iters = torch.full((total,), -1, dtype=torch.int32, device=device)
    fb = geneig.mixed_n54_batched(A, B, W, U, info, iters=iters, max_iter=1, tol=1e-300)
    W2, U2, i2 = geneig.empty_outputs(54, total, device)
    geneig.full_n54_batched(A, B, W2, U2, i2)
    torch.cuda.synchronize()
    print("all-fallback parts: fb %d of %d" % (fb, total), flush=True)
    assert fb == total and int((iters == 1).sum()) == total
    assert torch.equal(info, i2) and int(info.abs().max()) == 0
    assert torch.equal(W, W2) and torch.equal(U, U2)


def test_mixed_graph_capture(geneig, oracle, device):
    """The precision form reads its fallback count on the device, so with
    report_fallback=False the launch never waits for the host and captures
    into a graph (after n54_reserve on the capturing stream).  Two captures:
    16,384 spin-wave pairs (two parts, no fallback) and 512 generic pairs that
    all fall back (max_iter=1, tol=1e-300); each is replayed three times with
    eager work and synchronizations between, and every replay equals the
    eager launch bit for bit."""
    import torch

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    cases = [("spin-wave", geneig.gen_spinwave(27, 16384, seed=13), {}),
             ("all-fallback", geneig.gen_generic(54, 512, seed=14), dict(max_iter=1, tol=1e-300))]
    for name, (A, B), kw in cases:
        total = A.shape[0]
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            geneig.n54_reserve(total, 1, stream=s)
            W0, U0, i0 = geneig.empty_outputs(54, total, device)
            fb0 = geneig.mixed_n54_batched(A, B, W0, U0, i0, stream=s, **kw)
        s.synchronize()
        W, U, info = geneig.empty_outputs(54, total, device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            geneig.mixed_n54_batched(A, B, W, U, info, stream=s, report_fallback=False, **kw)
        torch.cuda.synchronize()
        for rep in range(3):
            W.fill_(float("nan"))
            U.fill_(float("nan"))
            info.fill_(-1)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(W, W0) and torch.equal(U, U0) and torch.equal(info, i0), (name, rep)
        print("%s: fallback %d of %d, 3 replays equal" % (name, fb0, total), flush=True)
        assert int(i0.abs().max()) == 0
        assert fb0 == (total if kw else 0)
        del g


@pytest.mark.parametrize("total", [1024, 30000])
def test_mixed_eta_equals_general_precision_form(geneig, oracle, device, total):
    """The (eta, S) launcher (A = eta not read; K1f and the fallback K1 form it
    in registers, the refinement scales X by eta instead of a matrix-core pass)
    against the general launcher on the materialized spin-wave pairs: W, U,
    info and the step counts bit for bit, at 1,024 pairs (one launch) and
    30,000 (three concurrent parts); the outputs meet the FP64 bars against
    the oracle."""
    import torch

    A, B = geneig.gen_spinwave(27, total, seed=31)
    eta = torch.zeros(54, dtype=torch.complex128, device=device)
    eta[:27], eta[27:] = 1.0, -1.0
    assert torch.equal(A[0], torch.diag(eta))    # the generator's A is eta
    W, U, info = geneig.empty_outputs(54, total, device)
    it = torch.full((total,), -1, dtype=torch.int32, device=device)
    fb = geneig.mixed_n54_batched(A, B, W, U, info, iters=it)
    W2, U2, i2 = geneig.empty_outputs(54, total, device)
    it2 = torch.full((total,), -1, dtype=torch.int32, device=device)
    fb2 = geneig.mixed_eta_n54_batched(B, W2, U2, i2, iters=it2)
    torch.cuda.synchronize()
    print("eta form, %d pairs: fallback %d / %d, steps %s" % (total, fb2, fb, hist(to_host(it2))), flush=True)
    assert fb == fb2 and torch.equal(it, it2) and torch.equal(info, i2)
    assert torch.equal(W, W2) and torch.equal(U, U2)
    if total == 1024:
        Ah, Bh = to_host(A), to_host(B)
        Wo, Xo, io = oracle.zhegv(Ah, Bh)
        check_pairs(oracle, Ah, Bh, to_host(W2), to_host(U2), Wo, Xo, 54, "eta form 1024")


def test_mixed_eta_fallback_matches_fp64_bitwise(geneig, oracle, device):
    """Every pair handed to the FP64 fallback (tolerance no refinement meets):
    its K1 factors S and forms eta in registers; the outputs equal the FP64
    drop-in's on (eta, S) bit for bit."""
    import torch

    A, B = geneig.gen_spinwave(27, 40, seed=3)
    W, U, info = geneig.empty_outputs(54, 40, device)
    fb = geneig.mixed_eta_n54_batched(B, W, U, info, max_iter=1, tol=1e-300)
    W2, U2, i2 = geneig.empty_outputs(54, 40, device)
    geneig.full_n54_batched(A, B, W2, U2, i2)
    torch.cuda.synchronize()
    assert fb == 40
    assert torch.equal(info, i2) and torch.equal(W, W2) and torch.equal(U, U2)


def test_mixed_eta_not_positive_definite(geneig, oracle, device):
    """An S whose leading minor of order 11 is not positive definite: the
    (eta, S) launcher reports info = n + 11 (zhegv numbering) and NaN for that
    pair, as the general launcher does on (eta, S); the other pairs' outputs
    equal the general launcher's bit for bit."""
    import torch

    A, B = geneig.gen_spinwave(27, 64, seed=17)
    B[5, 10, 10] = -1.0
    W, U, info = geneig.empty_outputs(54, 64, device)
    geneig.mixed_n54_batched(A, B, W, U, info)
    W2, U2, i2 = geneig.empty_outputs(54, 64, device)
    geneig.mixed_eta_n54_batched(B, W2, U2, i2)
    torch.cuda.synchronize()
    i2h = to_host(i2)
    assert i2h[5] == 54 + 11 and (np.delete(i2h, 5) == 0).all(), i2h
    assert torch.equal(info, i2)
    assert torch.isnan(W2[5]).all() and torch.isnan(U2[5].real).all()
    keep = [p for p in range(64) if p != 5]
    assert torch.equal(W[keep], W2[keep]) and torch.equal(U[keep], U2[keep])


def test_runtime_precision_launch_without_fallback_count(geneig, oracle, device):
    """The runtime precision launch with no host count (h_fallback NULL: the
    launch never waits) gives the same outputs, bit for bit, as one that
    reports the count; with every pair handed back (tol no step meets) the
    count read afterwards on the same inputs is the batch."""
    import torch

    A, B = oracle.gen_generic(40, 24, seed=21)
    dA, dB = to_dev(A, device), to_dev(B, device)
    s = geneig.GeneigMixedSolver(40, cache_mode=geneig.CACHE_NODISK)
    W, U, info = geneig.empty_outputs(40, 24, device)
    assert s.launch_mixed(dA, dB, W, U, info, report_fallback=False) is None
    W2, U2, i2 = geneig.empty_outputs(40, 24, device)
    fb = s.launch_mixed(dA, dB, W2, U2, i2)
    torch.cuda.synchronize()
    assert fb == 0
    assert torch.equal(W, W2) and torch.equal(U, U2) and torch.equal(info, i2)
    W3, U3, i3 = geneig.empty_outputs(40, 24, device)
    assert s.launch_mixed(dA, dB, W3, U3, i3, max_iter=1, tol=1e-300, report_fallback=False) is None
    assert s.launch_mixed(dA, dB, W3, U3, i3, max_iter=1, tol=1e-300) == 24


def paired_exact(W, U):
    """Per pair: whether W[53 - i] = -W[i] and X(:, i) = Sigma conj X(:, 53 - i)
    hold bit for bit (U column-major: U[p, col, row]; Sigma swaps rows
    0..26 and 27..53)."""
    import torch

    w_ok = (W + W.flip(1) == 0).all(dim=1)
    part = torch.cat([U[:, :, 27:], U[:, :, :27]], dim=2).conj().flip(1)
    return w_ok & (U == part).flatten(1).all(dim=1)


def spin_resid_orth(A, B, W, U):
    """Largest scale-invariant residual and B-orthogonality defect (units of
    n eps) over the pairs, on the device, in 4096-pair chunks; storage is
    column-major, so with Xt = X^T: (A X)^T = Xt At, (X^H B X)^T = (Xt Bt)
    conj(Xt)^T."""
    import torch

    eps = float(np.finfo(np.float64).eps)
    res = orth = 0.0
    eye = torch.eye(54, dtype=U.dtype, device=U.device)[None]
    for s in range(0, A.shape[0], 4096):
        At, Bt, Xt, w = A[s:s + 4096], B[s:s + 4096], U[s:s + 4096], W[s:s + 4096]
        XB = torch.matmul(Xt, Bt)
        R = torch.matmul(Xt, At) - w[:, :, None] * XB
        xn = torch.linalg.vector_norm(Xt, dim=2)
        fa, fb = torch.linalg.matrix_norm(At), torch.linalg.matrix_norm(Bt)
        res = max(res, float((torch.linalg.vector_norm(R, dim=2) /
                              ((fa[:, None] + w.abs() * fb[:, None]) * xn)).max()))
        G = torch.matmul(XB, Xt.conj().transpose(-1, -2).contiguous()) - eye
        orth = max(orth, float((G.abs() / (fb[:, None, None] * xn[:, :, None] * xn[:, None, :])).max()))
    return res / (54 * eps), orth / (54 * eps)


def test_mixed_eta_paired_vs_oracle(geneig, oracle, device):
    """The paired launcher on spin-wave pairs (S = [[P, Q], [conj Q, conj P]]
    as the generator builds H(q)): FP64 bars against the oracle, and every
    pair the refinement finished carries the pairing bit for bit (the FP64
    fallback's pairs carry it to rounding and phase)."""
    import torch

    A, B = geneig.gen_spinwave(27, 1024, seed=1, first_q=5000)
    W, U, info = geneig.empty_outputs(54, 1024, device)
    info.fill_(-1)
    it = torch.full((1024,), -1, dtype=torch.int32, device=device)
    fb = geneig.mixed_eta_paired_n54_batched(B, W, U, info, iters=it)
    torch.cuda.synchronize()
    ok = paired_exact(W, U)
    print("paired form 1024: fallback %d, steps %s, exactly paired %d" % (fb, hist(to_host(it)), int(ok.sum())),
          flush=True)
    assert int(info.abs().max()) == 0
    assert fb <= 1024 // 20 and int(ok.sum()) >= 1024 - fb
    Ah, Bh = to_host(A), to_host(B)
    Wo, Xo, io = oracle.zhegv(Ah, Bh)
    check_pairs(oracle, Ah, Bh, to_host(W), to_host(U), Wo, Xo, 54, "paired form 1024")


@pytest.mark.parametrize("total", [30000])
def test_mixed_eta_paired_close_to_eta_form(geneig, oracle, device, total):
    """30,000 spin-wave pairs (three concurrent parts): the paired launcher's
    eigenvalues agree with the eta form's to 1e-12 relative to the largest
    |lambda|, info agrees, residual and B-orthogonality of every pair stay
    <= 50 n eps, and all but the fallback pairs are paired bit for bit."""
    import torch

    A, B = geneig.gen_spinwave(27, total, seed=31)
    W, U, info = geneig.empty_outputs(54, total, device)
    fb = geneig.mixed_eta_n54_batched(B, W, U, info)
    W2, U2, i2 = geneig.empty_outputs(54, total, device)
    i2.fill_(-1)
    fb2 = geneig.mixed_eta_paired_n54_batched(B, W2, U2, i2)
    torch.cuda.synchronize()
    ok = paired_exact(W2, U2)
    dw = float(((W - W2).abs().amax(dim=1) / W.abs().amax(dim=1).clamp(min=1.0)).max())
    res, orth = spin_resid_orth(A, B, W2, U2)
    print("paired form %d: fallback %d (eta form %d), exactly paired %d, |dW| %.2e, resid %.3f, orth %.3f n eps"
          % (total, fb2, fb, int(ok.sum()), dw, res, orth), flush=True)
    assert torch.equal(info, i2) and int(info.abs().max()) == 0
    assert fb2 <= total // 20 and int(ok.sum()) >= total - fb2
    assert dw <= 1e-12 and res <= 50 and orth <= 50


def test_mixed_eta_paired_fallback_matches_fp64_bitwise(geneig, oracle, device):
    """Every pair handed to the FP64 fallback (tolerance no refinement step
    meets): the outputs equal the FP64 drop-in's on (eta, S) bit for bit."""
    import torch

    A, B = geneig.gen_spinwave(27, 40, seed=3)
    W, U, info = geneig.empty_outputs(54, 40, device)
    fb = geneig.mixed_eta_paired_n54_batched(B, W, U, info, max_iter=1, tol=1e-300)
    W2, U2, i2 = geneig.empty_outputs(54, 40, device)
    geneig.full_n54_batched(A, B, W2, U2, i2)
    torch.cuda.synchronize()
    assert fb == 40
    assert torch.equal(info, i2) and torch.equal(W, W2) and torch.equal(U, U2)


def test_mixed_eta_paired_not_positive_definite(geneig, oracle, device):
    """An S whose leading minor of order 11 is not positive definite: info =
    n + 11 and NaN outputs for that pair, as the eta form reports; the other
    pairs are solved and paired."""
    import torch

    A, B = geneig.gen_spinwave(27, 64, seed=17)
    B[5, 10, 10] = -1.0
    W, U, info = geneig.empty_outputs(54, 64, device)
    geneig.mixed_eta_n54_batched(B, W, U, info)
    W2, U2, i2 = geneig.empty_outputs(54, 64, device)
    fb = geneig.mixed_eta_paired_n54_batched(B, W2, U2, i2)
    torch.cuda.synchronize()
    i2h = to_host(i2)
    assert i2h[5] == 54 + 11 and (np.delete(i2h, 5) == 0).all(), i2h
    assert torch.equal(info, i2)
    assert torch.isnan(W2[5]).all() and torch.isnan(U2[5].real).all()
    keep = [p for p in range(64) if p != 5]
    ok = paired_exact(W2[keep], U2[keep])
    assert int(ok.sum()) >= 63 - fb
    dw = float(((W[keep] - W2[keep]).abs().amax(dim=1) / W[keep].abs().amax(dim=1).clamp(min=1.0)).max())
    assert dw <= 1e-12
