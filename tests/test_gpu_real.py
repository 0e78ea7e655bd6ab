"""GPU parity tests of the real-arithmetic solve of the spin-wave pencil
(geneig_eta_paired_n54_batched_launch): (eta, S) with S = [[P, Q], [conj Q,
conj P]] mapped by T = [[I, iI], [I, -iI]] / sqrt 2 to (iJ, T^H S T), solved
in real FP64 arithmetic, the 27 positive eigenpairs taken from the QL and
their partners written exactly.  Bars of the FP64 path against the oracle
(restated zhegv on (eta, S)): eigenvalue error <= 1e-12 relative to the
largest |lambda|, scale-invariant residual and S-orthogonality <= 50 n eps,
eigenvectors equal up to phase.
"""
import numpy as np
import pytest

from test_gpu_parity import check_pairs, to_host
from test_gpu_mixed import paired_exact, spin_resid_orth

pytestmark = pytest.mark.gpu


def real_solve(geneig, B, device):
    import torch

    batch = B.shape[0]
    W, U, info = geneig.empty_outputs(54, batch, device)
    info.fill_(-1)
    geneig.eta_paired_n54_batched(B, W, U, info)
    torch.cuda.synchronize()
    return W, U, info


def test_real_form_vs_oracle(geneig, oracle, device):
    A, B = geneig.gen_spinwave(27, 1024, seed=1, first_q=5000)
    W, U, info = real_solve(geneig, B, device)
    ok = paired_exact(W, U)
    print("real form 1024: info != 0 for %d, exactly paired %d" % (int((info != 0).sum()), int(ok.sum())), flush=True)
    assert int(info.abs().max()) == 0 and bool(ok.all())
    Ah, Bh = to_host(A), to_host(B)
    Wo, Xo, io = oracle.zhegv(Ah, Bh)
    check_pairs(oracle, Ah, Bh, to_host(W), to_host(U), Wo, Xo, 54, "real form 1024")


def test_real_form_against_plain_fp64(geneig, oracle, device):
    """30,000 pairs (three concurrent parts): eigenvalues within 1e-12 of the
    plain FP64 drop-in's on (eta, S), residual and S-orthogonality of every
    pair <= 50 n eps, every pair paired bit for bit, eigenvectors equal to the
    plain path's up to phase on a sample."""
    import torch

    A, B = geneig.gen_spinwave(27, 30000, seed=31)
    W, U, info = real_solve(geneig, B, device)
    W2, U2, i2 = geneig.empty_outputs(54, 30000, device)
    geneig.full_n54_batched(A, B, W2, U2, i2)
    torch.cuda.synchronize()
    dw = float(((W - W2).abs().amax(dim=1) / W2.abs().amax(dim=1).clamp(min=1.0)).max())
    res, orth = spin_resid_orth(A, B, W, U)
    ok = paired_exact(W, U)
    X, X2 = U[::97].transpose(1, 2), U2[::97].transpose(1, 2)
    ph = (X2.conj() * X).sum(1)
    dx = float((X - X2 * (ph / ph.abs())[:, None, :]).abs().max())
    print("real form 30000: |dW| %.2e, resid %.3f, orth %.3f n eps, exactly paired %d, |dX| %.2e"
          % (dw, res, orth, int(ok.sum()), dx), flush=True)
    assert torch.equal(info, i2) and int(info.abs().max()) == 0
    assert dw <= 1e-12 and res <= 50 and orth <= 50 and bool(ok.all()) and dx <= 1e-10


def test_real_form_agrees_with_paired_precision_form(geneig, oracle, device):
    """The two solvers of the block-form pencil (real FP64 and the precision
    form's paired refinement) agree to the FP64 bar on the same pairs."""
    import torch

    A, B = geneig.gen_spinwave(27, 2048, seed=9)
    W, U, info = real_solve(geneig, B, device)
    W2, U2, i2 = geneig.empty_outputs(54, 2048, device)
    geneig.mixed_eta_paired_n54_batched(B, W2, U2, i2)
    torch.cuda.synchronize()
    dw = float(((W - W2).abs().amax(dim=1) / W2.abs().amax(dim=1).clamp(min=1.0)).max())
    assert torch.equal(info, i2) and dw <= 1e-12, dw


def test_real_form_not_positive_definite(geneig, oracle, device):
    """A pair whose S is not positive definite: info > 54 (the failing order
    of T^H S T's leading minors) and NaN outputs for that pair; the other
    pairs are solved as by the plain path."""
    import torch

    A, B = geneig.gen_spinwave(27, 64, seed=17)
    B[5, 10, 10] = -1.0
    W, U, info = real_solve(geneig, B, device)
    ih = to_host(info)
    assert 54 < ih[5] <= 108 and (np.delete(ih, 5) == 0).all(), ih
    assert torch.isnan(W[5]).all() and torch.isnan(U[5].real).all()
    W2, U2, i2 = geneig.empty_outputs(54, 64, device)
    geneig.full_n54_batched(A, B, W2, U2, i2)
    torch.cuda.synchronize()
    keep = [p for p in range(64) if p != 5]
    dw = float(((W[keep] - W2[keep]).abs().amax(dim=1) / W2[keep].abs().amax(dim=1).clamp(min=1.0)).max())
    assert dw <= 1e-12 and bool(paired_exact(W[keep], U[keep]).all())


@pytest.mark.parametrize("batch", [1, 3, 65])
def test_real_form_small_batches(geneig, oracle, device, batch):
    A, B = geneig.gen_spinwave(27, batch, seed=40 + batch)
    W, U, info = real_solve(geneig, B, device)
    assert int(info.abs().max()) == 0 and bool(paired_exact(W, U).all())
    Ah, Bh = to_host(A), to_host(B)
    Wo, Xo, io = oracle.zhegv(Ah, Bh)
    check_pairs(oracle, Ah, Bh, to_host(W), to_host(U), Wo, Xo, 54, "real form %d" % batch)


def test_real_form_ql_log_overflow_path(geneig, oracle, device, monkeypatch):
    """With the QL rotation log capped below one QL iteration
    (GENEIG_QL_LOG_ENTRIES, read at every launch) every pair's QL runs in the
    one-pair overflow kernel: the real form still meets the FP64 bar with the
    pairing exact, and its eigenvalues agree with the logged path's to
    rounding."""
    A, B = geneig.gen_spinwave(27, 96, seed=5)
    Ah, Bh = to_host(A), to_host(B)
    Wo, Xo, _ = oracle.zhegv(Ah, Bh)
    res = {}
    for cap in (None, "16"):
        if cap is not None:
            monkeypatch.setenv("GENEIG_QL_LOG_ENTRIES", cap)
        W, U, info = real_solve(geneig, B, device)
        tag = "real form, log cap %s" % cap
        assert int(info.abs().max()) == 0 and bool(paired_exact(W, U).all()), tag
        check_pairs(oracle, Ah, Bh, to_host(W), to_host(U), Wo, Xo, 54, tag)
        res[cap] = to_host(W)
    assert np.abs(res[None] - res["16"]).max() <= 1e-13 * np.abs(res[None]).max()


def test_real_form_workspace_chunking(geneig, oracle, device):
    """A batch larger than one workspace chunk (GENEIG_WORKSPACE_MB=16: 141
    pairs per chunk) goes through several K1r-K2-K3r rounds on the caller's
    stream: the FP64 bar against the oracle and exact pairing on all 700
    pairs."""
    import os
    import subprocess
    import sys

    from conftest import GOLDEN

    code = r'''
import os, sys, numpy as np, torch
sys.path.insert(0, %r); sys.path.insert(0, %r)
import conftest, oracle as O
g = conftest.load_package()
A, B = g.gen_spinwave(27, 700, seed=12)
W, U, info = g.empty_outputs(54, 700, "cuda")
g.eta_paired_n54_batched(B, W, U, info); torch.cuda.synchronize()
Wh, Uh = W.cpu().numpy(), U.cpu().numpy()
part = np.concatenate([Uh[:, :, 27:], Uh[:, :, :27]], axis=2).conj()[:, ::-1, :]
assert (Wh + Wh[:, ::-1] == 0).all() and (Uh == part).all()
Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
Wo, Xo, io = O.zhegv(Ah, Bh)
m = O.compare(Ah, Bh, Wh, Uh, Wo, Xo)
assert int(info.abs().max()) == 0 and m["eig_rel_scaled"] < 1e-12 and m["phase_err"] < 1e-10, m
print("ok", m["eig_rel_scaled"])
''' % (os.path.dirname(GOLDEN), os.path.join(os.path.dirname(os.path.dirname(GOLDEN)), "oracle"))
    env = dict(os.environ, GENEIG_WORKSPACE_MB="16")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr[-2000:], flush=True)
    assert out.returncode == 0 and "ok" in out.stdout


@pytest.mark.parametrize("cond", [1e3, 1e4, 1e6, 1e8])
def test_real_form_spinwave_condition_sweep(geneig, oracle, device, cond):
    """The spin-wave family driven to cond(S) = 1e3 ... 1e8 by a diagonal
    shift (which keeps the block form): the real form against the oracle and
    the plain FP64 path on the same pairs, every metric within the FP64 bar or
    twice the plain path's figure, eigenvalues at cond >= 1e6 against
    long-double Rayleigh quotients; pairing exact."""
    import torch

    from test_gpu_mixed import compare_to_fp64_path, spinwave_with_cond
    from test_gpu_parity import to_dev

    A, B = spinwave_with_cond(geneig, 48, 21, cond, device)
    dB = to_dev(B, device)
    W, U, info = real_solve(geneig, dB, device)
    assert int(info.abs().max()) == 0 and bool(paired_exact(W, U).all())
    W2, U2, i2 = geneig.empty_outputs(54, 48, device)
    geneig.full_n54_batched(to_dev(A, device), dB, W2, U2, i2)
    torch.cuda.synchronize()
    assert int(i2.abs().max()) == 0
    compare_to_fp64_path(oracle, A, B, 54, cond, to_host(W), to_host(U), to_host(W2), to_host(U2),
                         "real form cond %.0e" % cond)
