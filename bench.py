"""Benchmark: n = 54 spin-wave generalized eigensolves per second on one MI355X,
chosen (precision) form against the plain FP64 form.

Workload (one "step"): solve every pair of a 131,072-point spin-wave batch,
n = 54 (27 sublattices), complex128, q on a 64 x 64 x 32 grid in [0,1)^3,
pair (A, B) = (eta, H(q)) with lambda_min(H) = 10^U(-3,-1) -- generated on the
device by this library (synthetic data).  Inputs are resident in HBM before
the timed region.  A step of the chosen form is one call of
geneig_mixed_n54_batched_launch (complex64 reduction / tridiagonalization /
eigenvectors, FP64 Cholesky, FP64 refinement on the FP64 matrix cores, FP64
re-solve of any pair the refinement hands back); a step of the plain form is
one launch of the drop-in geneig_full_n54_batched_launch.  Both run on the
same inputs; `value` is the chosen form, `plain` the FP64 form.

The JSON line carries the roofline of one solve launch (algorithmic FP64 work
F(n) = 32 n^3 flop per pair, LAPACK zhegv-equivalent, over the launch's
HIP-event duration on its own stream, against the 78.6 TFLOP/s FP64 peak) and
a CPU baseline: LAPACK zhegv (the routine the reference's CPU harness calls),
OpenMP over pairs on the host cores, median of >= 5 passes over a bounded
sample of the workload, with the oracle (this project's C port of the
pipeline) beside it.  The LAPACK solutions of that sample are also the
accuracy reference of both forms (eigenvalue error on 16,384 pairs), and the
residual and B-orthogonality of every pair of the batch are checked on the
device; `accuracy_gate` says whether both forms meet the metric's bars.

Usage: python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--batch 131072]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

FP64_PEAK = 78.6e12      # MI355X dense FP64 (vector and matrix), FLOP/s
HBM_PEAK = 8.0e12        # MI355X HBM3E, B/s


def flops_per_pair(n):
    return 32.0 * n ** 3


def bytes_per_pair(n):
    return 48.0 * n * n + 8.0 * n + 4.0


def load_package():
    import importlib.util

    pkg = os.path.join(ROOT, "sunny-batched-geneig_amd")
    name = "sunny_batched_geneig_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def accuracy(A, B, W, X):
    """Eigen-errors of sampled pairs against LAPACK (scipy.linalg.eigh(a, b),
    i.e. zhegvd) and the scale-invariant residual / B-orthogonality measures.
    Storage arrays: pair b, element (i, j) at [b, j, i]."""
    import scipy.linalg as sl

    n = A.shape[-1]
    eps = np.finfo(np.float64).eps
    e_max = e_sc = e_rel = res = orth = 0.0
    for b in range(A.shape[0]):
        a, bm = A[b].T, B[b].T
        w = sl.eigh(a, bm, eigvals_only=True)
        d = np.abs(W[b] - w)
        e_max = max(e_max, float(np.max(d) / np.max(np.abs(w))))
        e_sc = max(e_sc, float(np.max(d / np.maximum(np.abs(w), 1.0))))
        e_rel = max(e_rel, float(np.max(d / np.abs(w))))
        x = X[b].T
        bx = bm @ x
        r = a @ x - bx * W[b][None, :]
        xn = np.linalg.norm(x, axis=0)
        fa, fb = np.linalg.norm(a), np.linalg.norm(bm)
        res = max(res, float(np.max(np.linalg.norm(r, axis=0) / ((fa + np.abs(W[b]) * fb) * xn))))
        g = x.conj().T @ bx - np.eye(n)
        orth = max(orth, float(np.max(np.abs(g) / (fb * np.outer(xn, xn)))))
    return {"pairs_checked": int(A.shape[0]), "reference": "scipy.linalg.eigh (LAPACK zhegvd)",
            "max_eig_err_over_max_abs_eig": e_max,
            "max_eig_rel_err_floor1": e_sc, "max_eig_rel_err": e_rel,
            "max_resid_over_n_eps": float(res / (n * eps)), "max_orth_over_n_eps": float(orth / (n * eps))}


def cpu_baseline(A, B, target_s=12.0, min_passes=5):
    """LAPACK zhegv and the oracle port, OpenMP over pairs on every host core,
    on a bounded sample of the same inputs (every pair of the host copy given,
    16,384 by default: about a second per pass); the rate is the median over
    >= min_passes passes (more while under target_s).  Returns (lapack, port,
    (W, X, info, m, R) of LAPACK on the sample with R the long-double
    Rayleigh quotients of its eigenvectors)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    cores = len(os.sched_getaffinity(0))
    n = A.shape[-1]
    out, sol = {}, None
    for kind, fn in (("lapack", O.lapack_zhegv), ("port", O.zhegv)):
        try:
            if kind == "lapack" and O.lapack_load() is None:
                log("no LAPACK library with zhegv: the CPU baseline is the oracle port alone")
                continue
            m = A.shape[0]
            fn(A[:64], B[:64], nthreads=cores)       # threads and library warmed up
            times = []
            while len(times) < min_passes or (sum(times) < target_s and len(times) < 4 * min_passes):
                t = time.perf_counter()
                res = fn(A[:m], B[:m], nthreads=cores)
                times.append(time.perf_counter() - t)
            med = float(np.median(times))
            out[kind] = {"value": m / med, "unit": "solves/s", "cores": cores,
                         "sample": "median of %d passes over %d of the workload's n=%d pairs (%.2f s per pass)"
                                   % (len(times), m, n, med)}
            if kind == "lapack":
                out[kind]["library"] = O.lib().go_lapack_name().decode()
                sol = (res[0], res[1], res[2], m, O.rayleigh_ld_batch(A[:m], B[:m], res[1], nthreads=cores))
        except Exception as e:  # a baseline must never sink the GPU measurement
            log("cpu baseline %s failed: %s" % (kind, e))
    return out.get("lapack"), out.get("port"), sol


def eig_errors(W, Wr):
    """Eigenvalue errors against a reference, per the metric: the largest
    |dlambda| / max(|lambda|, 1) (the reference's scaling,
    tests/test_full_n54.cpp:147-148), |dlambda| / max|lambda| per pair, and the
    pure relative error."""
    d = np.abs(W - Wr)
    return {"max_eig_rel_err_floor1": float((d / np.maximum(np.abs(Wr), 1.0)).max()),
            "max_eig_err_over_max_abs_eig": float((d.max(axis=1) / np.abs(Wr).max(axis=1)).max()),
            "max_eig_rel_err": float((d / np.abs(Wr)).max())}


def device_check(torch, A, B, W, U, info, chunk=8192):
    """Residual and B-orthogonality of every pair on the device, in the
    scale-invariant per-eigenpair form of the reference harness
    (tests/lapack_reference.cpp:207-254): max_j ||A x_j - lambda_j B x_j|| /
    ((||A||_F + |lambda_j| ||B||_F) ||x_j||) and max_ij |x_i^H B x_j - delta_ij|
    / (||B||_F ||x_i|| ||x_j||), over pairs with info = 0, in units of n eps.
    Storage holds each matrix transposed (column-major), so with At = A^T,
    Xt = X^T: (A X)^T = Xt At and (X^H B X)^T = (Xt Bt) conj(Xt)^T."""
    n = A.shape[-1]
    eps = float(np.finfo(np.float64).eps)
    res = orth = 0.0
    bad = int((info != 0).sum().item())
    for off in range(0, A.shape[0], chunk):
        sl = slice(off, off + chunk)
        ok = info[sl] == 0
        if not bool(ok.any()):
            continue
        At, Bt, Xt, w = A[sl][ok], B[sl][ok], U[sl][ok], W[sl][ok]
        XB = torch.matmul(Xt, Bt)
        R = torch.matmul(Xt, At) - w[:, :, None] * XB
        xn = torch.linalg.vector_norm(Xt, dim=2)
        fa = torch.linalg.matrix_norm(At)
        fb = torch.linalg.matrix_norm(Bt)
        r = torch.linalg.vector_norm(R, dim=2) / ((fa[:, None] + w.abs() * fb[:, None]) * xn)
        res = max(res, float(r.max().item()))
        G = torch.matmul(XB, Xt.conj().transpose(-1, -2).contiguous())
        G = G - torch.eye(n, dtype=G.dtype, device=G.device)[None]
        o = G.abs() / (fb[:, None, None] * xn[:, :, None] * xn[:, None, :])
        orth = max(orth, float(o.max().item()))
    return {"pairs": int(A.shape[0]), "info_nonzero": bad, "max_resid_over_n_eps": res / (n * eps),
            "max_orth_over_n_eps": orth / (n * eps)}


def source_sha16():
    """Hash of the sources the library's kernels are built from (csrc/ and the
    Makefile): a measured traffic figure is used only for the build it came from."""
    import hashlib

    h = hashlib.sha256()
    pkg = os.path.join(ROOT, "sunny-batched-geneig_amd")
    for d in (os.path.join(pkg, "csrc"),):
        for f in sorted(os.listdir(d)):
            h.update(f.encode())
            with open(os.path.join(d, f), "rb") as fh:
                h.update(fh.read())
    with open(os.path.join(pkg, "Makefile"), "rb") as fh:
        h.update(fh.read())
    return h.hexdigest()[:16]


MIXED_SIZES = [8, 16, 24, 32, 48, 54, 64, 96]
MIXED_COUNTS = [72304, 46998, 30548, 19856, 12907, 8389, 5453, 3545]   # 0.35*0.65^k, 200,000 pairs


def bench_mixed_sizes(args, geneig, torch, dev):
    """Mixed-size batch: generic pairs (A = (M+M^H)/(2 sqrt n), B = G G^H/n + I)
    binned by n, solved by one graph launch (one branch per size, per-size
    pairs-per-workgroup from the autotuner) against one launch per size."""
    t0 = time.perf_counter()
    m = geneig.GeneigMulti(MIXED_SIZES, tuning_mode=geneig.TUNING_AUTOTUNE)
    log("solvers ready in %.1f s; pairs per workgroup %s"
        % (time.perf_counter() - t0, [m.batches_per_block(k) for k in range(len(MIXED_SIZES))]))
    bins = []
    for k, (n, c) in enumerate(zip(MIXED_SIZES, MIXED_COUNTS)):
        A, B = geneig.gen_generic(n, c, seed=args.seed + k)
        W, U, info = geneig.empty_outputs(n, c, dev)
        bins.append((A, B, W, U, info))
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    total = sum(MIXED_COUNTS)

    def timed(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        t_start = time.perf_counter()
        for e0, e1 in evs:
            e0.record(stream)
            step()
            e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        return wall, float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs])) * 1e-3

    wall_s, launch_s = timed(lambda: m.launch_serial(bins, stream=stream))
    wall_g, launch_g = timed(lambda: m.launch(bins, stream=stream))
    log("serial %.2f ms/step, graph %.2f ms/step (captures %d)"
        % (wall_s * 1e3 / args.steps, wall_g * 1e3 / args.steps, m.captures()))
    acc = {}
    for k, n in enumerate(MIXED_SIZES):
        A, B, W, U, info = bins[k]
        sel = torch.arange(0, A.shape[0], max(1, A.shape[0] // 16), device=dev)[:16]
        a = accuracy(A.index_select(0, sel).cpu().numpy(), B.index_select(0, sel).cpu().numpy(),
                     W.index_select(0, sel).cpu().numpy(), U.index_select(0, sel).cpu().numpy())
        acc[str(n)] = {"max_eig_err_over_max_abs_eig": a["max_eig_err_over_max_abs_eig"],
                       "max_resid_over_n_eps": a["max_resid_over_n_eps"],
                       "max_orth_over_n_eps": a["max_orth_over_n_eps"], "info_nonzero": int((info != 0).sum().item())}
    flops = sum(c * flops_per_pair(n) for n, c in zip(MIXED_SIZES, MIXED_COUNTS))
    nbytes = sum(c * bytes_per_pair(n) for n, c in zip(MIXED_SIZES, MIXED_COUNTS))
    line = {
        "metric": "mixed-size geneig solves/s (n in {8,16,24,32,48,54,64,96}, 200,000 pairs), one graph launch",
        "value": total / (wall_g / args.steps), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall_g * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (device-generated generic pairs, seed %d + bin)" % args.seed,
        "config": {"workload": "generic pairs binned by n", "sizes": MIXED_SIZES, "counts": MIXED_COUNTS,
                   "pairs_per_workgroup": [m.batches_per_block(k) for k in range(len(MIXED_SIZES))],
                   "parallelism": "none (1 GPU)"},
        "serial": {"value": total / (wall_s / args.steps), "ms_per_step": wall_s * 1e3 / args.steps},
        "speedup_vs_serial": wall_s / wall_g,
        "accuracy_sampled_16_per_size": acc,
        "roofline": {"bound": "mfma", "achieved": flops / launch_g / 1e12, "peak": FP64_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": flops / launch_g / FP64_PEAK, "traffic": None,
                     "kernel": "one graph launch (all bins)", "avg_launch_ms": launch_g * 1e3,
                     "hbm_frac": nbytes / launch_g / HBM_PEAK},
    }
    print(json.dumps(line), flush=True)


def bench_learned(args, geneig, torch, dev):
    """Learned warm start on the n = 54 spin-wave family with a coupling scale
    per q (s ~ U(0.8, 1.2)): train on 32,768 pairs (32^3 q grid), evaluate on
    262,144 unseen q (64^3 grid, same lattice), against the same pipeline from
    Q = I, the precision form and the plain FP64 form on the same pairs."""
    L = geneig.learn
    fam = 11
    t0 = time.perf_counter()
    A, B, P = L.spinwave_family(32768, fam, (32, 32, 32))
    Cm, info = L.reduced_matrices(A, B)
    X = L.features(P)
    del A, B
    torch.cuda.synchronize()
    log("training set ready in %.1f s (info nonzero %d)" % (time.perf_counter() - t0, int((info != 0).sum().item())))
    m = L.WarmStart(seed=args.seed, device=dev)
    t0 = time.perf_counter()
    hist = m.train(X, Cm, epochs=args.train_epochs, batch=256, lr=2e-3, lr_final=1e-4, seed=args.seed,
                   log=lambda msg: log(msg))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del Cm, X
    nev = args.eval_batch
    Ae, Be, Pe = L.spinwave_family(nev, fam, (64, 64, 64))
    Xe = L.features(Pe)
    W, U, info = geneig.empty_outputs(54, nev, dev)
    sw = torch.zeros(nev, dtype=torch.int32, device=dev)
    it = torch.zeros(nev, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    sel = torch.from_numpy(np.linspace(0, nev - 1, num=min(args.check, nev)).astype(np.int64)).to(dev)
    Ah, Bh = Ae.index_select(0, sel).cpu().numpy(), Be.index_select(0, sel).cpu().numpy()
    out = {}

    def run(name, step, steps):
        step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            r = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        acc = accuracy(Ah, Bh, W.index_select(0, sel).cpu().numpy(), U.index_select(0, sel).cpu().numpy())
        out[name] = {"solves_per_s": nev / dt, "ms_per_step": dt * 1e3, "fallback": r,
                     "max_eig_err_over_max_abs_eig": acc["max_eig_err_over_max_abs_eig"],
                     "max_resid_over_n_eps": acc["max_resid_over_n_eps"],
                     "max_orth_over_n_eps": acc["max_orth_over_n_eps"]}
        log("%s: %s" % (name, out[name]))

    steps = max(1, min(args.steps, 3))
    run("learned", lambda: L.learned_n54_batched(Ae, Be, W, U, info, X=Xe, model=m, sweeps=sw, iters=it,
                                                  max_sweeps=30), steps)
    sw_l = sw.cpu().numpy()
    it_l = it.cpu().numpy()
    run("jacobi_from_identity", lambda: L.learned_n54_batched(Ae, Be, W, U, info, sweeps=sw, iters=it, max_sweeps=30),
        steps)
    sw_i = sw.cpu().numpy()
    run("precision", lambda: geneig.mixed_n54_batched(Ae, Be, W, U, info), steps)
    run("plain", lambda: (geneig.full_n54_batched(Ae, Be, W, U, info), 0)[1], steps)

    def hist_of(a):
        v, c = np.unique(a, return_counts=True)
        return {str(int(k)): int(n) for k, n in zip(v, c)}

    line = {
        "metric": "learned warm start: n=54 spin-wave solves/s on unseen q, Jacobi sweeps saved, fallback rate",
        "value": out["learned"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": out["learned"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 network + c64 Jacobi + f64 refinement",
        "data": "synthetic (device-generated spin-wave family, lattice seed %d, s ~ U(0.8,1.2) per q)" % fam,
        "config": {"workload": "train 32,768 pairs (32^3 q grid), evaluate %d unseen q (64^3 grid)" % nev,
                   "network": "8 -> 256 -> 256 -> 256 (tanh) -> 5832, FP32", "train_epochs": args.train_epochs,
                   "train_batch": 256, "train_seconds": train_s, "parallelism": "none (1 GPU)"},
        "train_loss_first_last": [hist[0], hist[-1]],
        "train_loss_best": m.train_best, "train_rollbacks": m.rollbacks,
        "jacobi_sweeps_mean": {"network": float(sw_l.mean()), "identity": float(sw_i.mean())},
        "jacobi_sweeps_saved_mean": float(sw_i.mean() - sw_l.mean()),
        "jacobi_sweeps_histogram": {"network": hist_of(sw_l), "identity": hist_of(sw_i)},
        "refine_steps_histogram_network": hist_of(it_l),
        "fallback_rate": out["learned"]["fallback"] / nev,
        "forms": out,
    }
    print(json.dumps(line), flush=True)


def bench_generic(args, geneig, torch, dev):
    """Generic pairs at any n through the runtime (hipRTC) solvers: the FP64 form
    and the precision form on the same batch, device-resident, HIP-event timed."""
    n, batch = args.n, args.batch
    A, B = geneig.gen_generic(n, batch, seed=args.seed)
    W, U, info = geneig.empty_outputs(n, batch, dev)
    sp, sm = geneig.GeneigSolver(n), geneig.GeneigMixedSolver(n)
    iters = torch.zeros(batch, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def timed(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.steps

    plain_s = timed(lambda: sp.launch(A, B, W, U, info, stream=stream))
    Wp = W.clone()
    sel = torch.arange(0, batch, max(1, batch // args.check), device=dev)[: args.check]
    acc_p = accuracy(A.index_select(0, sel).cpu().numpy(), B.index_select(0, sel).cpu().numpy(),
                     W.index_select(0, sel).cpu().numpy(), U.index_select(0, sel).cpu().numpy())
    fb = [0]

    def mixed():
        # the fallback count stays on the device inside the timed steps; it is
        # read once afterwards
        sm.launch_mixed(A, B, W, U, info, iters=iters, stream=stream, report_fallback=False)

    mixed_s = timed(mixed)
    fb[0] = sm.launch_mixed(A, B, W, U, info, iters=iters, stream=stream)
    acc_m = accuracy(A.index_select(0, sel).cpu().numpy(), B.index_select(0, sel).cpu().numpy(),
                     W.index_select(0, sel).cpu().numpy(), U.index_select(0, sel).cpu().numpy())
    v, c = np.unique(iters.cpu().numpy(), return_counts=True)
    log("n=%d: FP64 %.1f ms, precision %.1f ms" % (n, plain_s * 1e3, mixed_s * 1e3))
    line = {
        "metric": "n=%d generic geneig solves/s, precision vs plain FP64 form (runtime solvers)" % n,
        "value": batch / mixed_s, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": mixed_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 bulk + f64 (Cholesky, refinement)", "data": "synthetic (device-generated generic pairs, seed %d)"
        % args.seed,
        "config": {"workload": "generic pairs A = (M + M^H)/(2 sqrt n), B = G G^H/n + I, n=%d, %d pairs" % (n, batch),
                   "n": n, "global_batch": batch, "parallelism": "none (1 GPU)"},
        "accuracy": acc_m, "fallback_pairs": int(fb[0]),
        "refine_steps_histogram": {str(int(a)): int(b) for a, b in zip(v, c)},
        "plain": {"value": batch / plain_s, "ms_per_step": plain_s * 1e3, "accuracy": acc_p},
        "speedup_vs_plain": plain_s / mixed_s,
        "roofline": {"bound": "mfma", "achieved": flops_per_pair(n) * batch / mixed_s / 1e12,
                     "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": flops_per_pair(n) * batch / mixed_s / FP64_PEAK, "traffic": None,
                     "kernel": "one precision-form launch (wall time per step)"},
    }
    print(json.dumps(line), flush=True)


def with_cond(torch, B, cond, geneig, n, chunk=8192):
    """Spin-wave B = H shifted to cond(H) = cond: H + (lambda_max / cond -
    lambda_min) I, the extremes from the library's own solve of (H, I)."""
    Ic = torch.zeros((chunk, n, n), dtype=torch.complex128, device=B.device)
    idx = torch.arange(n, device=B.device)
    Ic[:, idx, idx] = 1.0
    for off in range(0, B.shape[0], chunk):
        m = min(chunk, B.shape[0] - off)
        Wc, Uc, ic = geneig.empty_outputs(n, m, B.device)
        geneig.full_n54_batched(B[off:off + m], Ic[:m], Wc, Uc, ic)
        sh = Wc[:, -1] / cond - Wc[:, 0]
        B[off:off + m, idx, idx] += sh[:, None].to(torch.complex128)
    del Ic


def generic_with_cond(torch, geneig, n, batch, cond, seed, dev, chunk=8192):
    """Generic A = (M + M^H) / (2 sqrt n) (the library's generator) and
    B = Q D Q^H with D = diag(logspace(0, log10 cond, n)) and Q a product of
    four Householder reflectors with complex Gaussian vectors (exactly unitary
    to rounding, so cond(B) = cond)."""
    A, _ = geneig.gen_generic(n, batch, seed=seed)
    B = torch.empty_like(A)
    g = torch.Generator(device=dev).manual_seed(seed)
    d = torch.logspace(0.0, float(np.log10(cond)), n, dtype=torch.float64, device=dev)
    for off in range(0, batch, chunk):
        m = min(chunk, batch - off)
        M = torch.diag_embed(d.to(torch.complex128)).expand(m, n, n).contiguous()
        for _ in range(4):
            v = torch.complex(torch.randn((m, n, 1), dtype=torch.float64, device=dev, generator=g),
                              torch.randn((m, n, 1), dtype=torch.float64, device=dev, generator=g))
            v = v / torch.linalg.vector_norm(v, dim=1, keepdim=True)
            H = torch.eye(n, dtype=torch.complex128, device=dev)[None] - 2.0 * torch.matmul(
                v, v.conj().transpose(-1, -2).contiguous())
            M = torch.matmul(torch.matmul(H, M), H)
        B[off:off + m] = 0.5 * (M + M.conj().transpose(-1, -2))   # Hermitian to the last bit
    return A, B


def bench_precision_sweep(args, geneig, torch, dev):
    """The precision form against the plain FP64 form with cond(B) swept over
    1e2, 1e4, 1e6, 1e8: n = 54 spin-wave pairs (H shifted to the condition)
    and n = 96 generic pairs (B = Q D Q^H), --batch pairs each.  Per (family,
    cond): solves/s of both forms, fallback rate, refinement steps, eigenvalue
    errors of both forms against LAPACK zhegv on a sample, and residual /
    B-orthogonality of every pair on the device."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    have_lapack = O.lapack_load() is not None
    batch = args.batch
    stream = torch.cuda.current_stream()
    rows = []
    for fam, n in (("spin-wave", 54), ("generic", 96)):
        sp = geneig.GeneigSolver(n) if n != 54 else None
        sm = geneig.GeneigMixedSolver(n) if n != 54 else None
        if fam == "spin-wave":
            A, B0 = geneig.gen_spinwave(27, batch, seed=args.seed, grid=(64, 64, 32))
        for cond in (1e2, 1e4, 1e6, 1e8):
            t0 = time.perf_counter()
            if fam == "spin-wave":
                B = B0.clone()
                with_cond(torch, B, cond, geneig, n)
            else:
                A, B = generic_with_cond(torch, geneig, n, batch, cond, args.seed + int(np.log10(cond)), dev)
            torch.cuda.synchronize()
            log("%s n=%d cond %.0e: pairs ready in %.1f s" % (fam, n, cond, time.perf_counter() - t0))
            W, U, info = geneig.empty_outputs(n, batch, dev)
            iters = torch.zeros(batch, dtype=torch.int32, device=dev)

            def plain():
                if sp is None:
                    geneig.full_n54_batched(A, B, W, U, info, stream=stream)
                else:
                    sp.launch(A, B, W, U, info, stream=stream)

            def mixed(report=False):
                if sm is None:
                    return geneig.mixed_n54_batched(A, B, W, U, info, iters=iters, stream=stream,
                                                    report_fallback=report)
                return sm.launch_mixed(A, B, W, U, info, iters=iters, stream=stream, report_fallback=report)

            def timed(step):
                step()
                torch.cuda.synchronize()
                t = time.perf_counter()
                for _ in range(args.steps):
                    step()
                torch.cuda.synchronize()
                return (time.perf_counter() - t) / args.steps

            tp = timed(plain)
            Wp = W.clone()
            dp = device_check(torch, A, B, W, U, info)
            tm = timed(mixed)
            fb = mixed(True)
            torch.cuda.synchronize()
            dm = device_check(torch, A, B, W, U, info)
            k = min(batch, args.check)
            idx = torch.from_numpy(np.linspace(0, batch - 1, num=k).astype(np.int64)).to(dev)
            Ah, Bh = A.index_select(0, idx).cpu().numpy(), B.index_select(0, idx).cpu().numpy()
            Wr, Xr, ir = O.lapack_zhegv(Ah, Bh) if have_lapack else O.zhegv(Ah, Bh)
            ok = ir == 0
            Wph, Wmh = Wp.index_select(0, idx).cpu().numpy()[ok], W.index_select(0, idx).cpu().numpy()[ok]
            ep = eig_errors(Wph, Wr[ok])
            em = eig_errors(Wmh, Wr[ok])
            # at large cond(B) every FP64 eigenvalue, LAPACK's included, carries
            # ~eps cond(B): long-double Rayleigh quotients of the reference's
            # eigenvectors (second order in their error) are the sharp reference
            Rld = O.rayleigh_ld_batch(Ah[ok], Bh[ok], Xr[ok])
            ep["vs_long_double"] = eig_errors(Wph, Rld)["max_eig_rel_err_floor1"]
            em["vs_long_double"] = eig_errors(Wmh, Rld)["max_eig_rel_err_floor1"]
            ref_ld = eig_errors(Wr[ok], Rld)["max_eig_rel_err_floor1"]
            v, c = np.unique(iters.cpu().numpy(), return_counts=True)
            row = {"family": fam, "n": n, "cond_B": cond, "pairs": batch,
                   "precision": {"solves_per_s": batch / tm, "ms_per_step": tm * 1e3, "fallback_rate": fb / batch,
                                 "refine_steps_histogram": {str(int(a)): int(b) for a, b in zip(v, c)},
                                 "eig_vs_lapack": em, "device_full_batch": dm},
                   "plain": {"solves_per_s": batch / tp, "ms_per_step": tp * 1e3, "eig_vs_lapack": ep,
                             "device_full_batch": dp},
                   "speedup": tp / tm, "eig_pairs_checked": int(ok.sum()),
                   "eig_reference": "LAPACK zhegv" if have_lapack else "oracle",
                   "reference_vs_long_double": ref_ld}
            rows.append(row)
            log("%s n=%d cond %.0e: plain %.0f/s, precision %.0f/s (x%.2f), fallback %d, eig vs LAPACK %.2e / %.2e, "
                "vs long double %.2e / %.2e (LAPACK's own %.2e)"
                % (fam, n, cond, batch / tp, batch / tm, tp / tm, fb, ep["max_eig_rel_err_floor1"],
                   em["max_eig_rel_err_floor1"], ep["vs_long_double"], em["vs_long_double"], ref_ld))
            del W, U, info, Wp
            if fam == "generic":
                del A, B
            torch.cuda.empty_cache()
    head = [r for r in rows if r["family"] == "spin-wave" and r["cond_B"] == 1e4][0]
    line = {
        "metric": "precision form vs plain FP64 form with cond(B) swept 1e2..1e8 (n=54 spin-wave, n=96 generic)",
        "value": head["precision"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": head["precision"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 bulk + f64 (Cholesky, refinement)",
        "data": "synthetic (device-generated pairs, seed %d)" % args.seed,
        "config": {"workload": "n=54 spin-wave and n=96 generic pairs, %d each, cond(B) in 1e2..1e8" % batch,
                   "parallelism": "none (1 GPU)"},
        "rows": rows,
    }
    print(json.dumps(line), flush=True)


def bench_spinwave_fused(args, geneig, torch, dev):
    """The fused spin-wave solve (H(q) and eta built inside the solve from the
    bond tables) against the drop-in on materialized pairs and against
    generating the pairs then solving them, n = 54, --batch q."""
    n, batch = 54, args.batch
    grid = (64, 64, 32)
    model = geneig.spinwave_model(27, args.seed, device=dev)
    q = geneig.spinwave_q(batch, grid, device=dev)
    A, B = geneig.gen_spinwave(27, batch, seed=args.seed, grid=grid)
    H0 = geneig.spinwave_n54_build(model, q)
    diag = (B[:, 0, 0].real - H0[:, 0, 0].real).contiguous()
    del H0
    W, U, info = geneig.empty_outputs(n, batch, dev)
    stream = torch.cuda.current_stream()

    def timed(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.steps

    def gen_and_solve():
        Hs = geneig.spinwave_n54_build(model, q, diag, stream=stream)
        geneig.full_n54_batched(A, Hs, W, U, info, stream=stream)

    tf = timed(lambda: geneig.spinwave_n54(model, q, diag, W, U, info, stream=stream))
    dfu = device_check(torch, A, B, W, U, info)
    td = timed(lambda: geneig.full_n54_batched(A, B, W, U, info, stream=stream))
    tg = timed(gen_and_solve)
    log("fused %.2f ms, drop-in on stored pairs %.2f ms, build + drop-in %.2f ms" % (tf * 1e3, td * 1e3, tg * 1e3))
    line = {
        "metric": "n=54 spin-wave solves/s with H(q) built inside the solve (fused), vs the drop-in",
        "value": batch / tf, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": tf * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (spin-wave model of seed %d, %d q on a 64x64x32 grid)" % (args.seed, batch),
        "config": {"workload": "spin-wave (eta, H(q) + diag I), n=54, 12 bonds", "global_batch": batch,
                   "parallelism": "none (1 GPU)"},
        "device_full_batch": dfu,
        "drop_in_on_stored_pairs": {"value": batch / td, "ms_per_step": td * 1e3},
        "build_then_drop_in": {"value": batch / tg, "ms_per_step": tg * 1e3},
        "hbm_bytes_not_moved_per_step": batch * 2 * 16 * n * n,
    }
    print(json.dumps(line), flush=True)


def paired_exact_count(torch, W, U, chunk=16384):
    """Pairs whose outputs carry the spin-wave pairing bit for bit: W[n-1-i] =
    -W[i] and X(:, i) = Sigma conj X(:, n-1-i), Sigma swapping the halves of
    a column (storage is column-major: U[p, col, row])."""
    h = U.shape[-1] // 2
    cnt = 0
    for off in range(0, W.shape[0], chunk):
        w, u = W[off:off + chunk], U[off:off + chunk]
        part = torch.cat([u[:, :, h:], u[:, :, :h]], dim=2).conj().flip(1)
        cnt += int(((w + w.flip(1) == 0).all(dim=1) & (u == part).flatten(1).all(dim=1)).sum().item())
    return cnt


def max_rel_dw(torch, W1, W2):
    """max over pairs of max_i |W1 - W2| / max(max_i |W1|, 1)."""
    return float(((W1 - W2).abs().amax(dim=1) / W1.abs().amax(dim=1).clamp(min=1.0)).max().item())


def bench_spinwave_eta(args, geneig, torch, dev):
    """The precision form for the pencil (eta, S) (geneig_mixed_eta_n54_batched_launch:
    A = eta not read, the refinement's A X a row scaling) against the general
    precision form and the plain FP64 drop-in on the same n = 54 spin-wave
    pairs of the headline workload; the eta form's outputs must equal the
    general form's bit for bit."""
    n, batch = 54, args.batch
    A, B = geneig.gen_spinwave(27, batch, seed=args.seed, grid=(64, 64, 32))
    W, U, info = geneig.empty_outputs(n, batch, dev)
    W2, U2, i2 = geneig.empty_outputs(n, batch, dev)
    it, it2 = (torch.zeros(batch, dtype=torch.int32, device=dev) for _ in range(2))
    stream = torch.cuda.current_stream()

    def timed(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.steps

    tg = timed(lambda: geneig.mixed_n54_batched(A, B, W, U, info, iters=it, stream=stream, report_fallback=False))
    te = timed(lambda: geneig.mixed_eta_n54_batched(B, W2, U2, i2, iters=it2, stream=stream, report_fallback=False))
    same = bool(torch.equal(W, W2) and torch.equal(U, U2) and torch.equal(info, i2) and torch.equal(it, it2))
    dev_e = device_check(torch, A, B, W2, U2, i2)
    # the paired launcher on the same pairs (S of H(q)'s block form): refines
    # the 27 positive eigenpairs, writes their partners
    tq = timed(lambda: geneig.mixed_eta_paired_n54_batched(B, W, U, info, stream=stream, report_fallback=False))
    fb_q = geneig.mixed_eta_paired_n54_batched(B, W, U, info, stream=stream)
    torch.cuda.synchronize()
    dev_q = device_check(torch, A, B, W, U, info)
    paired_q = paired_exact_count(torch, W, U)
    dw_q = max_rel_dw(torch, W2, W)
    # the real-arithmetic FP64 solve of the same block-form pencil
    tr = timed(lambda: geneig.eta_paired_n54_batched(B, W, U, info, stream=stream))
    torch.cuda.synchronize()
    dev_r = device_check(torch, A, B, W, U, info)
    paired_r = paired_exact_count(torch, W, U)
    dw_r = max_rel_dw(torch, W2, W)
    Wr = W.clone()
    tp = timed(lambda: geneig.full_n54_batched(A, B, W, U, info, stream=stream))
    torch.cuda.synchronize()
    dw_rp = max_rel_dw(torch, W, Wr)
    log("eta form %.2f ms, paired %.2f ms, real FP64 %.2f ms, general precision form %.2f ms, plain FP64 %.2f ms, "
        "outputs equal: %s" % (te * 1e3, tq * 1e3, tr * 1e3, tg * 1e3, tp * 1e3, same))
    v, c = np.unique(it2.cpu().numpy(), return_counts=True)
    line = {
        "metric": "n=54 spin-wave solves/s, precision form for the pencil (eta, S) (A = eta not read)",
        "value": batch / te, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": te * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 bulk + f64 (Cholesky, refinement)",
        "data": "synthetic (device-generated spin-wave pairs, seed %d)" % args.seed,
        "config": {"workload": "spin-wave (eta, H(q)) pairs, n=54, %d q on a 64x64x32 grid" % batch,
                   "global_batch": batch, "parallelism": "none (1 GPU)"},
        "outputs_equal_general_precision_form": same, "device_full_batch": dev_e,
        "refine_steps_histogram": {str(int(a)): int(b) for a, b in zip(v, c)},
        "general_precision_form": {"value": batch / tg, "ms_per_step": tg * 1e3},
        "paired_form": {"value": batch / tq, "ms_per_step": tq * 1e3,
                        "entry": "geneig_mixed_eta_paired_n54_batched_launch",
                        "fallback_pairs": fb_q, "exactly_paired": paired_q, "device_full_batch": dev_q,
                        "max_rel_eig_diff_vs_eta_form": dw_q},
        "real_form": {"value": batch / tr, "ms_per_step": tr * 1e3,
                      "entry": "geneig_eta_paired_n54_batched_launch (real FP64 arithmetic via T = [[I, iI], [I, -iI]]"
                               " / sqrt 2)",
                      "exactly_paired": paired_r, "device_full_batch": dev_r,
                      "max_rel_eig_diff_vs_eta_form": dw_r, "max_rel_eig_diff_vs_plain": dw_rp},
        "plain_fp64": {"value": batch / tp, "ms_per_step": tp * 1e3},
        "speedup_vs_general": tg / te, "speedup_vs_plain": tp / te,
    }
    print(json.dumps(line), flush=True)


def bench_host_chunked(args, geneig, torch, dev):
    """n = 54 spin-wave pairs held in pinned host memory: launch_chunked_host copies
    each chunk in, solves it and copies it out on its stream, so with several
    streams the copies overlap the solves.  A PCIe-inclusive rate, reported
    beside the device-resident rate of the same runtime solver."""
    n, batch, chunk = 54, args.batch, args.host_chunk
    A, B = geneig.gen_spinwave(27, batch, seed=args.seed, grid=(64, 64, 32))
    s = geneig.GeneigSolver(n)
    Wd, Ud, idd = geneig.empty_outputs(n, batch, dev)
    stream = torch.cuda.current_stream()

    def wall(step, sync):
        for _ in range(args.warmup):
            step()
        sync()
        t = time.perf_counter()
        for _ in range(args.steps):
            step()
        sync()
        return (time.perf_counter() - t) / args.steps

    dev_s = wall(lambda: s.launch(A, B, Wd, Ud, idd, stream=stream), torch.cuda.synchronize)
    log("device-resident launch %.1f ms" % (dev_s * 1e3))
    Ah, Bh = A.cpu().pin_memory(), B.cpu().pin_memory()
    del A, B
    W = torch.empty((batch, n), dtype=torch.float64, pin_memory=True)
    U = torch.empty((batch, n, n), dtype=torch.complex128, pin_memory=True)
    info = torch.empty((batch,), dtype=torch.int32, pin_memory=True)
    moved = batch * (3 * n * n * 16 + 8 * n + 4)
    rates = {}
    for ns in (1, 2, 4):
        streams = geneig.make_streams(ns)
        dt = wall(lambda: s.launch_chunked_host(Ah, Bh, W, U, info, total=batch, chunk_size=chunk, streams=streams),
                  lambda: [st.synchronize() for st in streams])
        equal = bool(np.array_equal(W.numpy(), Wd.cpu().numpy()) and np.array_equal(U.numpy(), Ud.cpu().numpy()))
        rates[ns] = {"value": batch / dt, "ms_per_step": dt * 1e3, "pcie_GBps_both_ways": moved / dt / 1e9,
                     "equal_to_device_launch": equal}
        log("%d stream(s): %.1f ms, %.1f GB/s over PCIe, equal to the device launch: %s"
            % (ns, dt * 1e3, moved / dt / 1e9, equal))
    line = {
        "metric": "n=54 geneig solves/s from pinned host memory (PCIe-inclusive), launch_chunked_host",
        "value": rates[4]["value"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": rates[4]["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128", "data": "synthetic (device-generated spin-wave pairs, seed %d, copied to pinned host memory)"
        % args.seed,
        "config": {"workload": "spin-wave (eta, H(q)) pairs, n=54, %d q, complex128, host-resident" % batch,
                   "chunk_size": chunk, "streams": 4, "parallelism": "none (1 GPU)"},
        "by_streams": {str(k): v for k, v in rates.items()},
        "device_resident": {"value": batch / dev_s, "ms_per_step": dev_s * 1e3},
        "bytes_moved_per_step": moved,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=131072)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--check", type=int, default=256, help="pairs checked against the oracle")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-pairs", type=int, default=16384, help="pairs of the CPU baseline sample (also the "
                                                                  "LAPACK accuracy reference)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["spinwave54", "mixed-sizes", "learned", "host-chunked", "generic",
                                           "precision-sweep", "spinwave-fused", "spinwave-eta"],
                    default="spinwave54",
                    help="spinwave54: the headline metric; mixed-sizes: 200,000 generic pairs over "
                         "n in {8,...,96}, one graph launch vs one launch per size; learned: train the "
                         "warm-start network on 32,768 solver-produced C, evaluate on 262,144 unseen q; "
                         "host-chunked: the n=54 pairs from pinned host memory through launch_chunked_host "
                         "(PCIe-inclusive, 1/2/4 streams); generic: generic pairs at --n through the runtime "
                         "FP64 and precision solvers; precision-sweep: both forms with cond(B) 1e2..1e8 on "
                         "n=54 spin-wave and n=96 generic pairs; spinwave-fused: H(q) built inside the solve")
    ap.add_argument("--n", type=int, default=96, help="matrix size of --workload generic")
    ap.add_argument("--host-chunk", type=int, default=8192, help="chunk size of --workload host-chunked")
    ap.add_argument("--train-epochs", type=int, default=40)
    ap.add_argument("--eval-batch", type=int, default=262144)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: pairs are independent, a multi-GPU run would "
              "split the batch into per-GPU ranges with no exchange; "
              "run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calib_copy_mib:
        # Reference kernel of known traffic for the FETCH_SIZE / WRITE_SIZE
        # counters: reads and writes exactly calib_copy_mib MiB per dispatch.
        nbytes = args.calib_copy_mib << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_per_dispatch": nbytes}), flush=True)
        return
    geneig = load_package()
    if args.workload == "mixed-sizes":
        bench_mixed_sizes(args, geneig, torch, dev)
        return
    if args.workload == "learned":
        bench_learned(args, geneig, torch, dev)
        return
    if args.workload == "host-chunked":
        bench_host_chunked(args, geneig, torch, dev)
        return
    if args.workload == "generic":
        bench_generic(args, geneig, torch, dev)
        return
    if args.workload == "precision-sweep":
        bench_precision_sweep(args, geneig, torch, dev)
        return
    if args.workload == "spinwave-fused":
        bench_spinwave_fused(args, geneig, torch, dev)
        return
    if args.workload == "spinwave-eta":
        bench_spinwave_eta(args, geneig, torch, dev)
        return
    n, nsub, batch = 54, 27, args.batch
    grid = (64, 64, 32)

    log("generating %d spin-wave pairs (n=%d) on the device" % (batch, n))
    t0 = time.perf_counter()
    A, B = geneig.gen_spinwave(nsub, batch, seed=args.seed, grid=grid)
    W, U, info = geneig.empty_outputs(n, batch, dev)
    iters = torch.zeros(batch, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    log("generated in %.1f s" % (time.perf_counter() - t0))
    stream = torch.cuda.current_stream()
    fallback = [0]

    def plain():
        geneig.full_n54_batched(A, B, W, U, info, stream=stream)

    def chosen():
        # the fallback count stays on the device inside the timed steps (the
        # launch then never waits for the host); it is read once afterwards
        geneig.mixed_n54_batched(A, B, W, U, info, iters=iters, stream=stream, report_fallback=False)

    def timed(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        # K steps, synchronized on both sides; events on the launch stream
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for e0, e1 in evs:
            e0.record(stream)
            step()
            e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        launch_s = float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs])) * 1e-3
        return wall, launch_s

    wall_p, launch_p = timed(plain)
    torch.cuda.synchronize()
    Wp, Up, ip = W.clone(), U.clone(), info.clone()
    dev_p = device_check(torch, A, B, Wp, Up, ip)
    log("plain FP64: %.3f ms/step, launch avg %.3f ms, device check %s"
        % (wall_p * 1e3 / args.steps, launch_p * 1e3, dev_p))
    wall_c, launch_c = timed(chosen)
    fallback[0] = geneig.mixed_n54_batched(A, B, W, U, info, iters=iters, stream=stream)
    torch.cuda.synchronize()
    dev_c = device_check(torch, A, B, W, U, info)
    hist = np.bincount(iters.cpu().numpy().astype(np.int64), minlength=1)
    log("chosen: %.3f ms/step, launch avg %.3f ms, fallback %d, device check %s"
        % (wall_c * 1e3 / args.steps, launch_c * 1e3, fallback[0], dev_c))

    # the precision form for the pencil (eta, S) on the same pairs (A = eta not
    # read; outputs equal the chosen form's bit for bit): a side figure, the
    # headline value stays the general form's
    We, Ue, ie = geneig.empty_outputs(n, batch, dev)
    wall_e, launch_e = timed(lambda: geneig.mixed_eta_n54_batched(B, We, Ue, ie, stream=stream,
                                                                  report_fallback=False))
    torch.cuda.synchronize()
    eta_same = bool(torch.equal(W, We) and torch.equal(U, Ue) and torch.equal(info, ie))
    log("eta form: %.3f ms/step, launch avg %.3f ms, outputs equal the chosen form's: %s"
        % (wall_e * 1e3 / args.steps, launch_e * 1e3, eta_same))
    # the paired launcher (S of H(q)'s block form: the refinement works on the
    # 27 positive eigenpairs and writes their partners) on the same pairs
    wall_q, launch_q = timed(lambda: geneig.mixed_eta_paired_n54_batched(B, We, Ue, ie, stream=stream,
                                                                         report_fallback=False))
    fb_q = geneig.mixed_eta_paired_n54_batched(B, We, Ue, ie, stream=stream)
    torch.cuda.synchronize()
    dev_q = device_check(torch, A, B, We, Ue, ie)
    paired_q = paired_exact_count(torch, We, Ue)
    dw_q = max_rel_dw(torch, W, We)
    # the real-arithmetic FP64 solve of the same block-form pencil
    wall_r, launch_r = timed(lambda: geneig.eta_paired_n54_batched(B, We, Ue, ie, stream=stream))
    torch.cuda.synchronize()
    dev_r = device_check(torch, A, B, We, Ue, ie)
    paired_r = paired_exact_count(torch, We, Ue)
    dw_r = max_rel_dw(torch, Wp, We)
    log("real form: %.3f ms/step, launch avg %.3f ms, exactly paired %d, |dW| vs plain %.2e, device check %s"
        % (wall_r * 1e3 / args.steps, launch_r * 1e3, paired_r, dw_r, dev_r))
    del We, Ue, ie
    log("paired form: %.3f ms/step, launch avg %.3f ms, fallback %d, exactly paired %d, |dW| vs chosen %.2e, "
        "device check %s" % (wall_q * 1e3 / args.steps, launch_q * 1e3, fb_q, paired_q, dw_q, dev_q))

    value = batch / (wall_c / args.steps)
    plain_value = batch / (wall_p / args.steps)
    achieved = flops_per_pair(n) * batch / launch_c
    traffic, traffic_note = None, "no traffic profile for this build"
    tpath = os.path.join(ROOT, "profiles", "traffic_geneig_n54_spinwave.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                t = json.load(f)
            if t.get("batch") == batch and t.get("source_sha16") == source_sha16():
                traffic = t.get("mixed", {}).get("bytes_per_launch")
                traffic_note = "rocprofv3 FETCH_SIZE + WRITE_SIZE of this build (%s)" % os.path.basename(tpath)
            else:
                traffic_note = "traffic profile is of another build or batch: not used"
        except Exception:
            traffic = None
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK, "traffic": traffic, "traffic_source": traffic_note,
                "kernel": "one chosen-form solve launch (n=54): chol64 -> k1f/k2af/k2bf/k2xf/k3f -> refine on each "
                          "of three concurrent parts (+ FP64 k1l/k2al/k2bl/k2xl/k3l over fallback pairs)",
                "flop_per_pair": flops_per_pair(n), "pairs_per_launch": batch,
                "avg_launch_ms": launch_c * 1e3,
                "hbm_achieved_GBps": bytes_per_pair(n) * batch / launch_c / 1e9,
                "hbm_frac": bytes_per_pair(n) * batch / launch_c / HBM_PEAK,
                "plain_avg_launch_ms": launch_p * 1e3,
                "plain_frac": flops_per_pair(n) * batch / launch_p / FP64_PEAK}

    lapack = port = sol = None
    acc_p = acc_c = None
    if not args.no_cpu:
        cb = min(batch, args.cpu_pairs)
        log("cpu baselines and LAPACK reference solutions (%d pairs)" % cb)
        lapack, port, sol = cpu_baseline(A[:cb].cpu().numpy(), B[:cb].cpu().numpy(), target_s=args.cpu_seconds)
    if sol is not None:
        Wr, Xr, ir, m, Rld = sol
        ok = ir == 0
        ref = ("long-double Rayleigh quotients of the eigenvectors of LAPACK zhegv (%s) on every pair of the CPU "
               "sample; errors against LAPACK's own eigenvalues under 'vs_lapack'" % lapack.get("library", "?"))
        accs = []
        for Wg in (Wp, W):
            Wh = Wg[:m].cpu().numpy()[ok]
            a_ = dict(pairs_checked=int(ok.sum()), reference=ref, **eig_errors(Wh, Rld[ok]))
            a_["vs_lapack"] = eig_errors(Wh, Wr[ok])
            accs.append(a_)
        acc_p, acc_c = accs
        lapack_vs_ld = eig_errors(Wr[ok], Rld[ok])["max_eig_rel_err_floor1"]
        log("LAPACK's own eigenvalues vs the long-double quotients: %.2e" % lapack_vs_ld)
        for a_ in accs:
            a_["lapack_vs_long_double"] = lapack_vs_ld
    else:
        # no LAPACK on this host: scipy's zhegvd on a sample of --check pairs
        idx = torch.from_numpy(np.linspace(0, batch - 1, num=min(args.check, batch)).astype(np.int64)).to(dev)
        Ah, Bh = A.index_select(0, idx).cpu().numpy(), B.index_select(0, idx).cpu().numpy()
        acc_p = accuracy(Ah, Bh, Wp.index_select(0, idx).cpu().numpy(), Up.index_select(0, idx).cpu().numpy())
        acc_c = accuracy(Ah, Bh, W.index_select(0, idx).cpu().numpy(), U.index_select(0, idx).cpu().numpy())
    for a_, d_ in ((acc_p, dev_p), (acc_c, dev_c)):
        a_["device_full_batch"] = d_
    bar = 1e-12

    def passes(a_):
        d_ = a_["device_full_batch"]
        return (a_["max_eig_rel_err_floor1"] <= bar and d_["info_nonzero"] == 0
                and d_["max_resid_over_n_eps"] <= 50.0 and d_["max_orth_over_n_eps"] <= 50.0)

    gate = {"eig_bar_rel_floor1": bar, "resid_orth_bar_n_eps": 50.0, "eig_pairs_checked": acc_c["pairs_checked"],
            "resid_orth_pairs_checked": batch, "chosen_pass": passes(acc_c), "plain_pass": passes(acc_p)}
    gate["pass"] = gate["chosen_pass"] and gate["plain_pass"]
    log("accuracy gate: %s" % gate)

    line = {
        "metric": "n=54 geneig solves/s, chosen vs plain FP64 form, max rel eig err <=1e-12",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_c * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 bulk + f64 (Cholesky, refinement)",
        "data": "synthetic (device-generated spin-wave pairs, seed %d)" % args.seed,
        "config": {"workload": "spin-wave (eta, H(q)) pairs, n=54 (27 sublattices), %d q on a "
                               "64x64x32 grid, complex128" % batch,
                   "form": "precision: complex64 reduction/tridiagonalization/eigenvectors, FP64 Cholesky, "
                           "FP64 Ogita-Aishima refinement on FP64 MFMA, FP64 re-solve of non-converged pairs",
                   "global_batch": batch, "n": n, "parallelism": "none (1 GPU)"},
        "accuracy": acc_c,
        "accuracy_gate": gate,
        "fallback_pairs": fallback[0],
        "refine_steps_histogram": {str(k): int(v) for k, v in enumerate(hist) if v},
        "plain": {"value": plain_value, "unit": "solves/s", "ms_per_step": wall_p * 1e3 / args.steps,
                  "dtype": "c128", "entry": "geneig_full_n54_batched_launch", "accuracy": acc_p},
        "speedup_vs_plain": value / plain_value,
        "eta_form": {"value": batch / (wall_e / args.steps), "unit": "solves/s",
                     "ms_per_step": wall_e * 1e3 / args.steps, "avg_launch_ms": launch_e * 1e3,
                     "entry": "geneig_mixed_eta_n54_batched_launch (A = eta not read)",
                     "outputs_equal_chosen_form": eta_same,
                     "speedup_vs_plain": (batch / (wall_e / args.steps)) / plain_value},
        "paired_form": {"value": batch / (wall_q / args.steps), "unit": "solves/s",
                        "ms_per_step": wall_q * 1e3 / args.steps, "avg_launch_ms": launch_q * 1e3,
                        "entry": "geneig_mixed_eta_paired_n54_batched_launch (A = eta, S = [[P, Q], [conj Q, "
                                 "conj P]]: refinement of the 27 positive eigenpairs, partners written)",
                        "fallback_pairs": fb_q, "exactly_paired": paired_q, "device_full_batch": dev_q,
                        "max_rel_eig_diff_vs_chosen": dw_q,
                        "speedup_vs_plain": (batch / (wall_q / args.steps)) / plain_value},
        "real_form": {"value": batch / (wall_r / args.steps), "unit": "solves/s",
                      "ms_per_step": wall_r * 1e3 / args.steps, "avg_launch_ms": launch_r * 1e3,
                      "entry": "geneig_eta_paired_n54_batched_launch (A = eta, S = [[P, Q], [conj Q, conj P]] mapped "
                               "by T = [[I, iI], [I, -iI]] / sqrt 2 to a real pencil, solved in FP64)",
                      "dtype": "f64 (real)", "exactly_paired": paired_r, "device_full_batch": dev_r,
                      "max_rel_eig_diff_vs_plain": dw_r,
                      "speedup_vs_plain": (batch / (wall_r / args.steps)) / plain_value},
        "roofline": roofline,
        "cpu_baseline": (dict(lapack, kind="reference",
                              routine="LAPACK zhegv (itype 1, 'V', 'L'), the solver of the reference's CPU "
                                      "harness (tests/lapack_reference.cpp), one pair per OpenMP thread")
                         if lapack else (dict(port, kind="port") if port else None)),
        "cpu_port": (dict(port, kind="port", routine="oracle/geneig_oracle.c, this project's C restatement of "
                                                     "the pipeline") if port else None),
    }
    print(json.dumps(line), flush=True)
    if not gate["pass"]:
        log("accuracy gate FAILED: %s" % gate)
        sys.exit(3)


if __name__ == "__main__":
    main()
