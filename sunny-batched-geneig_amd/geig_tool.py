"""GEIG v1 fixture tooling with the reference's command lines
(tests/lapack_reference.cpp, tests/dump_pair.cpp), solving on the GPU.

  python sunny-batched-geneig_amd/geig_tool.py make --n N --batch B --seed S --out FILE
         [--s-mode random|near-identity] [--check-residuals|--no-check-residuals]
      lapack_reference's inputs for that seed, solved by this library's FP64
      runtime solver (the reference solved them with LAPACKE_zhegvd), written
      as GEIG v1 with the same summary line.
  python sunny-batched-geneig_amd/geig_tool.py dump FILE
      dump_pair's report.
  python sunny-batched-geneig_amd/geig_tool.py check FILE [--form plain|precision] [--tol 1e-12]
      solves the file's pairs on the GPU and compares eigenvalues (relative
      to max(|lambda|, 1) per pair, the reference's scaling) and the
      eigenvectors up to phase with the stored ones; exit status 1 if above tol.
"""
import argparse
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _package():
    name = "sunny_batched_geneig_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def _geig():
    _package()
    import importlib

    return importlib.import_module("sunny_batched_geneig_amd.geig")


def solve_gpu(H, S, form="plain"):
    """W, U, info of the pairs H, S (B x N x N, numpy) from the GPU library."""
    import torch

    g = _package()
    b, n, _ = H.shape
    dev = torch.device("cuda:0")
    col = lambda a: torch.from_numpy(np.ascontiguousarray(a.transpose(0, 2, 1))).to(dev)   # column-major pairs
    dA, dB = col(H), col(S)
    W, U, info = g.empty_outputs(n, b, dev)
    if form == "precision":
        g.GeneigMixedSolver(n).launch_mixed(dA, dB, W, U, info)
    else:
        g.GeneigSolver(n).launch(dA, dB, W, U, info)
    torch.cuda.synchronize()
    return W.cpu().numpy(), U.cpu().numpy().transpose(0, 2, 1), info.cpu().numpy()


def rel_residual(H, S, W, U):
    """max over pairs and eigenpairs of ||H u - l S u||_inf / (||H||_F + |l| ||S||_F)
    (tests/lapack_reference.cpp:201-245)."""
    worst = 0.0
    for p in range(H.shape[0]):
        hf, sf = np.linalg.norm(H[p]), np.linalg.norm(S[p])
        R = H[p] @ U[p] - (S[p] @ U[p]) * W[p][None, :]
        den = hf + np.abs(W[p]) * sf
        worst = max(worst, float((np.abs(R).max(axis=0) / np.where(den > 0, den, 1.0)).max()))
    return worst


def compare(W, U, Wr, Ur, S):
    """Eigenvalue error scaled by max(|lambda|, 1) per pair, and the eigenvector
    error after removing each column's phase (B-inner product with the stored one)."""
    scale = np.maximum(np.abs(Wr).max(axis=1, keepdims=True), 1.0)
    eig = float((np.abs(W - Wr) / scale).max())
    ph = 0.0
    for p in range(W.shape[0]):
        c = np.einsum("ij,ik,kj->j", Ur[p].conj(), S[p], U[p])   # u_ref^H S u per column
        phase = c / np.where(np.abs(c) > 0, np.abs(c), 1.0)
        ph = max(ph, float(np.abs(U[p] - Ur[p] * phase[None, :]).max()))
    return eig, ph


def main(argv=None):
    ap = argparse.ArgumentParser(prog="geig_tool.py")
    sub = ap.add_subparsers(dest="cmd", required=True)
    mk = sub.add_parser("make")
    mk.add_argument("--n", type=int, required=True)
    mk.add_argument("--batch", type=int, required=True)
    mk.add_argument("--seed", type=int, default=0)
    mk.add_argument("--out", required=True)
    mk.add_argument("--s-mode", choices=["random", "near-identity"], default="random")
    mk.add_argument("--check-residuals", dest="check", action="store_true", default=True)
    mk.add_argument("--no-check-residuals", dest="check", action="store_false")
    dp = sub.add_parser("dump")
    dp.add_argument("path")
    ck = sub.add_parser("check")
    ck.add_argument("path")
    ck.add_argument("--form", choices=["plain", "precision"], default="plain")
    ck.add_argument("--tol", type=float, default=1e-12)
    a = ap.parse_args(argv)
    G = _geig()
    if a.cmd == "dump":
        print("\n".join(G.dump(a.path)))
        return 0
    if a.cmd == "make":
        if a.n <= 0 or a.batch <= 0:
            ap.error("--n and --batch must be positive")
        H, S = G.reference_pairs(a.n, a.batch, a.seed, a.s_mode)
        W, U, info = solve_gpu(H, S)
        if (info != 0).any():
            print("solve failed for %d pairs (first info %d)" % (int((info != 0).sum()), int(info[info != 0][0])),
                  file=sys.stderr)
            return 3
        G.write(a.out, H, S, W, U, seed=a.seed)
        line = ("GPU reference computed: N=%d, B=%d, seed=%d, s_mode=%s, output=%s, total eigenvalues_norm=%.6e"
                % (a.n, a.batch, a.seed, a.s_mode, a.out, float(np.sqrt((W * W).sum()))))
        if a.check:
            line += ", max_rel_residual=%.3e" % rel_residual(H, S, W, U)
        print(line)
        return 0
    d = G.read(a.path)
    W, U, info = solve_gpu(d["H"], d["S"], a.form)
    eig, ph = compare(W, U, d["W"], d["U"], d["S"])
    res = rel_residual(d["H"], d["S"], W, U)
    ok = eig <= a.tol and (info == 0).all()
    print("%s: N=%d B=%d form=%s  eig %.3e  phase %.3e  max_rel_residual %.3e  info nonzero %d  %s"
          % (a.path, d["n"], d["batch"], a.form, eig, ph, res, int((info != 0).sum()), "ok" if ok else "FAIL"))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
