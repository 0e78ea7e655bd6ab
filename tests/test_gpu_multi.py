"""GPU tests of mixed-size batches (geneig_multi_*): every size bin solved by
one graph launch (one branch per size, captured once and replayed) must give
outputs bit-identical to one solve launch per size, and both must meet the
FP64 bar against the oracle.  Graph replays are separated by eager work and
synchronizations and every replay is checked."""
import numpy as np
import pytest

from test_gpu_parity import check_pairs, to_dev, to_host

pytestmark = pytest.mark.gpu

SIZES = [8, 16, 24, 32, 48, 54, 64, 96]


def make_bins(geneig, oracle, device, counts, seed):
    bins, host = [], []
    for k, (n, c) in enumerate(zip(SIZES, counts)):
        if c == 0:
            bins.append(None)
            host.append(None)
            continue
        A, B = oracle.gen_generic(n, c, seed=seed + k)
        W, U, info = geneig.empty_outputs(n, c, device)
        bins.append((to_dev(A, device), to_dev(B, device), W, U, info))
        host.append((A, B))
    return bins, host


def snapshot(bins):
    return [None if b is None else (to_host(b[2]).copy(), to_host(b[3]).copy(), to_host(b[4]).copy())
            for b in bins]


def poison(bins):
    for b in bins:
        if b is not None:
            b[2].fill_(float("nan"))
            b[3].fill_(float("nan"))
            b[4].fill_(-99)


def test_multi_graph_replays_match_serial_and_oracle(geneig, oracle, device):
    import torch

    counts = [23, 17, 11, 9, 6, 5, 4, 3]
    m = geneig.GeneigMulti(SIZES, tuning_mode=geneig.TUNING_SUGGESTED)
    bins, host = make_bins(geneig, oracle, device, counts, seed=100)
    m.launch_serial(bins)
    torch.cuda.synchronize()
    ref = snapshot(bins)
    for k, (n, c) in enumerate(zip(SIZES, counts)):
        A, B = host[k]
        W, U, info = ref[k]
        assert (info == 0).all(), n
        Wo, Xo, _ = oracle.zhegv(A, B)
        check_pairs(oracle, A, B, W, U, Wo, Xo, n, "serial n=%d" % n)
    scratch = torch.randn(1 << 20, device=device)
    for rep in range(3):
        poison(bins)
        torch.cuda.synchronize()
        m.launch(bins)
        scratch.mul_(1.0001).add_(0.5)        # eager work between replays
        torch.cuda.synchronize()
        got = snapshot(bins)
        for k in range(len(SIZES)):
            for a, b in zip(got[k], ref[k]):
                assert a.tobytes() == b.tobytes(), "replay %d, n=%d" % (rep, SIZES[k])
    assert m.captures() == 1 and m.replays() == 3


def test_multi_new_buffers_capture_again(geneig, oracle, device):
    import torch

    m = geneig.GeneigMulti(SIZES, tuning_mode=geneig.TUNING_SUGGESTED)
    for seed, counts in ((7, [5, 4, 3, 2, 2, 2, 1, 1]), (9, [3, 0, 6, 1, 0, 4, 2, 1])):
        bins, host = make_bins(geneig, oracle, device, counts, seed=seed)
        m.launch(bins)
        m.launch(bins)
        torch.cuda.synchronize()
        for k, (n, c) in enumerate(zip(SIZES, counts)):
            if c == 0:
                continue
            A, B = host[k]
            Wo, Xo, _ = oracle.zhegv(A, B)
            check_pairs(oracle, A, B, to_host(bins[k][2]), to_host(bins[k][3]), Wo, Xo, n,
                        "graph n=%d seed %d" % (n, seed))
    assert m.captures() == 2 and m.replays() == 4


def test_multi_rejects_bad_bins(geneig, device):
    m = geneig.GeneigMulti([8, 16], tuning_mode=geneig.TUNING_SUGGESTED)
    with pytest.raises(geneig.GeneigError):
        m.launch([None])
