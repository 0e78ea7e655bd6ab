"""The allocation-free batched solve (geneig_full_n54_batched_launch_noalloc)
through the C ABI: its outputs equal geneig_full_n54_batched_launch's bit for
bit on two calls in a row (the second smaller, on other data, so nothing the
first left in the shared array may leak into it), it creates no workspace for
its stream, and a batch beyond the array is refused with a clear error."""
import pytest

from test_gpu_parity import check_pairs, to_host

pytestmark = pytest.mark.gpu


def test_noalloc_matches_batched_launch(geneig, oracle, device):
    import torch

    mx = geneig.noalloc_max_batch()
    assert mx == 4485
    s = torch.cuda.Stream()          # a stream the library has never seen
    s.wait_stream(torch.cuda.current_stream())
    cases = [("spin-wave", geneig.gen_spinwave(27, 3000, seed=31)),
             ("generic", geneig.gen_generic(54, 1000, seed=32))]
    for name, (A, B) in cases:
        total = A.shape[0]
        torch.cuda.synchronize()
        W, U, info = geneig.empty_outputs(54, total, device)
        W.fill_(float("nan"))
        info.fill_(-1)
        with torch.cuda.stream(s):
            geneig.full_n54_batched_noalloc(A, B, W, U, info, stream=s)
        s.synchronize()
        W2, U2, i2 = geneig.empty_outputs(54, total, device)
        geneig.full_n54_batched(A, B, W2, U2, i2)
        torch.cuda.synchronize()
        print("%s: %d pairs, info max %d" % (name, total, int(info.abs().max())), flush=True)
        assert torch.equal(info, i2) and int(info.abs().max()) == 0, name
        assert torch.equal(W, W2) and torch.equal(U, U2), name
        Ah, Bh = to_host(A[:16]), to_host(B[:16])
        Wo, Xo, _ = oracle.zhegv(Ah, Bh)
        check_pairs(oracle, Ah, Bh, to_host(W[:16]), to_host(U[:16]), Wo, Xo, 54, "noalloc " + name)
    # nothing was allocated for the stream: no workspace to release
    assert geneig.workspace_release(stream=s) == 0
