"""Python host binding of libgeneig.so (the MI355X batched generalized eigensolver).

Mirrors the reference's host interface for the solve path:
  * the static n = 54 extern "C" launchers of src/geneig_fixed.cu
    (``full_n54_batched``, ``chol_n54``, ``reduce_n54``, ``reduce_eig_n54``,
    ``full_n54``, ``finish_n54``), and
  * ``GeneigSolver``, the runtime-n solver with the constructor arguments,
    ``launch``/``launch_chunked`` and introspection of
    ``NvrtcGeneigSolver`` (src/nvrtc/nvrtc_solver.hpp:81-217).
Device buffers are torch tensors (complex128 / float64 / int32 on a ROCm
device); they are plumbing only -- every solve runs the library's HIP kernels
through the C ABI, and a missing or unloadable library raises at once.
Errors surface as ``GeneigError`` carrying the library's message (the
reference threw std::runtime_error).
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# GENEIG_LIBRARY names another build of the library (the host-sanitized one
# of ``make SANITIZE=1``); by default the in-tree libgeneig.so is loaded.
LIB_PATH = os.environ.get("GENEIG_LIBRARY") or os.path.join(HERE, "libgeneig.so")

TUNING_SUGGESTED, TUNING_AUTOTUNE = 0, 1
CACHE_AUTO, CACHE_NODISK, CACHE_NOCACHE = 0, 1, 2


class GeneigError(RuntimeError):
    pass


_lib = None


def build(jobs=4):
    """Compile libgeneig.so in-tree (hipcc, gfx950)."""
    import subprocess

    subprocess.check_call(["make", "-s", "-C", HERE, "-j%d" % jobs])


def lib():
    """Load libgeneig.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GeneigError("libgeneig.so is not built (run __graft_entry__.build() or make -C %s)" % HERE)
    L = C.CDLL(LIB_PATH)
    P, I, U64, D = C.c_void_p, C.c_int, C.c_uint64, C.c_double
    sig = {
        "chol_n54_launch": (I, [P, P, P, P]),
        "geneig_reduce_n54_launch": (I, [P, P, P, P, P]),
        "geneig_reduce_eig_n54_launch": (I, [P, P, P, P, P, P]),
        "geneig_full_n54_launch": (I, [P, P, P, P, P, P]),
        "geneig_full_n54_batched_launch": (I, [P, P, P, P, P, I, P]),
        "geneig_full_n54_batched_inplace_launch": (I, [P, P, P, P, I, P]),
        "geneig_solver_launch_inplace": (I, [P, P, P, P, P, I, P]),
        "geneig_finish_n54_launch": (I, [P, P, P, P, P, I, P]),
        "geneig_mixed_n54_batched_launch": (I, [P, P, P, P, P, I, I, D, P, P, P]),
        "geneig_mixed_eta_n54_batched_launch": (I, [P, P, P, P, I, I, D, P, P, P]),
        "geneig_mixed_eta_paired_n54_batched_launch": (I, [P, P, P, P, I, I, D, P, P, P]),
        "geneig_eta_paired_n54_batched_launch": (I, [P, P, P, P, I, P]),
        "geneig_mixed_default_max_iter": (I, []),
        "geneig_mixed_default_tol": (D, []),
        "geneig_mixed_solver_create": (I, [P, I, I, I]),
        "geneig_mixed_solver_create_bdx": (I, [P, I, I, I, I]),
        "geneig_spinwave_params_launch": (I, [U64, I, I, I, U64, I, D, P, P]),
        "geneig_learn_num_features": (I, []),
        "geneig_learn_hidden": (I, []),
        "geneig_learn_num_outputs": (I, []),
        "geneig_learn_features": (I, [P, P, I, P]),
        "geneig_learn_gemm": (I, [I, I, I, I, I, P, I, P, I, P, I, P, I, P]),
        "geneig_learn_tanh_backward": (I, [P, P, C.c_longlong, P]),
        "geneig_learn_colsum": (I, [P, I, I, P, P]),
        "geneig_learn_adam": (I, [P, P, P, P, C.c_longlong, C.c_float, C.c_float, C.c_float, C.c_float, I, P]),
        "geneig_learn_mlp_forward": (I, [P, P, I, P, P, P]),
        "geneig_learn_loss": (I, [P, P, P, I, C.c_float, P, P, P]),
        "geneig_learn_reduce_n54": (I, [P, P, P, P, P, I, P]),
        "geneig_learned_n54_batched_launch": (I, [P, P, P, P, P, I, P, P, I, C.c_float, I, D, P, P, P, P]),
        "geneig_mixed_solver_launch": (I, [P, P, P, P, P, P, I, I, D, P, P, P]),
        "geneig_solver_create": (I, [P, I, I, I, I, I, I]),
        "geneig_solver_destroy": (None, [P]),
        "geneig_solver_launch": (I, [P, P, P, P, P, P, I, P]),
        "geneig_solver_launch_chunked": (I, [P, P, P, P, P, P, I, I, I, P]),
        "geneig_solver_launch_chunked_host": (I, [P, P, P, P, P, P, I, I, I, P]),
        "geneig_solver_matrix_size": (I, [P]),
        "geneig_solver_batches_per_block": (I, [P]),
        "geneig_solver_block_dim": (None, [P, P, P, P]),
        "geneig_solver_shared_mem_bytes": (C.c_uint, [P]),
        "geneig_solver_autotuned_block_dim_x": (I, [P]),
        "geneig_solver_was_autotuned": (I, [P]),
        "geneig_solver_autotune_sweep_ran": (I, [P]),
        "geneig_solver_cache_layer_used": (C.c_char_p, [P]),
        "geneig_solver_acquire_module_ms": (D, [P]),
        "geneig_solver_variant": (I, [P]),
        "geneig_solver_k1_chunk_rows": (I, [P]),
        "geneig_solver_reserve": (I, [P, I, P]),
        "geneig_solver_workspace_bytes": (C.c_size_t, [P, I]),
        "geneig_n54_workspace_bytes": (C.c_size_t, [I, I]),
        "geneig_n54_reserve": (I, [I, I, P]),
        "geneig_workspace_release": (I, [P, P]),
        "geneig_full_n54_batched_launch_noalloc": (I, [P, P, P, P, P, I, P]),
        "geneig_full_n54_batched_noalloc_max_batch": (I, []),
        "geneig_multi_create": (I, [P, I, P, I, I, I]),
        "geneig_multi_destroy": (None, [P]),
        "geneig_multi_launch": (I, [P, P, P, P, P, P, P, P]),
        "geneig_multi_launch_serial": (I, [P, P, P, P, P, P, P, P]),
        "geneig_multi_num_sizes": (I, [P]),
        "geneig_multi_solver": (P, [P, I]),
        "geneig_multi_captures": (C.c_longlong, [P]),
        "geneig_multi_replays": (C.c_longlong, [P]),
        "geneig_cache_dir": (C.c_size_t, [C.c_char_p, C.c_size_t]),
        "geneig_cache_path": (C.c_size_t, [I, I, C.c_char_p, I, I, C.c_char_p, C.c_size_t]),
        "geneig_rtc_compile": (I, [I, I, I, C.c_char_p, P, C.c_size_t, P]),
        "geneig_rtc_compile_rows": (I, [I, I, I, I, C.c_char_p, P, C.c_size_t, P]),
        "geneig_rtc_compile_bdx": (I, [I, I, I, I, I, C.c_char_p, P, C.c_size_t, P]),
        "geneig_source_hash": (U64, []),
        "geneig_fnv1a64": (U64, [P, C.c_size_t, U64]),
        "geneig_make_streams": (I, [I, P]),
        "geneig_destroy_streams": (None, [I, P]),
        "geneig_last_error": (C.c_char_p, []),
        "geneig_gen_generic_launch": (I, [I, I, U64, U64, P, P, P]),
        "geneig_gen_spinwave_h0_launch": (I, [I, I, U64, I, I, I, U64, I, D, P, P, P]),
        "geneig_spinwave_shift_launch": (I, [I, I, U64, U64, P, P, P]),
        "geneig_fill_identity_launch": (I, [I, I, P, P]),
        "geneig_spinwave_model_launch": (I, [I, I, U64, D, P, P, P, P]),
        "geneig_spinwave_n54_launch": (I, [I, P, P, P, P, P, P, P, P, I, P]),
        "geneig_spinwave_n54_build": (I, [I, P, P, P, P, P, P, I, P]),
    }
    for name in ("chol_n54", "geneig_reduce_n54", "geneig_reduce_eig_n54", "geneig_full_n54",
                 "geneig_full_n54_batched", "geneig_finish_n54"):
        sig[name + "_block_dim_x"] = (C.c_uint, [])
        sig[name + "_shared_mem_bytes"] = (C.c_uint, [])
    for extra in ("geneig_reduce_eig_n54_chol_smem_bytes", "geneig_reduce_eig_n54_heev_smem_bytes",
                  "geneig_reduce_eig_n54_heev_workspace"):
        sig[extra] = (C.c_uint, [])
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    return lib().geneig_last_error().decode(errors="replace")


def _check_hip(rc, what):
    if rc != 0:
        raise GeneigError("%s failed: hipError %d" % (what, rc))


def _check(rc, what):
    if rc != 0:
        raise GeneigError("%s failed (code %d): %s" % (what, rc, last_error()))


# ---- torch plumbing ------------------------------------------------------------
def _torch():
    import torch

    return torch


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def _require(t, dtype, name):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise GeneigError("%s must be a device tensor" % name)
    if t.dtype != dtype:
        raise GeneigError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise GeneigError("%s must be contiguous" % name)


def _require_host(t, dtype, name):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or t.is_cuda:
        raise GeneigError("%s must be a host tensor" % name)
    if t.dtype != dtype:
        raise GeneigError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise GeneigError("%s must be contiguous" % name)


def _check_batch(n, batch, A, B, W, U, info, host=False):
    torch = _torch()
    req = _require_host if host else _require
    for t, nm in ((A, "A"), (B, "B"), (U, "U")):
        req(t, torch.complex128, nm)
        if t.numel() < batch * n * n:
            raise GeneigError("%s holds %d elements, batch*n*n = %d" % (nm, t.numel(), batch * n * n))
    if W is not None:
        req(W, torch.float64, "W")
        if W.numel() < batch * n:
            raise GeneigError("W holds %d elements, batch*n = %d" % (W.numel(), batch * n))
    req(info, torch.int32, "info")
    if info.numel() < batch:
        raise GeneigError("info holds %d elements, batch = %d" % (info.numel(), batch))


def empty_outputs(n, batch, device="cuda"):
    torch = _torch()
    W = torch.empty((batch, n), dtype=torch.float64, device=device)
    U = torch.empty((batch, n, n), dtype=torch.complex128, device=device)
    info = torch.empty((batch,), dtype=torch.int32, device=device)
    return W, U, info


# ---- static n = 54 entry points (src/geneig_fixed.cu) ---------------------------
def full_n54_batched(H, S, W, U, info, batch=None, stream=None):
    """geneig_full_n54_batched_launch (geneig_fixed.cu:474-492)."""
    batch = H.shape[0] if batch is None else batch
    _check_batch(54, batch, H, S, W, U, info)
    _check_hip(lib().geneig_full_n54_batched_launch(_ptr(H), _ptr(S), _ptr(W), _ptr(U), _ptr(info),
                                                    int(batch), _stream(stream)),
               "geneig_full_n54_batched_launch")


def full_n54_batched_inplace(H, S, W, info, batch=None, stream=None):
    """geneig_full_n54_batched_inplace_launch: zhegvd storage -- H receives the
    eigenvectors, the lower triangle of S the Cholesky factor."""
    batch = H.shape[0] if batch is None else batch
    _check_batch(54, batch, H, S, W, H, info)
    _check_hip(lib().geneig_full_n54_batched_inplace_launch(_ptr(H), _ptr(S), _ptr(W), _ptr(info), int(batch),
                                                            _stream(stream)), "geneig_full_n54_batched_inplace_launch")


def mixed_n54_batched(H, S, W, U, info, batch=None, max_iter=0, tol=0.0, iters=None, stream=None,
                      report_fallback=True):
    """Precision form of the n = 54 batched solve (geneig_mixed_n54_batched_launch):
    complex64 bulk, FP64 Cholesky and FP64 refinement, FP64 re-solve of the
    pairs the refinement hands back.  Same outputs as ``full_n54_batched``.
    ``iters`` (optional int32 device tensor of ``batch`` entries) receives the
    refinement steps per pair.  Returns the number of pairs solved again in
    FP64, for which the launch waits on the device; with
    ``report_fallback=False`` it returns None and does not wait (the launch
    is then asynchronous and capturable)."""
    batch = H.shape[0] if batch is None else batch
    _check_batch(54, batch, H, S, W, U, info)
    if iters is not None:
        _require(iters, _torch().int32, "iters")
        if iters.numel() < batch:
            raise GeneigError("iters holds %d entries, batch is %d" % (iters.numel(), batch))
    fb = C.c_int(0)
    _check_hip(lib().geneig_mixed_n54_batched_launch(_ptr(H), _ptr(S), _ptr(W), _ptr(U), _ptr(info), int(batch),
                                                     int(max_iter), float(tol),
                                                     _ptr(iters) if iters is not None else None,
                                                     C.byref(fb) if report_fallback else None, _stream(stream)),
               "geneig_mixed_n54_batched_launch")
    return fb.value if report_fallback else None


def mixed_eta_n54_batched(S, W, U, info, batch=None, max_iter=0, tol=0.0, iters=None, stream=None,
                          report_fallback=True):
    """The precision form for the spin-wave pencil (eta, S), eta = diag(I_27,
    -I_27), which is not passed (geneig_mixed_eta_n54_batched_launch): the
    arguments and outputs of ``mixed_n54_batched`` without H; outputs equal
    ``mixed_n54_batched(eta, S, ...)``."""
    batch = S.shape[0] if batch is None else batch
    _check_batch(54, batch, S, S, W, U, info)
    if iters is not None:
        _require(iters, _torch().int32, "iters")
        if iters.numel() < batch:
            raise GeneigError("iters holds %d entries, batch is %d" % (iters.numel(), batch))
    fb = C.c_int(0)
    _check_hip(lib().geneig_mixed_eta_n54_batched_launch(_ptr(S), _ptr(W), _ptr(U), _ptr(info), int(batch),
                                                         int(max_iter), float(tol),
                                                         _ptr(iters) if iters is not None else None,
                                                         C.byref(fb) if report_fallback else None, _stream(stream)),
               "geneig_mixed_eta_n54_batched_launch")
    return fb.value if report_fallback else None


def mixed_eta_paired_n54_batched(S, W, U, info, batch=None, max_iter=0, tol=0.0, iters=None, stream=None,
                                 report_fallback=True):
    """The precision form for the spin-wave pencil (eta, S) with S = [[P, Q],
    [conj Q, conj P]] (geneig_mixed_eta_paired_n54_batched_launch): the
    refinement works on the 27 positive eigenpairs and writes their partners
    (-lambda, Sigma conj(x)); arguments and outputs as ``mixed_eta_n54_batched``."""
    batch = S.shape[0] if batch is None else batch
    _check_batch(54, batch, S, S, W, U, info)
    if iters is not None:
        _require(iters, _torch().int32, "iters")
        if iters.numel() < batch:
            raise GeneigError("iters holds %d entries, batch is %d" % (iters.numel(), batch))
    fb = C.c_int(0)
    _check_hip(lib().geneig_mixed_eta_paired_n54_batched_launch(_ptr(S), _ptr(W), _ptr(U), _ptr(info), int(batch),
                                                                int(max_iter), float(tol),
                                                                _ptr(iters) if iters is not None else None,
                                                                C.byref(fb) if report_fallback else None,
                                                                _stream(stream)),
               "geneig_mixed_eta_paired_n54_batched_launch")
    return fb.value if report_fallback else None


def eta_paired_n54_batched(S, W, U, info, batch=None, stream=None):
    """The spin-wave pencil (eta, S) with S = [[P, Q], [conj Q, conj P]] in
    real FP64 arithmetic (geneig_eta_paired_n54_batched_launch): outputs as
    ``full_n54_batched(eta, S, ...)`` to rounding, with the pairing
    W[53 - i] = -W[i], X(:, i) = Sigma conj X(:, 53 - i) exact."""
    batch = S.shape[0] if batch is None else batch
    _check_batch(54, batch, S, S, W, U, info)
    _check_hip(lib().geneig_eta_paired_n54_batched_launch(_ptr(S), _ptr(W), _ptr(U), _ptr(info), int(batch),
                                                          _stream(stream)),
               "geneig_eta_paired_n54_batched_launch")


def full_n54_batched_noalloc(H, S, W, U, info, batch=None, stream=None):
    """geneig_full_n54_batched_launch_noalloc: the batched FP64 solve with its workspace
    in the library's fixed device array (no device allocation; at most
    ``noalloc_max_batch()`` pairs; calls overlapping on one GPU race)."""
    batch = H.shape[0] if batch is None else batch
    _check_batch(54, batch, H, S, W, U, info)
    rc = lib().geneig_full_n54_batched_launch_noalloc(_ptr(H), _ptr(S), _ptr(W), _ptr(U), _ptr(info), int(batch),
                                                      _stream(stream))
    if rc != 0:
        raise GeneigError("geneig_full_n54_batched_launch_noalloc failed: hipError %d: %s" % (rc, last_error()))


def noalloc_max_batch():
    return int(lib().geneig_full_n54_batched_noalloc_max_batch())


def n54_workspace_bytes(batch, form=0):
    """Workspace bytes of a batched n = 54 launch (form 0 plain FP64, 1 precision)."""
    return int(lib().geneig_n54_workspace_bytes(int(batch), int(form)))


def n54_reserve(batch, form=0, stream=None):
    """Allocate the workspace of a batched n = 54 launch of ``batch`` pairs on ``stream``
    ahead of time (the launch then allocates nothing and can be captured)."""
    _check_hip(lib().geneig_n54_reserve(int(batch), int(form), _stream(stream)), "geneig_n54_reserve")


def workspace_release(stream=None):
    """Free the workspace the library holds for ``stream`` on the current device; returns bytes freed."""
    freed = C.c_size_t(0)
    _check_hip(lib().geneig_workspace_release(_stream(stream), C.byref(freed)), "geneig_workspace_release")
    return freed.value


def full_n54(H, S, W, U, info, stream=None):
    """geneig_full_n54_launch (geneig_fixed.cu:385-400), one pair."""
    _check_batch(54, 1, H, S, W, U, info)
    _check_hip(lib().geneig_full_n54_launch(_ptr(H), _ptr(S), _ptr(W), _ptr(U), _ptr(info), _stream(stream)),
               "geneig_full_n54_launch")


def chol_n54(S, L, info, stream=None):
    """chol_n54_launch (geneig_fixed.cu:134-147), one pair."""
    _check_batch(54, 1, S, S, None, L, info)
    _check_hip(lib().chol_n54_launch(_ptr(S), _ptr(L), _ptr(info), _stream(stream)), "chol_n54_launch")


def reduce_n54(H, S, M, info, stream=None):
    """geneig_reduce_n54_launch (geneig_fixed.cu:208-222), one pair."""
    _check_batch(54, 1, H, S, None, M, info)
    _check_hip(lib().geneig_reduce_n54_launch(_ptr(H), _ptr(S), _ptr(M), _ptr(info), _stream(stream)),
               "geneig_reduce_n54_launch")


def reduce_eig_n54(H, S, W, V, info, stream=None):
    """geneig_reduce_eig_n54_launch (geneig_fixed.cu:302-317), one pair."""
    _check_batch(54, 1, H, S, W, V, info)
    _check_hip(lib().geneig_reduce_eig_n54_launch(_ptr(H), _ptr(S), _ptr(W), _ptr(V), _ptr(info),
                                                  _stream(stream)), "geneig_reduce_eig_n54_launch")


def finish_n54(L, M, W, U, info, batch=None, stream=None):
    """geneig_finish_n54_launch (geneig_fixed.cu:569-587)."""
    batch = L.shape[0] if batch is None else batch
    _check_batch(54, batch, L, M, W, U, info)
    _check_hip(lib().geneig_finish_n54_launch(_ptr(L), _ptr(M), _ptr(W), _ptr(U), _ptr(info), int(batch),
                                              _stream(stream)), "geneig_finish_n54_launch")


# ---- runtime-n solver (src/nvrtc/nvrtc_solver.hpp:81-217) -------------------------
class GeneigSolver:
    """Per-n specialized solver: GeneigSolver(n, device_id, bpb, force_block_dim_x, tuning_mode, cache_mode)."""

    def __init__(self, n, device_id=0, batches_per_block_request=0, force_block_dim_x=0,
                 tuning_mode=TUNING_SUGGESTED, cache_mode=CACHE_AUTO):
        self._h = C.c_void_p()
        rc = lib().geneig_solver_create(C.byref(self._h), int(n), int(device_id), int(batches_per_block_request),
                                        int(force_block_dim_x), int(tuning_mode), int(cache_mode))
        _check(rc, "GeneigSolver(n=%d)" % n)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.geneig_solver_destroy(h)
            self._h = None

    def launch(self, H, S, W, U, info, batch=None, stream=None):
        batch = H.shape[0] if batch is None else batch
        _check_batch(self.matrix_size(), batch, H, S, W, U, info)
        _check(lib().geneig_solver_launch(self._h, _ptr(H), _ptr(S), _ptr(W), _ptr(U), _ptr(info),
                                          int(batch), _stream(stream)), "GeneigSolver.launch")

    def launch_inplace(self, H, S, W, info, batch=None, stream=None):
        """zhegvd storage: H receives the eigenvectors, the lower triangle of S the factor."""
        batch = H.shape[0] if batch is None else batch
        _check_batch(self.matrix_size(), batch, H, S, W, H, info)
        _check(lib().geneig_solver_launch_inplace(self._h, _ptr(H), _ptr(S), _ptr(W), _ptr(info), int(batch),
                                                  _stream(stream)), "GeneigSolver.launch_inplace")

    def launch_chunked(self, H, S, W, U, info, total, chunk_size, streams):
        _check_batch(self.matrix_size(), total, H, S, W, U, info)
        arr = (C.c_void_p * len(streams))(*[s.cuda_stream for s in streams])
        _check(lib().geneig_solver_launch_chunked(self._h, _ptr(H), _ptr(S), _ptr(W), _ptr(U), _ptr(info),
                                                  int(total), int(chunk_size), len(streams),
                                                  C.cast(arr, C.c_void_p)), "GeneigSolver.launch_chunked")

    def launch_chunked_host(self, H, S, W, U, info, total, chunk_size, streams):
        """launch_chunked over host tensors (pinned for copy overlap): per chunk, on its
        stream, H and S are copied in, solved and W, U, info copied out.  Outputs are
        complete once every stream is synchronized."""
        _check_batch(self.matrix_size(), total, H, S, W, U, info, host=True)
        arr = (C.c_void_p * len(streams))(*[s.cuda_stream for s in streams])
        _check(lib().geneig_solver_launch_chunked_host(self._h, _ptr(H), _ptr(S), _ptr(W), _ptr(U), _ptr(info),
                                                       int(total), int(chunk_size), len(streams),
                                                       C.cast(arr, C.c_void_p)), "GeneigSolver.launch_chunked_host")

    def matrix_size(self):
        return lib().geneig_solver_matrix_size(self._h)

    def workspace_bytes(self, batch):
        return int(lib().geneig_solver_workspace_bytes(self._h, int(batch)))

    def batches_per_block(self):
        return lib().geneig_solver_batches_per_block(self._h)

    def block_dim(self):
        x, y, z = C.c_uint(), C.c_uint(), C.c_uint()
        lib().geneig_solver_block_dim(self._h, C.byref(x), C.byref(y), C.byref(z))
        return (x.value, y.value, z.value)

    def shared_mem_bytes(self):
        return lib().geneig_solver_shared_mem_bytes(self._h)

    def autotuned_block_dim_x(self):
        return lib().geneig_solver_autotuned_block_dim_x(self._h)

    def was_autotuned(self):
        return bool(lib().geneig_solver_was_autotuned(self._h))

    def autotune_sweep_ran(self):
        return bool(lib().geneig_solver_autotune_sweep_ran(self._h))

    def cache_layer_used(self):
        return lib().geneig_solver_cache_layer_used(self._h).decode()

    def acquire_module_ms(self):
        return lib().geneig_solver_acquire_module_ms(self._h)

    def variant(self):
        return lib().geneig_solver_variant(self._h)

    def k1_chunk_rows(self):
        """Rows per chunk of K1's register walks (the autotuned tile; 0 above n = 64)."""
        return lib().geneig_solver_k1_chunk_rows(self._h)

    def reserve(self, batch, stream=None):
        """Pre-allocate the workspace a launch of `batch` pairs on `stream` uses."""
        _check(lib().geneig_solver_reserve(self._h, int(batch), _stream(stream)), "GeneigSolver.reserve")

    @staticmethod
    def cache_dir():
        buf = C.create_string_buffer(4096)
        lib().geneig_cache_dir(buf, 4096)
        return buf.value.decode()

    @staticmethod
    def cache_path(n, bpb, arch, block_dim_x=64, variant=0):
        buf = C.create_string_buffer(4096)
        lib().geneig_cache_path(int(n), int(bpb), arch.encode(), int(block_dim_x), int(variant), buf, 4096)
        return buf.value.decode()


class GeneigMixedSolver(GeneigSolver):
    """Precision form for any n (geneig_mixed_solver_*): complex64 bulk, FP64
    Cholesky and refinement, FP64 re-solve of the pairs the refinement hands
    back.  ``launch`` (inherited) runs it with the default controls;
    ``launch_mixed`` exposes them and returns the number of re-solved pairs."""

    def __init__(self, n, device_id=0, cache_mode=CACHE_AUTO, block_dim_x=0):
        """block_dim_x: threads per pair of the Cholesky, K1 and K3 kernels (64
        or 256; 0: 64 up to n = 64, 256 above)."""
        self._h = C.c_void_p()
        _check(lib().geneig_mixed_solver_create_bdx(C.byref(self._h), int(n), int(device_id), int(block_dim_x),
                                                    int(cache_mode)),
               "GeneigMixedSolver(n=%d, block_dim_x=%d)" % (n, block_dim_x))

    def launch_mixed(self, H, S, W, U, info, batch=None, max_iter=0, tol=0.0, iters=None, stream=None,
                     report_fallback=True):
        """Returns the number of pairs re-solved in FP64, which waits for the
        launch; with report_fallback=False nothing waits and None is returned."""
        batch = H.shape[0] if batch is None else batch
        _check_batch(self.matrix_size(), batch, H, S, W, U, info)
        if iters is not None:
            _require(iters, _torch().int32, "iters")
            if iters.numel() < batch:
                raise GeneigError("iters holds %d entries, batch is %d" % (iters.numel(), batch))
        fb = C.c_int(0)
        _check(lib().geneig_mixed_solver_launch(self._h, _ptr(H), _ptr(S), _ptr(W), _ptr(U), _ptr(info), int(batch),
                                                int(max_iter), float(tol), _ptr(iters),
                                                C.byref(fb) if report_fallback else None, _stream(stream)),
               "GeneigMixedSolver.launch_mixed")
        return fb.value if report_fallback else None


class GeneigMulti:
    """Mixed-size batches (geneig_multi_*): one solver per size; ``launch`` solves
    every bin in one graph launch (one branch per size, captured once and
    replayed), ``launch_serial`` one solve launch per size."""

    def __init__(self, sizes, device_id=0, tuning_mode=TUNING_AUTOTUNE, cache_mode=CACHE_AUTO):
        self.sizes = [int(n) for n in sizes]
        self._h = C.c_void_p()
        arr = (C.c_int * len(self.sizes))(*self.sizes)
        _check(lib().geneig_multi_create(C.byref(self._h), len(self.sizes), C.cast(arr, C.c_void_p),
                                         int(device_id), int(tuning_mode), int(cache_mode)),
               "GeneigMulti(%s)" % self.sizes)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.geneig_multi_destroy(h)
            self._h = None

    def _args(self, bins):
        """bins: one (A, B, W, U, info) tuple of device tensors per size (None or
        an empty A for a size with no pairs)."""
        K = len(self.sizes)
        if len(bins) != K:
            raise GeneigError("need %d bins, got %d" % (K, len(bins)))
        counts = (C.c_int * K)()
        ptrs = [(C.c_void_p * K)() for _ in range(5)]
        for k, b in enumerate(bins):
            if b is None or b[0].shape[0] == 0:
                continue
            A, B, W, U, info = b
            _check_batch(self.sizes[k], A.shape[0], A, B, W, U, info)
            counts[k] = A.shape[0]
            for j, t in enumerate((A, B, W, U, info)):
                ptrs[j][k] = t.data_ptr()
        return [C.cast(counts, C.c_void_p)] + [C.cast(p, C.c_void_p) for p in ptrs]

    def launch(self, bins, stream=None):
        _check(lib().geneig_multi_launch(self._h, *self._args(bins), _stream(stream)), "GeneigMulti.launch")

    def launch_serial(self, bins, stream=None):
        _check(lib().geneig_multi_launch_serial(self._h, *self._args(bins), _stream(stream)),
               "GeneigMulti.launch_serial")

    def batches_per_block(self, k):
        return lib().geneig_solver_batches_per_block(lib().geneig_multi_solver(self._h, int(k)))

    def captures(self):
        return lib().geneig_multi_captures(self._h)

    def replays(self):
        return lib().geneig_multi_replays(self._h)

def rtc_compile(n, bpb=1, variant=0, arch="gfx950", k1_rows=0, block_dim_x=64):
    """Device-free hipRTC compile of the per-n kernel; returns the code object
    bytes.  k1_rows: K1's row-chunk width (4, 8, 16; 0 = the build's default);
    block_dim_x: threads per pair of K1 / K3 (64, or 256 with bpb = 1)."""
    size = C.c_size_t(0)
    args = (int(n), int(bpb), int(variant), int(k1_rows), int(block_dim_x), arch.encode())
    _check(lib().geneig_rtc_compile_bdx(*args, None, 0, C.byref(size)), "rtc_compile(n=%d)" % n)
    buf = C.create_string_buffer(size.value)
    _check(lib().geneig_rtc_compile_bdx(*args, buf, size.value, C.byref(size)), "rtc_compile(n=%d)" % n)
    return buf.raw[: size.value]


def make_streams(n):
    torch = _torch()
    return [torch.cuda.Stream() for _ in range(n)]


# ---- device generators ------------------------------------------------------------
def gen_generic(n, batch, seed, first_pair=0, device="cuda", stream=None):
    """A = (M + M^H)/(2 sqrt n), B = G G^H / n + I on the device."""
    torch = _torch()
    A = torch.empty((batch, n, n), dtype=torch.complex128, device=device)
    B = torch.empty_like(A)
    _check_hip(lib().geneig_gen_generic_launch(int(n), int(batch), int(seed), int(first_pair), _ptr(A), _ptr(B),
                                               _stream(stream)), "geneig_gen_generic_launch")
    return A, B


def spinwave_params(batch, seed, grid=(64, 32, 32), first_q=0, jscale=1.0, device="cuda", stream=None):
    """Per-pair (q0, q1, q2, s, t) of the spin-wave family as a (batch, 5) float32
    tensor: grid point, coupling scale (per q when jscale < 0), lambda_min target."""
    torch = _torch()
    out = torch.empty((batch, 5), dtype=torch.float32, device=device)
    _check_hip(lib().geneig_spinwave_params_launch(int(seed), grid[0], grid[1], grid[2], int(first_q), int(batch),
                                                   float(jscale), _ptr(out), _stream(stream)),
               "geneig_spinwave_params_launch")
    return out


def gen_spinwave(nsub, batch, seed, grid=(64, 32, 32), first_q=0, nbonds=12, jscale=1.0, device="cuda",
                 chunk=8192, stream=None):
    """Spin-wave pairs (eta, H(q)) with lambda_min(H) = 10^U(-3,-1) per q.

    lambda_min(H0) is taken from this library's own solve of (H0, I), chunked
    to bound the scratch memory; the shift kernel then sets lambda_min.
    """
    torch = _torch()
    n = 2 * nsub
    A = torch.empty((batch, n, n), dtype=torch.complex128, device=device)
    B = torch.empty_like(A)
    s = _stream(stream)
    _check_hip(lib().geneig_gen_spinwave_h0_launch(int(nsub), int(nbonds), int(seed), grid[0], grid[1], grid[2],
                                                   int(first_q), int(batch), float(jscale), _ptr(A), _ptr(B), s),
               "geneig_gen_spinwave_h0_launch")
    solver = None if n == 54 else GeneigSolver(n, device_id=torch.device(device).index or 0)
    c = min(chunk, batch)
    Hc = torch.empty((c, n, n), dtype=torch.complex128, device=device)
    Ic = torch.empty_like(Hc)
    Wc, Uc, infoc = empty_outputs(n, c, device)
    _check_hip(lib().geneig_fill_identity_launch(n, c, _ptr(Ic), s), "geneig_fill_identity_launch")
    worst = torch.zeros((), dtype=torch.int32, device=device)
    for off in range(0, batch, c):
        m = min(c, batch - off)
        Hc[:m].copy_(B[off:off + m])
        if solver is None:
            full_n54_batched(Hc, Ic, Wc, Uc, infoc, batch=m, stream=stream)
        else:
            solver.launch(Hc, Ic, Wc, Uc, infoc, batch=m, stream=stream)
        _check_hip(lib().geneig_spinwave_shift_launch(n, m, int(seed), int(first_q) + off, _ptr(Wc),
                                                      C.c_void_p(B[off].data_ptr()), s),
                   "geneig_spinwave_shift_launch")
        worst = torch.maximum(worst, infoc[:m].abs().max())
    if int(worst.item()) != 0:
        raise GeneigError("spin-wave generator: lambda_min solve reported info != 0")
    del Hc, Ic, Wc, Uc, infoc
    return A, B


def spinwave_model(nsub, seed, nbonds=12, jscale=1.0, device="cuda", stream=None):
    """The spin-wave family's bond tables (shift, r, J) on the device, as
    geneig_spinwave_n54_launch takes them (the draws of gen_spinwave)."""
    torch = _torch()
    shift = torch.empty(nbonds, dtype=torch.int32, device=device)
    r = torch.empty((nbonds, 3), dtype=torch.int32, device=device)
    J = torch.empty((nbonds, nsub), dtype=torch.float64, device=device)
    _check_hip(lib().geneig_spinwave_model_launch(int(nsub), int(nbonds), int(seed), float(jscale), _ptr(shift),
                                                  _ptr(r), _ptr(J), _stream(stream)), "geneig_spinwave_model_launch")
    return shift, r, J


def spinwave_q(batch, grid=(64, 32, 32), first_q=0, device="cuda"):
    """q of grid points first_q .. first_q + batch - 1, index (i g1 + j) g2 + k,
    q = (i/g0, j/g1, k/g2) (exact in float64), as a (batch, 3) device tensor."""
    import numpy as np

    torch = _torch()
    g0, g1, g2 = grid
    qi = np.arange(first_q, first_q + batch, dtype=np.int64)
    q = np.stack([(qi // (g2 * g1)) % g0 / g0, (qi // g2) % g1 / g1, qi % g2 / g2], axis=1)
    return torch.from_numpy(np.ascontiguousarray(q, dtype=np.float64)).to(device)


def spinwave_n54(model, q, diag, W, U, info, batch=None, stream=None):
    """geneig_spinwave_n54_launch: solve (eta, H(q) + diag I) with H(q) built inside
    the solve from the model tables (shift, r, J); diag may be None."""
    torch = _torch()
    shift, r, J = model
    batch = q.shape[0] if batch is None else batch
    _require(q, torch.float64, "q")
    if diag is not None:
        _require(diag, torch.float64, "diag")
    _check_batch(54, batch, U, U, W, U, info)
    _check_hip(lib().geneig_spinwave_n54_launch(int(shift.numel()), _ptr(shift), _ptr(r), _ptr(J), _ptr(q),
                                                _ptr(diag), _ptr(W), _ptr(U), _ptr(info), int(batch), _stream(stream)),
               "geneig_spinwave_n54_launch")


def spinwave_n54_build(model, q, diag=None, stream=None):
    """The H(q) + diag I matrices the fused solve builds, as a (batch, 54, 54) tensor."""
    torch = _torch()
    shift, r, J = model
    batch = q.shape[0]
    H = torch.empty((batch, 54, 54), dtype=torch.complex128, device=q.device)
    _check_hip(lib().geneig_spinwave_n54_build(int(shift.numel()), _ptr(shift), _ptr(r), _ptr(J), _ptr(q),
                                               _ptr(diag), _ptr(H), int(batch), _stream(stream)),
               "geneig_spinwave_n54_build")
    return H


from . import learn  # noqa: E402  (learned warm start; needs the names above)
