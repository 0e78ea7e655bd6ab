"""CPU tests of the product library: it loads, exports every symbol the
public headers declare, compiles the runtime-n kernel with hipRTC without a
device, and its cache/host logic behaves as the reference's
(src/nvrtc/nvrtc_solver.cpp:321-431).  No compute runs here."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("geneig.h", "geneig_solver.h", "geneig_gen.h", "geneig_mixed.h", "geneig_multi.h", "geneig_learn.h")]


def declared_functions():
    names = []
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", txt, flags=re.M):
            if m.group(1) not in ("defined",):
                names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("geneig_full_n54_batched_launch", "chol_n54_launch", "geneig_finish_n54_launch",
                 "geneig_solver_create", "geneig_solver_launch_chunked", "geneig_cache_path",
                 "geneig_mixed_n54_batched_launch"):
        assert must in names


def test_library_exports_every_declared_symbol(geneig):
    L = geneig.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_headers_compile_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text("".join('#include "%s"\n' % h for h in HEADERS) + "int main(void){return 0;}\n")
    subprocess.check_call(["gcc", "-std=c11", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                           "-c", str(src), "-o", str(tmp_path / "t.o")])


def test_static_queries(geneig):
    L = geneig.lib()
    assert L.geneig_full_n54_batched_block_dim_x() == 64
    lds = L.geneig_full_n54_batched_shared_mem_bytes()
    assert 16 * 54 * 55 // 2 <= lds <= 160 * 1024   # the packed lower triangle of C at least
    assert L.geneig_reduce_eig_n54_chol_smem_bytes() + L.geneig_reduce_eig_n54_heev_smem_bytes() == \
        L.geneig_reduce_eig_n54_shared_mem_bytes()


@pytest.mark.parametrize("n", [1, 2, 8, 16, 32, 48, 54, 64, 96])
def test_rtc_compiles_without_device(geneig, n):
    code = geneig.rtc_compile(n, 1)
    assert code[:4] == b"\x7fELF" and len(code) > 1024


@pytest.mark.parametrize("n", [3, 16, 33, 54, 64, 96, 128])
def test_rtc_precision_module_compiles_without_device(geneig, n):
    """The precision-form module (variant 1: FP64 Cholesky, complex64 stages,
    refinement with FP64 and FP32 matrix-core products, FP64 list kernels) for n
    up to 128."""
    code = geneig.rtc_compile(n, 1, variant=1)
    assert code[:4] == b"\x7fELF" and len(code) > 1024


@pytest.mark.parametrize("n,rows", [(32, 4), (32, 16), (54, 4), (54, 16), (64, 16)])
def test_rtc_k1_row_chunks_compile_without_device(geneig, n, rows):
    """The autotuner's tile-shape candidates (K1's row-chunk width 4 / 16
    besides the default 8) build for n <= 64, and differ from the default."""
    code = geneig.rtc_compile(n, 1, k1_rows=rows)
    assert code[:4] == b"\x7fELF" and len(code) > 1024
    assert code != geneig.rtc_compile(n, 1)


@pytest.mark.parametrize("n,variant", [(1, 0), (8, 0), (54, 0), (54, 1), (96, 1), (100, 0), (128, 0), (128, 1)])
def test_rtc_four_wave_modules_compile_without_device(geneig, n, variant):
    """Both module variants with K1 / K3 (and the Cholesky stages) on one
    workgroup of four waves per pair, n up to 128, and distinct from the
    one-wave module."""
    code = geneig.rtc_compile(n, 1, variant=variant, block_dim_x=256)
    assert code[:4] == b"\x7fELF" and len(code) > 1024
    if n <= 64:
        assert code != geneig.rtc_compile(n, 1, variant=variant)


def test_rtc_rejects_bad_arguments(geneig):
    with pytest.raises(geneig.GeneigError):
        geneig.rtc_compile(0)
    with pytest.raises(geneig.GeneigError):
        geneig.rtc_compile(200)
    with pytest.raises(geneig.GeneigError):
        geneig.rtc_compile(32, 1, variant=7)
    with pytest.raises(geneig.GeneigError, match="row chunk"):
        geneig.rtc_compile(32, 1, k1_rows=32)        # would read past the vectors' zero pad
    with pytest.raises(geneig.GeneigError, match="one pair per workgroup"):
        geneig.rtc_compile(32, 2, block_dim_x=256)
    with pytest.raises(geneig.GeneigError, match="threads per pair"):
        geneig.rtc_compile(32, 1, block_dim_x=128)


def test_cache_dir_resolution(geneig, monkeypatch):
    monkeypatch.setenv("SUNNY_GENEIG_CACHE_DIR", "/x/y")
    assert geneig.GeneigSolver.cache_dir() == "/x/y"
    monkeypatch.delenv("SUNNY_GENEIG_CACHE_DIR")
    monkeypatch.setenv("XDG_CACHE_HOME", "/xdg")
    assert geneig.GeneigSolver.cache_dir() == "/xdg/sunny_geneig/hsaco"
    monkeypatch.delenv("XDG_CACHE_HOME")
    monkeypatch.setenv("HOME", "/home/u")
    assert geneig.GeneigSolver.cache_dir() == "/home/u/.cache/sunny_geneig/hsaco"


def test_cache_path_key(geneig, monkeypatch):
    monkeypatch.setenv("SUNNY_GENEIG_CACHE_DIR", "/c")
    p = geneig.GeneigSolver.cache_path(54, 2, "gfx950:sramecc+:xnack-", 64, 0)
    src = geneig.lib().geneig_source_hash()
    assert re.fullmatch(r"/c/54_2_gfx950_sramecc__xnack-_64_v0_lib[0-9a-f]{16}_src%016x\.hsaco" % src, p), p
    assert p != geneig.GeneigSolver.cache_path(54, 1, "gfx950:sramecc+:xnack-", 64, 0)


def test_source_hash_is_fnv1a_of_device_source(geneig, oracle):
    csrc = os.path.join(ROOT, "sunny-batched-geneig_amd", "csrc")
    src = b"".join(open(os.path.join(csrc, f), "rb").read() for f in ("geneig_device.inc", "geneig_tile.inc", "geneig_rtc0.inc", "geneig_refine.inc"))
    assert geneig.lib().geneig_source_hash() == oracle.fnv1a64(src)
    buf = C.create_string_buffer(b"foobar", 6)
    assert geneig.lib().geneig_fnv1a64(C.cast(buf, C.c_void_p), 6, 0) == 0x85944171f73967e8


def test_solver_create_reports_errors(geneig):
    with pytest.raises(geneig.GeneigError, match="outside the supported range"):
        geneig.GeneigSolver(0)
    with pytest.raises(geneig.GeneigError, match="one wave per pair"):
        geneig.GeneigSolver(54, force_block_dim_x=128)
    with pytest.raises(geneig.GeneigError, match="one pair per workgroup"):
        geneig.GeneigSolver(54, batches_per_block_request=2, force_block_dim_x=256)


def test_product_does_not_touch_oracle():
    """The product sources never reference the oracle (it only checks results)."""
    pkg = os.path.join(ROOT, "sunny-batched-geneig_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".inc", ".h", "Makefile")) and "_build" not in dirpath:
                txt = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in txt.lower(), f


def _r256(b):
    return (b + 255) // 256 * 256


def _pair_bytes(n, real):
    """Workspace bytes per pair of the solve stages: d, e (n reals each), Q
    (n^2 complex) and the slot that holds the QL rotation log (guard 8 +
    capacity 3n^2/2 + 64 + read-ahead 64 ceil(n/64) entries of two reals)
    and later Z (n^2 reals)."""
    log = 2 * (8 + 3 * n * n // 2 + 64 + 64 * ((n + 63) // 64))
    return real * (2 * n + 2 * n * n + max(log, n * n))


def _region(slots, pb):
    return _r256(slots * pb) + _r256(4 * (slots + 1))


@pytest.mark.parametrize("batch", [1, 37, 8191, 16385, 24575, 24576, 131071])
def test_n54_workspace_sizing(geneig, batch):
    """The bytes a batched n = 54 launch reserves: one region per concurrent
    part (three when every part keeps >= 8192 pairs, else two, else one
    chunked region), each holding its slots' stage arrays and the overflow
    list; the precision form adds the fallback list and takes the larger of
    its complex64 regions and the FP64 re-solve region.  Odd batches round
    the part size up, so every part's pairs fit its region."""
    pd, pf = _pair_bytes(54, 8), _pair_bytes(54, 4)
    parts = 3 if batch // 3 >= 8192 else (2 if batch // 2 >= 8192 else 1)
    slots = -(-batch // parts)
    want0 = parts * _region(slots, pd)
    assert geneig.n54_workspace_bytes(batch, 0) == want0
    for i in range(parts):   # the parts the launcher cuts
        P = batch * (i + 1) // parts - batch * i // parts
        assert _region(P, pd) <= _region(slots, pd)
    want1 = _r256(4 * (batch + 1)) + max(parts * _region(slots, pf), _region(batch, pd))
    assert geneig.n54_workspace_bytes(batch, 1) == want1
    assert geneig.n54_workspace_bytes(0, 0) == 0 and geneig.n54_workspace_bytes(-5, 1) == 0


def test_noalloc_refuses_oversized_batch_before_device_work(geneig):
    """The allocation-free launch checks the batch against its fixed array
    before any HIP call: a batch one past the maximum returns an error (here
    with null pointers, on a host without a device) naming the limit."""
    L = geneig.lib()
    mx = L.geneig_full_n54_batched_noalloc_max_batch()
    assert mx == 4485
    rc = L.geneig_full_n54_batched_launch_noalloc(None, None, None, None, None, mx + 1, None)
    assert rc != 0
    assert "split the batch" in geneig.last_error() and str(mx) in geneig.last_error()
    assert L.geneig_full_n54_batched_launch_noalloc(None, None, None, None, None, 0, None) == 0


def test_rtc_refuses_non_cdna_target(geneig):
    """The precision-form module builds gfx9-format buffer descriptors and uses
    CDNA matrix-core instructions: compiling it for another target fails
    with a message naming the requirement instead of producing wrong code."""
    with pytest.raises(geneig.GeneigError, match="gfx9"):
        geneig.rtc_compile(54, 1, variant=1, arch="gfx1100")


@pytest.mark.parametrize("batch", [0, -5])
def test_launchers_treat_empty_batches_as_no_ops(geneig, batch):
    """Every batched launcher returns 0 for batch <= 0 before any HIP call
    (here with null pointers on a host without a device), as the library's
    contract and the reference's launchers (batch <= 0 returns early) have it;
    a requested fallback count reads 0."""
    L = geneig.lib()
    fb = C.c_int(7)
    assert L.geneig_full_n54_batched_launch(None, None, None, None, None, batch, None) == 0
    assert L.geneig_full_n54_batched_inplace_launch(None, None, None, None, batch, None) == 0
    assert L.geneig_eta_paired_n54_batched_launch(None, None, None, None, batch, None) == 0
    assert L.geneig_mixed_n54_batched_launch(None, None, None, None, None, batch, 0, C.c_double(0.0), None,
                                             C.byref(fb), None) == 0 and fb.value == 0
    fb.value = 7
    assert L.geneig_mixed_eta_n54_batched_launch(None, None, None, None, batch, 0, C.c_double(0.0), None,
                                                 C.byref(fb), None) == 0 and fb.value == 0
    fb.value = 7
    assert L.geneig_mixed_eta_paired_n54_batched_launch(None, None, None, None, batch, 0, C.c_double(0.0), None,
                                                        C.byref(fb), None) == 0 and fb.value == 0
    fb.value = 7
    assert L.geneig_learned_n54_batched_launch(None, None, None, None, None, batch, None, None, 0,
                                               C.c_float(0.0), 0, C.c_double(0.0), None, None, C.byref(fb),
                                               None) == 0 and fb.value == 0
    assert L.geneig_spinwave_n54_launch(0, None, None, None, None, None, None, None, None, batch, None) == 0
