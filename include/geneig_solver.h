/*
 * geneig_solver.h -- C ABI of the runtime-n solver (per-n specialization by
 * hipRTC, module cache, autotuner).  Replaces the reference's
 * NvrtcGeneigSolver (src/nvrtc/nvrtc_solver.hpp:81-217).
 *
 * Errors: every function returning int returns 0 on success and a nonzero
 * GENEIG_E_* code otherwise; geneig_last_error() gives the message of the
 * last failure on the calling thread (the reference threw std::runtime_error,
 * src/nvrtc/nvrtc_solver.cpp:49-91).
 */
#ifndef GENEIG_SOLVER_H
#define GENEIG_SOLVER_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GENEIG_E_OK 0
#define GENEIG_E_ARG 1       /* invalid argument */
#define GENEIG_E_HIP 2       /* HIP runtime / module error */
#define GENEIG_E_RTC 3       /* hipRTC compile error */
#define GENEIG_E_FIT 4       /* configuration does not fit the device */
#define GENEIG_E_TUNE 5      /* autotune found no valid candidate */

/* TuningMode (nvrtc_solver.hpp:38): how the launch shape is chosen. */
#define GENEIG_TUNING_SUGGESTED 0
#define GENEIG_TUNING_AUTOTUNE 1
/* CacheMode (nvrtc_solver.hpp:47): module cache layers. */
#define GENEIG_CACHE_AUTO 0
#define GENEIG_CACHE_NODISK 1
#define GENEIG_CACHE_NOCACHE 2

typedef struct geneig_solver* geneig_solver_t;

/* Constructor, NvrtcGeneigSolver(n, device_id, bpb, force_block_dim_x,
 * TuningMode, CacheMode) (nvrtc_solver.hpp:99-104, nvrtc_solver.cpp:927-1047).
 * bpb_request <= 0 selects pairs-per-workgroup automatically; force_block_dim_x
 * > 0 fixes threads per pair (64 = one wavefront, the only shape built) and
 * skips autotuning. */
int geneig_solver_create(geneig_solver_t* out, int n, int device_id, int bpb_request,
                         int force_block_dim_x, int tuning_mode, int cache_mode);
void geneig_solver_destroy(geneig_solver_t s);

/* NvrtcGeneigSolver::launch (nvrtc_solver.cpp:1063-1076).  Unlike the
 * reference, a partial last workgroup is handled (no padding needed). */
int geneig_solver_launch(geneig_solver_t s, const void* d_H, const void* d_S, double* d_W,
                         void* d_U, int* d_info, int batch_size, hipStream_t stream);
/* launch in place, with LAPACK zhegvd's storage (the runtime-n counterpart of
 * geneig_full_n54_batched_inplace_launch): d_H receives the eigenvectors,
 * the lower triangle of d_S the Cholesky factor.  FP64-form solvers only (a
 * precision-form solver refines against the original H and S and refuses). */
int geneig_solver_launch_inplace(geneig_solver_t s, void* d_H, void* d_S, double* d_W, int* d_info,
                                 int batch_size, hipStream_t stream);
/* NvrtcGeneigSolver::launch_chunked (nvrtc_solver.cpp:1078-1107). */
int geneig_solver_launch_chunked(geneig_solver_t s, const void* d_H, const void* d_S,
                                 double* d_W, void* d_U, int* d_info, int total_matrices,
                                 int chunk_size, int num_streams, const hipStream_t* streams);

/* launch_chunked for host-resident pairs: chunks of chunk_size pairs go
 * round-robin over the streams, and on its stream each chunk is copied in
 * (H, S), solved and copied out (W, U, info), so with two or more streams the
 * copies of one chunk overlap the solve of another.  The reference's
 * launch_chunked (nvrtc_solver.cpp:1078-1107) takes device pointers only and
 * its copy overlap was deferred (docs/PHASE3_NVRTC.md:76-89); this entry has
 * no counterpart there.  Host buffers should be pinned (hipHostMalloc):
 * pageable ones give the same results without overlap.  The solver keeps one
 * device staging slot of chunk_size pairs per stream (allocated on first use,
 * grown on demand, freed by geneig_solver_destroy).  Returns once everything
 * is enqueued; the host outputs are complete after every stream is
 * synchronized. */
int geneig_solver_launch_chunked_host(geneig_solver_t s, const void* h_H, const void* h_S,
                                      double* h_W, void* h_U, int* h_info, int total_matrices,
                                      int chunk_size, int num_streams, const hipStream_t* streams);

/* Allocates, ahead of time, the device workspace a launch of batch_size
 * pairs on `stream` will use, so that the launch itself allocates nothing
 * and can be captured into a graph.  No counterpart in the reference (whose
 * kernel used shared memory only). */
int geneig_solver_reserve(geneig_solver_t s, int batch_size, hipStream_t stream);
/* Workspace bytes a launch of batch_size pairs uses on its stream (what
 * geneig_solver_reserve allocates); geneig_workspace_release (geneig.h)
 * frees a stream's workspace. */
size_t geneig_solver_workspace_bytes(geneig_solver_t s, int batch_size);

/* Accessors (nvrtc_solver.hpp:120-140). */
int geneig_solver_matrix_size(geneig_solver_t s);
int geneig_solver_batches_per_block(geneig_solver_t s);
void geneig_solver_block_dim(geneig_solver_t s, unsigned* x, unsigned* y, unsigned* z);
unsigned geneig_solver_shared_mem_bytes(geneig_solver_t s);
int geneig_solver_autotuned_block_dim_x(geneig_solver_t s);
int geneig_solver_was_autotuned(geneig_solver_t s);
int geneig_solver_autotune_sweep_ran(geneig_solver_t s);
const char* geneig_solver_cache_layer_used(geneig_solver_t s); /* "in-process"|"disk"|"compile" */
double geneig_solver_acquire_module_ms(geneig_solver_t s);
/* Eigen-algorithm variant of the loaded module (0 = tridiagonal QL). */
int geneig_solver_variant(geneig_solver_t s);
/* Rows per chunk of K1's register walks in the loaded module (n <= 64; 0
 * above, where K1 works in LDS): the tile-shape knob the autotuner sweeps (4,
 * 8 or 16 at the winning pairs per workgroup), else the build's default (8 in
 * FP64, 16 in complex64).  No counterpart in the reference, whose autotuner
 * sweeps threads per pair (nvrtc_solver.cpp:649-925). */
int geneig_solver_k1_chunk_rows(geneig_solver_t s);

/* NvrtcGeneigSolver::cache_dir / cubin_path (nvrtc_solver.cpp:403-431).
 * Copies a NUL-terminated path into buf; returns the full length. */
size_t geneig_cache_dir(char* buf, size_t len);
size_t geneig_cache_path(int n, int bpb, const char* arch, int block_dim_x, int variant,
                         char* buf, size_t len);

/* Device-free hipRTC compile of the per-n kernel into a code object (ELF).
 * Used by the build check and by the cache tests; *size receives the object
 * size, and when buf != NULL and cap >= size the object is copied there. */
int geneig_rtc_compile(int n, int bpb, int variant, const char* arch, void* buf, size_t cap,
                       size_t* size);
/* The same with K1's row-chunk width set (4, 8 or 16; 0 = the build's
 * default), as the autotuner's candidates are built. */
int geneig_rtc_compile_rows(int n, int bpb, int variant, int k1_rows, const char* arch, void* buf, size_t cap,
                            size_t* size);
/* The same with the threads per pair of K1 / K3 (block_dim_x 64: one wave per
 * pair; 256: one pair per workgroup of four waves, bpb must be 1). */
int geneig_rtc_compile_bdx(int n, int bpb, int variant, int k1_rows, int block_dim_x, const char* arch, void* buf,
                           size_t cap, size_t* size);
uint64_t geneig_source_hash(void);
uint64_t geneig_fnv1a64(const void* data, size_t len, uint64_t seed);

/* Helpers for launch_chunked callers (make_streams/destroy_streams,
 * nvrtc_solver.cpp:1109-1125). */
int geneig_make_streams(int n, hipStream_t* out);
void geneig_destroy_streams(int n, hipStream_t* streams);

const char* geneig_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GENEIG_SOLVER_H */
