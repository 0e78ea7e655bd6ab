/* geneig_multi.h -- mixed-size batches of generalized eigensolves in one
 * graph launch (gfx950).
 *
 * No counterpart in the reference: it solves one n per NvrtcGeneigSolver
 * (src/nvrtc/nvrtc_solver.hpp:81-217) and leaves mixing sizes to the caller.
 * A geneig_multi_t holds one runtime-n solver per size (hipRTC-specialized,
 * pairs per workgroup from the autotuner when tuning_mode is
 * GENEIG_TUNING_AUTOTUNE) and solves a batch binned by size:
 *
 *   geneig_multi_launch          one hipGraphLaunch.  The per-size solves are
 *                                captured once into a graph with one branch per
 *                                size (the branches run concurrently) and the
 *                                graph is replayed while the counts and buffer
 *                                pointers stay the same; other arguments capture
 *                                a new graph.
 *   geneig_multi_launch_serial   one solve launch per size, in order, on the
 *                                caller's stream (the comparison).
 *
 * Bin k holds counts[k] pairs of size sizes[k] in its own buffers d_A[k],
 * d_B[k], d_W[k], d_U[k], d_info[k] (the layout of geneig_solver_launch).
 * The outputs of both launch forms are bit-identical.  Returns GENEIG_E_* codes
 * (geneig_solver.h); the message is in geneig_last_error().
 */
#ifndef GENEIG_MULTI_H
#define GENEIG_MULTI_H

#include <hip/hip_runtime_api.h>

#include "geneig_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct geneig_multi* geneig_multi_t;

int geneig_multi_create(geneig_multi_t* out, int num_sizes, const int* sizes, int device_id, int tuning_mode,
                        int cache_mode);
void geneig_multi_destroy(geneig_multi_t m);

int geneig_multi_launch(geneig_multi_t m, const int* counts, const void* const* d_A, const void* const* d_B,
                        double* const* d_W, void* const* d_U, int* const* d_info, hipStream_t stream);
int geneig_multi_launch_serial(geneig_multi_t m, const int* counts, const void* const* d_A,
                               const void* const* d_B, double* const* d_W, void* const* d_U, int* const* d_info,
                               hipStream_t stream);

int geneig_multi_num_sizes(geneig_multi_t m);
/* The solver of bin k (owned by m): its accessors report the tuned shape. */
geneig_solver_t geneig_multi_solver(geneig_multi_t m, int k);
/* Graphs captured and graph launches made so far. */
long long geneig_multi_captures(geneig_multi_t m);
long long geneig_multi_replays(geneig_multi_t m);

#ifdef __cplusplus
}
#endif

#endif /* GENEIG_MULTI_H */
