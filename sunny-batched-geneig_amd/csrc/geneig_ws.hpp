// geneig_ws.hpp -- device workspace shared by the solve kernels.
//
// The solve runs as K1 (factor, reduce, tridiagonalize, form Q), K2 (QL with
// the rotation product in registers) and K3 (sort, Y = Q Z, back-transform);
// they hand d, e, Q and Z over in a workspace of Ws<n>::PAIR_BYTES per pair.
// A batch is processed in chunks so the workspace stays bounded (default
// 16 GiB, GENEIG_WORKSPACE_MB overrides).  Workspaces are kept per (device,
// stream), allocated on first use and grown when a larger launch needs more,
// so launches on different streams never share one; launches that overlap on
// the same stream are ordered by the stream itself.
//
// Every launcher and every reserve entry point sizes its workspace with the
// functions below, so a reserve followed by a launch of the same batch never
// grows the buffer (the condition for capturing the launch into a graph).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace geneig_host {

constexpr size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// Chunk size (pairs) for a batch given the per-pair workspace bytes.
int chunk_pairs(int batch, unsigned long long pair_bytes);

// Control block of a solver region of `slots` slots: the K2 overflow list
// (its length, then one slot index per entry).
constexpr size_t ctl_bytes(int slots) { return round256(sizeof(int) * ((size_t)slots + 1)); }
// One solver region: `slots` x pair_bytes of stage arrays, then the control block.
constexpr size_t region_bytes(int slots, unsigned long long pair_bytes) {
    return round256((size_t)slots * pair_bytes) + ctl_bytes(slots);
}
// The control block of a region that starts at `base`.
inline int* region_ctl(void* base, int slots, unsigned long long pair_bytes) {
    return reinterpret_cast<int*>(static_cast<unsigned char*>(base) + round256((size_t)slots * pair_bytes));
}

// Concurrent parts of one launch.  A batch that fits one workspace chunk can
// be cut into parts whose kernel pipelines run on separate streams, so one
// part's memory-bound or low-occupancy stages overlap another's compute-bound
// ones.  fork_streams hands out `parts` (<= kMaxParts) internal streams of the
// caller's stream, kept per (device, caller stream) for the process lifetime
// (non-blocking, default priority), each made to wait for the work already
// queued on the caller's stream; join_streams makes the caller's stream wait
// for everything queued on them.  The record/wait pairs go through events,
// so the pattern is also valid under stream capture.  prepare_parts creates
// the streams and events ahead of time (the reserve entry points call it, so
// a captured launch creates nothing).
constexpr int kMaxParts = 4;
hipError_t prepare_parts(hipStream_t stream, int parts);
hipError_t fork_streams(hipStream_t stream, int parts, hipStream_t* subs);
hipError_t join_streams(hipStream_t stream, int parts, const hipStream_t* subs);

// Parts to cut a batch into: `want` when every part keeps at least
// kMinPartPairs pairs and the batch is one workspace chunk, else 1.
constexpr int kMinPartPairs = 8192;
int part_count(int batch, int chunk, int want);

// Layout of a plain (FP64 pipeline) launch: `parts` regions of `part_slots`
// slots, or one region of `chunk` slots walked chunk by chunk.
struct PlainLayout {
    int chunk = 0, parts = 1, part_slots = 0;
    size_t region = 0, bytes = 0;
};
PlainLayout plain_layout(int batch, unsigned long long pair_bytes, int want_parts);

// Layout of a precision-form launch: the fallback list (its length, then
// one pair index per entry), then the solver area shared by the complex64
// stages (parts regions or one chunked region) and, after the refinement, by
// the FP64 re-solve of the listed pairs (one region of chunk_d slots).
// extra_pair_bytes: per-pair bytes the complex64 stages need beside their
// stage arrays (the learned form's network activations).
struct MixedLayout {
    int chunk_f = 0, chunk_d = 0, parts = 1, part_slots = 0, rounds = 0;
    size_t list_bytes = 0, region_f = 0, region_d = 0, bytes = 0;
};
MixedLayout mixed_layout(int batch, unsigned long long pair_bytes_f, unsigned long long pair_bytes_d,
                         int want_parts, unsigned long long extra_pair_bytes = 0);
// The layout of the static n = 54 precision-form launch (geneig_mixed.hip).
MixedLayout mixed_n54_layout(int batch);

// Workspace of at least `bytes` for `stream` on the current device.  Growing
// it while `stream` is being captured is refused (hipErrorStreamCaptureUnsupported:
// reserve before capturing).  Once a stream has been seen under capture, a
// buffer it outgrows is retired, not freed, since a captured graph may still
// use it; geneig_workspace_release frees both.
hipError_t workspace(hipStream_t stream, size_t bytes, void** out);

// Frees the workspace (current and retired buffers) of `stream` on the
// current device after the stream has drained; *freed receives the bytes.
hipError_t release(hipStream_t stream, size_t* freed);

// Entries the QL rotation log may take per pair: GENEIG_QL_LOG_ENTRIES when set
// (read at every launch; a small value sends every pair through the one-pair
// QL that handles an overflowed log, which is how that path is tested), else
// no limit below the compiled capacity.
int ql_log_entries();

}  // namespace geneig_host
