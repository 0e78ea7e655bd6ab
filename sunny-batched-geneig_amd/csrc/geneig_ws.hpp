// geneig_ws.hpp -- device workspace shared by the three solve kernels.
//
// The solve runs as K1 (factor, reduce, tridiagonalize, form Q), K2 (QL with
// the rotation product in registers) and K3 (sort, Y = Q Z, back-transform);
// they hand d, e, Q and Z over in a workspace of Ws<n>::PAIR_BYTES per pair.
// A batch is processed in chunks so the workspace stays bounded (default
// 16 GiB, GENEIG_WORKSPACE_MB overrides).  Workspaces are kept per (device,
// stream), allocated on first use and grown when a larger chunk is needed,
// so launches on different streams never share one; launches that overlap on
// the same stream are ordered by the stream itself.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace geneig_host {

// Chunk size (pairs) for a batch given the per-pair workspace bytes.
int chunk_pairs(int batch, unsigned long long pair_bytes);

// Workspace of at least `bytes` for `stream` on the current device.
hipError_t workspace(hipStream_t stream, size_t bytes, void** out);

// Entries the QL rotation log may take per pair: GENEIG_QL_LOG_ENTRIES when set
// (read at every launch; a small value sends every pair through the one-pair
// QL that handles an overflowed log, which is how that path is tested), else
// no limit below the compiled capacity.
int ql_log_entries();

// Concurrent parts of one launch.  A batch that fits one workspace chunk can
// be cut into parts whose kernel pipelines run on separate streams, so one
// part's memory-bound or low-occupancy stages overlap another's compute-bound
// ones.  fork_streams hands out `parts` (<= kMaxParts) internal streams of the
// caller's stream, kept per (device, caller stream) for the process lifetime
// (non-blocking, default priority), each made to wait for the work already
// queued on the caller's stream; join_streams makes the caller's stream wait
// for everything queued on them.  The record/wait pairs go through events,
// so the pattern is also valid under stream capture.
constexpr int kMaxParts = 4;
hipError_t fork_streams(hipStream_t stream, int parts, hipStream_t* subs);
hipError_t join_streams(hipStream_t stream, int parts, const hipStream_t* subs);

// Parts to cut a batch into: `want` when every part keeps at least
// kMinPartPairs pairs and the batch is one workspace chunk, else 1.
constexpr int kMinPartPairs = 8192;
int part_count(int batch, int chunk, int want);

}  // namespace geneig_host
