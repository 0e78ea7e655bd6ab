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

}  // namespace geneig_host
