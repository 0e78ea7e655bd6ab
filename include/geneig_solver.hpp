// geneig_solver.hpp -- C++ interface of the runtime-n solver, header-only over
// the C ABI of geneig_solver.h.  Same constructor arguments, launch methods,
// introspection and statics as the reference's NvrtcGeneigSolver
// (src/nvrtc/nvrtc_solver.hpp:81-217), with HIP types: device pointers are
// void* (hipDeviceptr_t) instead of CUdeviceptr and streams are hipStream_t.
// Errors throw std::runtime_error carrying the library's message, as the
// reference does (src/nvrtc/nvrtc_solver.cpp:49-91).
//
// The classes live in namespace geneig.  The reference's own spellings
// (NvrtcGeneigSolver, TuningMode, CacheMode, ModuleCacheKey, make_streams,
// destroy_streams, all at global scope: nvrtc_solver.hpp:27-79) are declared
// at the end as aliases, so a caller written against the reference compiles
// once its CUDA types are replaced by HIP ones; define
// GENEIG_NO_REFERENCE_NAMES to leave the global namespace alone.
#pragma once

#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "geneig_solver.h"

namespace geneig {

enum class TuningMode { Suggested = GENEIG_TUNING_SUGGESTED, Autotune = GENEIG_TUNING_AUTOTUNE };
enum class CacheMode { Auto = GENEIG_CACHE_AUTO, NoDisk = GENEIG_CACHE_NODISK, NoCache = GENEIG_CACHE_NOCACHE };

// Cache key of one compiled module (nvrtc_solver.hpp:67-79): two solvers with
// equal keys share one loaded module.  arch is the gfx target number, 950 for
// gfx950 (the reference's int compute capability, 90 for sm_90); variant
// selects the module (0 the FP64 solve, 1 the precision-form one) and is not
// part of the reference's key.
struct ModuleCacheKey {
    int n = 0;
    int bpb = 0;
    int arch = 0;
    int block_dim_x = 64;
    int variant = 0;
    bool operator<(const ModuleCacheKey& o) const {
        return std::tie(n, bpb, arch, block_dim_x, variant) < std::tie(o.n, o.bpb, o.arch, o.block_dim_x, o.variant);
    }
    // The target name the library keys its modules on: the full name of a
    // visible device of this target (gfx950:sramecc+:xnack-, features
    // included, since code objects are built for them), else gfx<arch>.
    std::string arch_name() const {
        const std::string base = "gfx" + std::to_string(arch);
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess) return base;
        for (int d = 0; d < count; ++d) {
            hipDeviceProp_t p;
            if (hipGetDeviceProperties(&p, d) != hipSuccess) continue;
            const std::string full = p.gcnArchName;
            if (full.compare(0, base.size(), base) == 0 && (full.size() == base.size() || full[base.size()] == ':'))
                return full;
        }
        return base;
    }
};

inline void check(int rc, const char* what) {
    if (rc != GENEIG_E_OK) throw std::runtime_error(std::string(what) + ": " + geneig_last_error());
}

class GeneigSolver {
public:
    explicit GeneigSolver(int n, int device_id, int batches_per_block_request = 0, int force_block_dim_x = 0,
                          TuningMode tuning_mode = TuningMode::Suggested, CacheMode cache_mode = CacheMode::Auto) {
        check(geneig_solver_create(&h_, n, device_id, batches_per_block_request, force_block_dim_x,
                                   static_cast<int>(tuning_mode), static_cast<int>(cache_mode)),
              "GeneigSolver");
    }
    ~GeneigSolver() { geneig_solver_destroy(h_); }
    GeneigSolver(const GeneigSolver&) = delete;
    GeneigSolver& operator=(const GeneigSolver&) = delete;
    GeneigSolver(GeneigSolver&&) = delete;
    GeneigSolver& operator=(GeneigSolver&&) = delete;

    void launch(const void* d_H, const void* d_S, double* d_W, void* d_U, int* d_info, int batch_size,
                hipStream_t stream) {
        check(geneig_solver_launch(h_, d_H, d_S, d_W, d_U, d_info, batch_size, stream), "GeneigSolver::launch");
    }
    void launch_chunked(const void* d_H, const void* d_S, double* d_W, void* d_U, int* d_info,
                        int total_matrices, int chunk_size, int num_streams,
                        const std::vector<hipStream_t>& streams) {
        if ((int)streams.size() != num_streams)
            throw std::runtime_error("launch_chunked: streams.size() must equal num_streams");
        check(geneig_solver_launch_chunked(h_, d_H, d_S, d_W, d_U, d_info, total_matrices, chunk_size,
                                           num_streams, streams.data()),
              "GeneigSolver::launch_chunked");
    }
    // Host-resident pairs (pinned for copy overlap): copy in, solve, copy out
    // per chunk on its stream; outputs complete once the streams are synchronized.
    void launch_chunked_host(const void* h_H, const void* h_S, double* h_W, void* h_U, int* h_info,
                             int total_matrices, int chunk_size, const std::vector<hipStream_t>& streams) {
        check(geneig_solver_launch_chunked_host(h_, h_H, h_S, h_W, h_U, h_info, total_matrices, chunk_size,
                                                (int)streams.size(), streams.data()),
              "GeneigSolver::launch_chunked_host");
    }

    int matrix_size() const { return geneig_solver_matrix_size(h_); }
    int batches_per_block() const { return geneig_solver_batches_per_block(h_); }
    dim3 block_dim() const {
        unsigned x = 0, y = 0, z = 0;
        geneig_solver_block_dim(h_, &x, &y, &z);
        return dim3(x, y, z);
    }
    unsigned int shared_mem_bytes() const { return geneig_solver_shared_mem_bytes(h_); }
    int autotuned_block_dim_x() const { return geneig_solver_autotuned_block_dim_x(h_); }
    // Rows per chunk of K1's register walks: the tile the autotuner picked, or
    // the build's default (no counterpart in the reference).
    int k1_chunk_rows() const { return geneig_solver_k1_chunk_rows(h_); }
    bool was_autotuned() const { return geneig_solver_was_autotuned(h_) != 0; }
    bool autotune_sweep_ran() const { return geneig_solver_autotune_sweep_ran(h_) != 0; }
    const char* cache_layer_used() const { return geneig_solver_cache_layer_used(h_); }
    double acquire_module_ms() const { return geneig_solver_acquire_module_ms(h_); }

    static std::string cache_dir() {
        std::string s(geneig_cache_dir(nullptr, 0) + 1, '\0');
        geneig_cache_dir(&s[0], s.size());
        s.resize(s.size() - 1);
        return s;
    }
    // The on-disk code-object path of a key (the reference names it a cubin).
    static std::string cubin_path(const ModuleCacheKey& key) {
        const std::string arch = key.arch_name();
        const size_t len = geneig_cache_path(key.n, key.bpb, arch.c_str(), key.block_dim_x, key.variant, nullptr, 0);
        std::string s(len + 1, '\0');
        geneig_cache_path(key.n, key.bpb, arch.c_str(), key.block_dim_x, key.variant, &s[0], s.size());
        s.resize(len);
        return s;
    }

private:
    geneig_solver_t h_ = nullptr;
};

// make_streams / destroy_streams (nvrtc_solver.cpp:1109-1125).
inline std::vector<hipStream_t> make_streams(int n) {
    std::vector<hipStream_t> v(n, nullptr);
    check(geneig_make_streams(n, v.data()), "make_streams");
    return v;
}
inline void destroy_streams(std::vector<hipStream_t>& streams) {
    geneig_destroy_streams((int)streams.size(), streams.data());
    streams.clear();
}

}  // namespace geneig

#ifndef GENEIG_NO_REFERENCE_NAMES
// The reference's global names (nvrtc_solver.hpp:27-79, 81).
using NvrtcGeneigSolver = geneig::GeneigSolver;
using TuningMode = geneig::TuningMode;
using CacheMode = geneig::CacheMode;
using ModuleCacheKey = geneig::ModuleCacheKey;
using geneig::destroy_streams;
using geneig::make_streams;
#endif
