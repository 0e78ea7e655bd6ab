// solver_hpp_driver.cpp -- a C++ caller of include/geneig_solver.hpp written
// against the reference's spellings (NvrtcGeneigSolver, TuningMode, CacheMode,
// ModuleCacheKey, make_streams / destroy_streams: src/nvrtc/nvrtc_solver.hpp
// :27-217), the way the reference's own C++ tests drive its solver
// (tests/test_nvrtc_n54.cpp, tests/test_nvrtc_multiN.cpp, tests/test_cache_n54.cpp).
//
//   solver_hpp_driver info
//       Device-free: the aliases name the library's classes, cache_dir() and
//       cubin_path() give the library's paths, and a constructor that cannot
//       reach a device throws std::runtime_error with the library's message.
//   solver_hpp_driver solve <n> <batch> <in.bin> <out.bin>
//       in.bin: H then S, batch x n x n complex128 each (column-major per
//       pair).  Solves with launch(), again with launch_chunked() over
//       make_streams(2) (which must agree bit for bit), checks the accessors
//       and the module cache (a second solver of the same key comes from the
//       in-process layer; cubin_path() of the key is the file on disk), and
//       writes W (batch x n double), U (batch x n x n complex128) and info
//       (batch int32) of the launch() call to out.bin.
// Every check prints one "check <name> ok|FAIL ..." line; the exit status is
// the number of failed checks (2 for a usage or I/O error).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

#include "geneig_solver.hpp"

static_assert(std::is_same<NvrtcGeneigSolver, geneig::GeneigSolver>::value, "NvrtcGeneigSolver alias");
static_assert(std::is_same<ModuleCacheKey, geneig::ModuleCacheKey>::value, "ModuleCacheKey alias");
static_assert(std::is_same<decltype(TuningMode::Autotune), geneig::TuningMode>::value, "TuningMode alias");
static_assert(std::is_same<decltype(CacheMode::NoDisk), geneig::CacheMode>::value, "CacheMode alias");

namespace {

int failures = 0;

void check(bool ok, const char* name, const std::string& detail = "") {
    std::printf("check %s %s%s%s\n", name, ok ? "ok" : "FAIL", detail.empty() ? "" : " ", detail.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) check(false, what, hipGetErrorString(e));
}

bool ends_with(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

bool file_exists(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    return (bool)f;
}

int run_info() {
    const std::string dir = NvrtcGeneigSolver::cache_dir();
    check(!dir.empty(), "cache_dir", dir);
    // aggregate initialization in the reference's field order {n, bpb, arch, block_dim_x}
    const ModuleCacheKey key{32, 1, 950, 64};
    const std::string path = NvrtcGeneigSolver::cubin_path(key);
    check(path.compare(0, dir.size(), dir) == 0 && ends_with(path, ".hsaco") &&
              path.find("/32_1_gfx950") != std::string::npos,
          "cubin_path", path);
    const ModuleCacheKey other{32, 2, 950, 64};
    check(key < other && !(other < key), "key_order");
    bool threw = false;
    std::string msg;
    try {
        NvrtcGeneigSolver s(32, 9999);          // no such device
    } catch (const std::runtime_error& e) {
        threw = true;
        msg = e.what();
    }
    check(threw && msg.find("GeneigSolver") != std::string::npos, "ctor_bad_device_throws", msg);
    return failures;
}

template <typename T>
bool read_all(std::ifstream& f, std::vector<T>& v) {
    return (bool)f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

int run_solve(int n, int batch, const char* in_path, const char* out_path) {
    const size_t mat = (size_t)n * n, cplx = 16;
    std::vector<double> H(2 * mat * batch), S(2 * mat * batch);
    {
        std::ifstream f(in_path, std::ios::binary);
        if (!f || !read_all(f, H) || !read_all(f, S)) {
            std::fprintf(stderr, "cannot read %s\n", in_path);
            return 2;
        }
    }
    void *dH = nullptr, *dS = nullptr, *dU = nullptr, *dU2 = nullptr;
    double *dW = nullptr, *dW2 = nullptr;
    int *dI = nullptr, *dI2 = nullptr;
    if (hipMalloc(&dH, mat * batch * cplx) || hipMalloc(&dS, mat * batch * cplx) || hipMalloc(&dU, mat * batch * cplx) ||
        hipMalloc(&dU2, mat * batch * cplx) || hipMalloc(&dW, sizeof(double) * n * batch) ||
        hipMalloc(&dW2, sizeof(double) * n * batch) || hipMalloc(&dI, sizeof(int) * batch) ||
        hipMalloc(&dI2, sizeof(int) * batch)) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 2;
    }
    hip_ok(hipMemcpy(dH, H.data(), mat * batch * cplx, hipMemcpyHostToDevice), "copy_in");
    hip_ok(hipMemcpy(dS, S.data(), mat * batch * cplx, hipMemcpyHostToDevice), "copy_in");

    NvrtcGeneigSolver solver(n, 0, 0, 0, TuningMode::Suggested, CacheMode::Auto);
    check(solver.matrix_size() == n, "matrix_size");
    const int bpb = solver.batches_per_block();
    const dim3 bd = solver.block_dim();
    // 64 threads per pair (bpb pairs per workgroup) up to n = 64, one pair on
    // 256 threads above
    check(bpb >= 1 && bd.x == (n > 64 ? 256u : 64u) && bd.y == (unsigned)bpb && bd.z == 1 &&
              (bd.x == 64 || bpb == 1),
          "block_dim", std::to_string(bd.x) + "x" + std::to_string(bd.y) + "x" + std::to_string(bd.z));
    check(solver.shared_mem_bytes() > 0, "shared_mem_bytes", std::to_string(solver.shared_mem_bytes()));
    check(!solver.was_autotuned() && !solver.autotune_sweep_ran(), "not_autotuned");
    const std::string layer = solver.cache_layer_used();
    check(layer == "compile" || layer == "disk" || layer == "in-process", "cache_layer", layer);
    check(solver.acquire_module_ms() >= 0.0, "acquire_module_ms");

    hipStream_t stream = nullptr;
    solver.launch(dH, dS, dW, dU, dI, batch, stream);
    check(hipStreamSynchronize(stream) == hipSuccess, "launch");

    std::vector<hipStream_t> streams = make_streams(2);
    check(streams.size() == 2 && streams[0] && streams[1], "make_streams");
    solver.launch_chunked(dH, dS, dW2, dU2, dI2, batch, (batch + 1) / 2, 2, streams);
    for (hipStream_t s : streams) hip_ok(hipStreamSynchronize(s), "chunked_sync");
    destroy_streams(streams);
    check(streams.empty(), "destroy_streams");

    std::vector<double> W(n * (size_t)batch), W2(n * (size_t)batch), U(2 * mat * batch), U2(2 * mat * batch);
    std::vector<int> info(batch), info2(batch);
    hip_ok(hipMemcpy(W.data(), dW, W.size() * 8, hipMemcpyDeviceToHost), "copy_out");
    hip_ok(hipMemcpy(W2.data(), dW2, W2.size() * 8, hipMemcpyDeviceToHost), "copy_out");
    hip_ok(hipMemcpy(U.data(), dU, U.size() * 8, hipMemcpyDeviceToHost), "copy_out");
    hip_ok(hipMemcpy(U2.data(), dU2, U2.size() * 8, hipMemcpyDeviceToHost), "copy_out");
    hip_ok(hipMemcpy(info.data(), dI, info.size() * 4, hipMemcpyDeviceToHost), "copy_out");
    hip_ok(hipMemcpy(info2.data(), dI2, info2.size() * 4, hipMemcpyDeviceToHost), "copy_out");
    check(std::memcmp(W.data(), W2.data(), W.size() * 8) == 0 && std::memcmp(U.data(), U2.data(), U.size() * 8) == 0 &&
              info == info2,
          "launch_chunked_bitwise");

    // a size the API refuses throws with the library's message
    bool threw = false;
    std::string msg;
    try {
        NvrtcGeneigSolver bad(129, 0);
    } catch (const std::runtime_error& e) {
        threw = true;
        msg = e.what();
    }
    check(threw, "ctor_n129_throws", msg);

    // the module cache: a second solver of the same key shares the loaded module
    {
        NvrtcGeneigSolver again(n, 0, bpb, 0, TuningMode::Suggested, CacheMode::Auto);
        check(std::string(again.cache_layer_used()) == "in-process", "second_solver_in_process",
              again.cache_layer_used());
    }
    const ModuleCacheKey key{n, bpb, 950, (int)bd.x};
    const std::string path = NvrtcGeneigSolver::cubin_path(key);
    check(file_exists(path), "cubin_path_on_disk", path);

    std::ofstream o(out_path, std::ios::binary);
    o.write(reinterpret_cast<const char*>(W.data()), (std::streamsize)(W.size() * 8));
    o.write(reinterpret_cast<const char*>(U.data()), (std::streamsize)(U.size() * 8));
    o.write(reinterpret_cast<const char*>(info.data()), (std::streamsize)(info.size() * 4));
    check((bool)o, "write_output");
    for (void* p : {dH, dS, dU, dU2, static_cast<void*>(dW), static_cast<void*>(dW2), static_cast<void*>(dI),
                    static_cast<void*>(dI2)})
        hip_ok(hipFree(p), "free");
    return failures;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "info") == 0) return run_info();
    if (argc == 6 && std::strcmp(argv[1], "solve") == 0)
        return run_solve(std::atoi(argv[2]), std::atoi(argv[3]), argv[4], argv[5]);
    std::fprintf(stderr, "usage: %s info | solve <n> <batch> <in.bin> <out.bin>\n", argv[0]);
    return 2;
}
