// geneig_solver.cpp -- runtime-n solver: hipRTC per-n specialization, the
// three-layer module cache and the autotuner.  MI355X counterpart of the
// reference's NvrtcGeneigSolver (src/nvrtc/nvrtc_solver.{hpp,cpp}).
//
// What is specialized per n: the whole pipeline of geneig_device.inc is
// compiled by hipRTC with -DGENEIG_RTC_N=n and -DGENEIG_RTC_BPB=pairs per
// workgroup, so every loop bound and LDS offset is a compile-time constant.
// A compile takes seconds (no linker step against a closed library), so the
// disk cache mostly saves start-up time rather than minutes.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/geneig_solver.h"
#include "../../include/geneig_mixed.h"
#include "geneig_err.hpp"
#include "geneig_ws.hpp"
#include "geneig_device_src.inc"  // kGeneigDeviceSource: geneig_device.inc as a string

namespace {
thread_local std::string g_last_error;
}  // namespace

int geneig_host::fail(int code, const char* fmt, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    std::fprintf(stderr, "[geneig] %s\n", buf);
    return code;
}

namespace {

using geneig_host::fail;
// Grids of the dispatches that walk device-built lists (the same values as
// kListGrid* of geneig_device.inc, which the host part of this file does not
// include).
constexpr int kListGridK1 = 4096, kListGridK2a = 512, kListGridK2x = 1024;
// Threads per pair of the four-wave K1 / K3 (kTileThreads of geneig_tile.inc).
constexpr int kTileThreadsHost = 256;
#define HIP_TRY(expr) GENEIG_HIP_TRY(expr)

// ---- content hashes (FNV-1a 64, the reference's cache fingerprint,
//      src/nvrtc/nvrtc_solver.cpp:321-333) ------------------------------------
constexpr uint64_t kFnvOffset = 0xcbf29ce484222325ULL;
constexpr uint64_t kFnvPrime = 0x100000001b3ULL;

uint64_t fnv1a(const void* data, size_t n, uint64_t h = kFnvOffset) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= kFnvPrime;
    }
    return h;
}

uint64_t source_hash() {
    static const uint64_t h = fnv1a(kGeneigDeviceSource, std::strlen(kGeneigDeviceSource));
    return h;
}

// Toolchain fingerprint: hipRTC version plus the compile options, so a code
// object built by another compiler or with other flags is never reused.
const char* const kRtcFlags[] = {"-O3", "-std=c++17", "-Xclang", "-target-feature",
                                 "-Xclang", "-packed-fp32-ops"};

uint64_t toolchain_hash() {
    static const uint64_t h = [] {
        int major = 0, minor = 0;
        hiprtcVersion(&major, &minor);
        uint64_t x = fnv1a(&major, sizeof major);
        x = fnv1a(&minor, sizeof minor, x);
        for (const char* f : kRtcFlags) x = fnv1a(f, std::strlen(f), x);
        return x;
    }();
    return h;
}

std::string sanitize(const std::string& s) {
    std::string o = s;
    for (char& c : o)
        if (c == ':' || c == '/' || c == '+' || c == ' ') c = '_';
    return o;
}

std::string cache_dir_str() {
    if (const char* env = std::getenv("SUNNY_GENEIG_CACHE_DIR")) return env;
    if (const char* xdg = std::getenv("XDG_CACHE_HOME")) return std::string(xdg) + "/sunny_geneig/hsaco";
    if (const char* home = std::getenv("HOME")) return std::string(home) + "/.cache/sunny_geneig/hsaco";
    return "/tmp/sunny_geneig/hsaco";
}

// k1_rows: K1's row-chunk width when the autotuner set one (0: the build's
// default, and the reference's path format).
std::string cache_path_str(int n, int bpb, const std::string& arch, int bdx, int variant, int k1_rows = 0) {
    char rows[16] = "";
    if (k1_rows > 0) std::snprintf(rows, sizeof rows, "_r%d", k1_rows);
    char buf[1024];
    std::snprintf(buf, sizeof buf, "%s/%d_%d_%s_%d_v%d%s_lib%016llx_src%016llx.hsaco",
                  cache_dir_str().c_str(), n, bpb, sanitize(arch).c_str(), bdx, variant, rows,
                  (unsigned long long)toolchain_hash(), (unsigned long long)source_hash());
    return buf;
}

std::string tune_path_str(int n, const std::string& arch) {
    char buf[1024];
    std::snprintf(buf, sizeof buf, "%s/autotune_%d_%s_lib%016llx_src%016llx.txt",
                  cache_dir_str().c_str(), n, sanitize(arch).c_str(),
                  (unsigned long long)toolchain_hash(), (unsigned long long)source_hash());
    return buf;
}

void mkdir_p(const std::string& path) {
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') ::mkdir(path.substr(0, i).c_str(), 0755);
    ::mkdir(path.c_str(), 0755);
}

// Atomic write: tmp file in the same directory, then rename (multi-process
// safe, as src/nvrtc/nvrtc_solver.cpp:481-520).
bool write_atomic(const std::string& path, const void* data, size_t size) {
    mkdir_p(path.substr(0, path.find_last_of('/')));
    char tmp[1200];
    std::snprintf(tmp, sizeof tmp, "%s.tmp.%d.%lld", path.c_str(), (int)::getpid(),
                  (long long)std::chrono::steady_clock::now().time_since_epoch().count());
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return false;
        f.write(static_cast<const char*>(data), (std::streamsize)size);
        if (!f) {
            f.close();
            std::remove(tmp);
            return false;
        }
    }
    if (std::rename(tmp, path.c_str()) != 0) {
        std::remove(tmp);
        return false;
    }
    return true;
}

// ---- hipRTC ----------------------------------------------------------------
// bdx: threads per pair of K1 / K3 (64: one wave, bpb pairs per workgroup;
// 256: the four-wave register-tiled kernels, one pair per workgroup).
int rtc_compile(int n, int bpb, int variant, const std::string& arch, std::vector<char>& code, int k1_rows = 0,
                int bdx = 64) {
    if (n < 1 || n > 128) return fail(GENEIG_E_ARG, "n=%d outside the supported range 1..128", n);
    if (bpb < 1 || bpb > 16) return fail(GENEIG_E_ARG, "pairs per workgroup %d outside 1..16", bpb);
    if (variant != 0 && variant != 1) return fail(GENEIG_E_ARG, "variant %d is not built", variant);
    if (k1_rows != 0 && k1_rows != 4 && k1_rows != 8 && k1_rows != 16)
        return fail(GENEIG_E_ARG, "K1 row chunk %d is not 4, 8 or 16", k1_rows);
    if (bdx != 64 && bdx != kTileThreadsHost)
        return fail(GENEIG_E_ARG, "threads per pair %d: 64 (one wave) and %d (four waves) are built", bdx,
                    kTileThreadsHost);
    if (bdx == kTileThreadsHost && bpb != 1)
        return fail(GENEIG_E_ARG, "the four-wave kernels take one pair per workgroup (bpb=%d)", bpb);
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, kGeneigDeviceSource, "geneig_device.hip", 0, nullptr, nullptr) !=
        HIPRTC_SUCCESS)
        return fail(GENEIG_E_RTC, "hiprtcCreateProgram failed");
    std::vector<std::string> opts = {"--offload-arch=" + arch, "-DGENEIG_RTC_N=" + std::to_string(n),
                                     "-DGENEIG_RTC_BPB=" + std::to_string(bpb),
                                     "-DGENEIG_RTC_VARIANT=" + std::to_string(variant)};
    if (k1_rows > 0) opts.push_back("-DGENEIG_RTC_K1_RCH=" + std::to_string(k1_rows));
    if (bdx == kTileThreadsHost) opts.push_back("-DGENEIG_RTC_TILE=1");
    for (const char* f : kRtcFlags) opts.push_back(f);
    std::vector<const char*> p;
    for (auto& s : opts) p.push_back(s.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)p.size(), p.data());
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls + 1, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        return fail(GENEIG_E_RTC, "hipRTC compile failed (n=%d bpb=%d variant=%d bdx=%d arch=%s): %s\n%s", n,
                    bpb, variant, bdx, arch.c_str(), hiprtcGetErrorString(r), log.c_str());
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return GENEIG_E_OK;
}

// ---- loaded module -----------------------------------------------------------
struct Module {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;    // K1: factor, reduce, tridiagonalize, form Q
    hipFunction_t fn2a = nullptr;  // K2a: QL chases, 64 pairs per wave, rotations logged
    hipFunction_t fn2b = nullptr;  // K2b: Z replayed from the log
    hipFunction_t fn2x = nullptr;  // K2x: one-pair QL for a pair whose log overflowed
    hipFunction_t fn3 = nullptr;   // K3: sort, Y = Q Z, back-transform
    // in-place solve (zhegvd storage): Cholesky of B into B, K1 / K3 with L from B
    hipFunction_t fci = nullptr, f1i = nullptr, f3i = nullptr;
    unsigned lds_per_pair = 0;     // K1 dynamic LDS bytes per pair
    unsigned lds_total = 0;        // larger of K1 / K3 dynamic LDS + static vectors, per pair
    unsigned k2_threads = 0;
    unsigned k2a_lds = 0, f_k2a_lds = 0;   // K2a dynamic LDS (FP64, complex64 form)
    unsigned chol64_lds = 0;               // FP64 Cholesky stage dynamic LDS (precision form)
    unsigned k3_lds = 0, f_k3_lds = 0;     // K3 dynamic LDS per pair (FP64, complex64 form)
    unsigned long long ws_pair_bytes = 0;
    // precision form (variant 1): FP64 Cholesky, complex64 K1/K2/K3, FP64
    // refinement, FP64 K1/K2/K3 over a pair list, counter clear
    hipFunction_t fc = nullptr, f1f = nullptr, f2af = nullptr, f2bf = nullptr, f2xf = nullptr, f3f = nullptr;
    hipFunction_t fr = nullptr, f1l = nullptr, f2al = nullptr, f2bl = nullptr, f2xl = nullptr, f3l = nullptr;
    hipFunction_t fz = nullptr;            // clears device counters (both variants)
    unsigned f_lds_per_pair = 0, f_k2_threads = 0, refine_lds = 0, refine_threads = 0;
    unsigned k13_threads = 64;     // threads per pair of K1, K3 and the Cholesky stages
    unsigned f_lds_total = 0;      // larger of K1f / K3f dynamic LDS + static vectors, per pair
    unsigned long long f_ws_pair_bytes = 0;
    int n = 0, bpb = 0, variant = 0;
    Module() = default;
    Module(const Module&) = delete;
    Module& operator=(const Module&) = delete;
    ~Module() {
        if (mod) (void)hipModuleUnload(mod);
    }
};

int load_module(const std::vector<char>& code, int n, int bpb, int variant,
                std::shared_ptr<Module>& out) {
    auto m = std::make_shared<Module>();
    m->n = n;
    m->bpb = bpb;
    m->variant = variant;
    HIP_TRY(hipModuleLoadData(&m->mod, code.data()));
    if (variant == 1) {
        HIP_TRY(hipModuleGetFunction(&m->fc, m->mod, "geneig_chol64"));
        HIP_TRY(hipModuleGetFunction(&m->f1f, m->mod, "geneig_k1f"));
        HIP_TRY(hipModuleGetFunction(&m->f2af, m->mod, "geneig_k2af"));
        HIP_TRY(hipModuleGetFunction(&m->f2bf, m->mod, "geneig_k2bf"));
        HIP_TRY(hipModuleGetFunction(&m->f2xf, m->mod, "geneig_k2xf"));
        HIP_TRY(hipModuleGetFunction(&m->f3f, m->mod, "geneig_k3f"));
        HIP_TRY(hipModuleGetFunction(&m->fr, m->mod, "geneig_refine"));
        HIP_TRY(hipModuleGetFunction(&m->f1l, m->mod, "geneig_k1l"));
        HIP_TRY(hipModuleGetFunction(&m->f2al, m->mod, "geneig_k2al"));
        HIP_TRY(hipModuleGetFunction(&m->f2bl, m->mod, "geneig_k2bl"));
        HIP_TRY(hipModuleGetFunction(&m->f2xl, m->mod, "geneig_k2xl"));
        HIP_TRY(hipModuleGetFunction(&m->f3l, m->mod, "geneig_k3l"));
    } else {
        HIP_TRY(hipModuleGetFunction(&m->fn, m->mod, "geneig_k1"));
        HIP_TRY(hipModuleGetFunction(&m->fn2a, m->mod, "geneig_k2a"));
        HIP_TRY(hipModuleGetFunction(&m->fn2b, m->mod, "geneig_k2b"));
        HIP_TRY(hipModuleGetFunction(&m->fn2x, m->mod, "geneig_k2x"));
        HIP_TRY(hipModuleGetFunction(&m->fn3, m->mod, "geneig_k3"));
        HIP_TRY(hipModuleGetFunction(&m->fci, m->mod, "geneig_chol_ip"));
        HIP_TRY(hipModuleGetFunction(&m->f1i, m->mod, "geneig_k1_ip"));
        HIP_TRY(hipModuleGetFunction(&m->f3i, m->mod, "geneig_k3_ip"));
    }
    HIP_TRY(hipModuleGetFunction(&m->fz, m->mod, "geneig_zero"));
    auto read_global = [&](const char* name, void* dst, size_t want) -> int {
        hipDeviceptr_t ptr = nullptr;
        size_t size = 0;
        HIP_TRY(hipModuleGetGlobal(&ptr, &size, m->mod, name));
        if (size != want) return fail(GENEIG_E_HIP, "%s has size %zu, expected %zu", name, size, want);
        HIP_TRY(hipMemcpyDtoH(dst, ptr, size));
        return GENEIG_E_OK;
    };
    int rc = read_global("geneig_lds_bytes_per_pair", &m->lds_per_pair, sizeof(unsigned));
    if (!rc) rc = read_global("geneig_lds_total_per_pair", &m->lds_total, sizeof(unsigned));
    if (!rc) rc = read_global("geneig_k2_threads", &m->k2_threads, sizeof(unsigned));
    if (!rc) rc = read_global("geneig_k2a_lds_bytes", &m->k2a_lds, sizeof(unsigned));
    if (!rc) rc = read_global("geneig_k3_lds_bytes", &m->k3_lds, sizeof(unsigned));
    if (!rc) rc = read_global("geneig_ws_bytes_per_pair", &m->ws_pair_bytes, sizeof(unsigned long long));
    if (!rc) rc = read_global("geneig_chol64_lds_bytes", &m->chol64_lds, sizeof(unsigned));
    if (!rc) rc = read_global("geneig_k13_threads", &m->k13_threads, sizeof(unsigned));
    if (variant == 1) {
        if (!rc) rc = read_global("geneig_f_lds_bytes_per_pair", &m->f_lds_per_pair, sizeof(unsigned));
        if (!rc) rc = read_global("geneig_f_lds_total_per_pair", &m->f_lds_total, sizeof(unsigned));
        if (!rc) rc = read_global("geneig_f_k2_threads", &m->f_k2_threads, sizeof(unsigned));
        if (!rc) rc = read_global("geneig_f_k2a_lds_bytes", &m->f_k2a_lds, sizeof(unsigned));
        if (!rc) rc = read_global("geneig_f_k3_lds_bytes", &m->f_k3_lds, sizeof(unsigned));
        if (!rc) rc = read_global("geneig_f_ws_bytes_per_pair", &m->f_ws_pair_bytes, sizeof(unsigned long long));
        if (!rc) rc = read_global("geneig_refine_lds_bytes", &m->refine_lds, sizeof(unsigned));
        if (!rc) rc = read_global("geneig_refine_threads", &m->refine_threads, sizeof(unsigned));
    }
    if (rc) return rc;
    out = m;
    return GENEIG_E_OK;
}

struct Key {
    int n, bpb, variant, bdx;
    std::string arch;
    int k1_rows;
    bool operator<(const Key& o) const {
        return std::tie(n, bpb, variant, bdx, arch, k1_rows) < std::tie(o.n, o.bpb, o.variant, o.bdx, o.arch, o.k1_rows);
    }
};

std::mutex g_module_mu;
std::map<Key, std::weak_ptr<Module>> g_modules;

struct TuneWinner {
    int bpb = 0, variant = 0, k1_rows = 0, bdx = 64;
    double pairs_per_s = 0.0;
};
// One autotune candidate: pairs per workgroup, module variant, K1 row chunk
// (0: the build's default), threads per pair of K1 / K3.
struct TuneCand {
    int bpb, variant, k1_rows, bdx;
};
std::mutex g_tune_mu;
std::map<std::pair<int, std::string>, TuneWinner> g_tune;

// Synthetic tuning pairs (the reference's generators, nvrtc_solver.cpp:606-638,
// restated with a counter-based RNG): A = M + M^H, B = G G^H + n I.
uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double gauss(uint64_t key) {
    double u1 = ((double)(mix64(key) >> 11) + 0.5) * 0x1.0p-53;
    double u2 = ((double)(mix64(key ^ 0xA5A5A5A5A5A5A5A5ULL) >> 11) + 0.5) * 0x1.0p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
void make_tune_pairs(int n, int batch, std::vector<std::complex<double>>& A,
                     std::vector<std::complex<double>>& B) {
    const size_t nn = (size_t)n * n;
    A.assign(nn * batch, 0.0);
    B.assign(nn * batch, 0.0);
    std::vector<std::complex<double>> M(nn), G(nn);
    for (int b = 0; b < batch; ++b) {
        for (size_t t = 0; t < nn; ++t) {
            uint64_t k = ((uint64_t)n << 48) ^ ((uint64_t)b << 22) ^ (t << 1);
            M[t] = {gauss(4 * k), gauss(4 * k + 1)};
            G[t] = {gauss(4 * k + 2), gauss(4 * k + 3)};
        }
        auto* Ab = &A[nn * b];
        auto* Bb = &B[nn * b];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) Ab[i + (size_t)j * n] = M[i + (size_t)j * n] + std::conj(M[j + (size_t)i * n]);
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                const std::complex<double> gjk = std::conj(G[j + (size_t)k * n]);
                for (int i = 0; i < n; ++i) Bb[i + (size_t)j * n] += G[i + (size_t)k * n] * gjk;
            }
        for (int i = 0; i < n; ++i) Bb[i + (size_t)i * n] += (double)n;
    }
}

double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
}

}  // namespace

// ---- the solver object ---------------------------------------------------------
struct geneig_solver {
    int n = 0, device = 0, bpb = 0, variant = 0;
    std::string arch;
    size_t lds_per_cu = 0, lds_per_block = 0;
    std::shared_ptr<Module> module;
    // the bpb = 1 module that sized the LDS per pair, kept so that a second
    // solver for the same n finds both of its modules in-process
    std::shared_ptr<Module> probe;
    unsigned shared_bytes = 0;
    int autotuned_bdx = 0;
    int bdx = 64;                  // threads per pair of K1 / K3 (64 or 256)
    int k1_rows = 0;               // K1 row-chunk width of the module (0: the build's default)
    bool was_autotuned = false, sweep_ran = false;
    const char* layer = "compile";
    double acquire_ms = 0.0;
    // Device staging of launch_chunked_host, one slot per stream: work
    // queued on a stream is ordered, so a slot is reused only after the
    // copies of its previous chunk.  Grown on demand, freed with the solver.
    struct Staging {
        void* base = nullptr;
        size_t bytes = 0;
    };
    std::mutex stage_mu;
    std::map<hipStream_t, Staging> stage;

    ~geneig_solver() {
        if (stage.empty()) return;
        (void)hipSetDevice(device);
        for (auto& kv : stage)
            if (kv.second.base) (void)hipFree(kv.second.base);   // waits for the device
    }

    // A staging slot of at least `bytes` for `stream`.
    int staging(hipStream_t stream, size_t bytes, void** out) {
        std::lock_guard<std::mutex> lk(stage_mu);
        Staging& st = stage[stream];
        if (st.bytes < bytes) {
            if (st.base) {
                HIP_TRY(hipStreamSynchronize(stream));   // the old slot may still be in use on this stream
                (void)hipFree(st.base);
                st.base = nullptr;
                st.bytes = 0;
            }
            HIP_TRY(hipMalloc(&st.base, bytes));
            st.bytes = bytes;
        }
        *out = st.base;
        return GENEIG_E_OK;
    }

    // Module acquisition through the cache layers (nvrtc_solver.cpp:522-597).
    int acquire(int want_bpb, int want_variant, int cache_mode, std::shared_ptr<Module>& out,
                const char** layer_out, int k1_rows = 0, int bdx = 64) {
        const Key key{n, want_bpb, want_variant, bdx, arch, k1_rows};
        if (cache_mode != GENEIG_CACHE_NOCACHE) {
            std::lock_guard<std::mutex> lk(g_module_mu);
            auto it = g_modules.find(key);
            if (it != g_modules.end()) {
                if (auto h = it->second.lock()) {
                    out = h;
                    *layer_out = "in-process";
                    return GENEIG_E_OK;
                }
                g_modules.erase(it);
            }
        }
        const std::string path = cache_path_str(n, want_bpb, arch, bdx, want_variant, k1_rows);
        if (cache_mode == GENEIG_CACHE_AUTO) {
            std::ifstream f(path, std::ios::binary | std::ios::ate);
            if (f) {
                const std::streamsize sz = f.tellg();
                f.seekg(0);
                std::vector<char> code((size_t)std::max<std::streamsize>(sz, 0));
                bool ok = sz >= 1024 && (bool)f.read(code.data(), sz);
                // Code objects are ELF: 0x7F 'E' 'L' 'F' (nvrtc_solver.cpp:466-474).
                ok = ok && (unsigned char)code[0] == 0x7F && code[1] == 'E' && code[2] == 'L' &&
                     code[3] == 'F';
                f.close();
                std::shared_ptr<Module> h;
                if (ok && load_module(code, n, want_bpb, want_variant, h) == GENEIG_E_OK) {
                    std::lock_guard<std::mutex> lk(g_module_mu);
                    g_modules[key] = h;
                    out = h;
                    *layer_out = "disk";
                    return GENEIG_E_OK;
                }
                std::fprintf(stderr, "[cache] invalid code object, removing: %s\n", path.c_str());
                std::remove(path.c_str());
            }
        }
        std::vector<char> code;
        int rc = rtc_compile(n, want_bpb, want_variant, arch, code, k1_rows, bdx);
        if (rc) return rc;
        std::shared_ptr<Module> h;
        rc = load_module(code, n, want_bpb, want_variant, h);
        if (rc) return rc;
        if (cache_mode == GENEIG_CACHE_AUTO && !write_atomic(path, code.data(), code.size()))
            std::fprintf(stderr, "[cache] could not write %s\n", path.c_str());
        if (cache_mode != GENEIG_CACHE_NOCACHE) {
            std::lock_guard<std::mutex> lk(g_module_mu);
            g_modules[key] = h;
        }
        out = h;
        *layer_out = "compile";
        return GENEIG_E_OK;
    }

    bool fits(int b, unsigned per_pair) const { return (size_t)b * per_pair <= lds_per_block; }

    // Pairs per workgroup that maximizes resident pairs per CU (LDS-limited);
    // ties go to the larger workgroup (fewer workgroups to dispatch).
    int suggest_bpb(unsigned per_pair) const {
        int best = 0;
        size_t best_res = 0;
        for (int b : {16, 8, 4, 2, 1}) {
            if (!fits(b, per_pair)) continue;
            size_t res = (size_t)b * (lds_per_cu / ((size_t)b * per_pair));
            if (res > best_res) {
                best_res = res;
                best = b;
            }
        }
        return best;
    }

    // Workspace layouts of a launch of `batch` pairs (shared with reserve).
    geneig_host::PlainLayout plain_layout_of(const Module& m, int batch) const {
        return geneig_host::plain_layout(batch, m.ws_pair_bytes, 1);
    }
    geneig_host::MixedLayout mixed_layout_of(const Module& m, int batch) const {
        return geneig_host::mixed_layout(batch, m.f_ws_pair_bytes, m.ws_pair_bytes, 1);
    }
    size_t workspace_bytes(int batch) const {
        if (batch <= 0) return 0;
        return variant == 1 ? mixed_layout_of(*module, batch).bytes : plain_layout_of(*module, batch).bytes;
    }

    // K1 -> K2 -> K3 over the batch, chunked to bound the per-stream workspace.
    // inplace: LAPACK zhegvd storage (U is H: the eigenvectors overwrite H,
    // the Cholesky factor S), with the Cholesky as its own kernel.
    int launch_raw(const Module& m, int b, const void* H, const void* S, double* W, void* U, int* info,
                   int batch, hipStream_t stream, bool inplace = false) {
        if (batch <= 0) return GENEIG_E_OK;
        const geneig_host::PlainLayout L = plain_layout_of(m, batch);
        void* ws = nullptr;
        HIP_TRY(geneig_host::workspace(stream, L.bytes, &ws));
        const size_t mat = (size_t)n * n * 16;
        const unsigned bytes = m.lds_per_pair * (unsigned)b;
        for (int off = 0; off < batch; off += L.chunk) {
            int P = std::min(L.chunk, batch - off);
            const void* h = static_cast<const char*>(H) + mat * off;
            const void* sp = static_cast<const char*>(S) + mat * off;
            double* w = W + (size_t)n * off;
            void* u = static_cast<char*>(U) + mat * off;
            int* inf = info + off;
            int* ovf = geneig_host::region_ctl(ws, P, m.ws_pair_bytes);
            int one = 1;
            void* az[] = {&ovf, &one};
            HIP_TRY(hipModuleLaunchKernel(m.fz, 1, 1, 1, 64, 1, 1, 0, stream, az, nullptr));
            if (inplace) {
                void* ac[] = {&sp, &w, &u, &inf, &P};
                HIP_TRY(hipModuleLaunchKernel(m.fci, (unsigned)P, 1, 1, m.k13_threads, 1, 1, m.chol64_lds, stream, ac,
                                              nullptr));
            }
            void* a1[] = {&h, &sp, &w, &u, &inf, &ws, &P};
            // one wave per pair, b pairs per workgroup; or one pair per
            // workgroup of four waves (k13_threads = 256, b = 1)
            const unsigned g13 = (unsigned)((P + b - 1) / b);
            const unsigned bx = m.k13_threads == 64 ? 64u : m.k13_threads, by = m.k13_threads == 64 ? (unsigned)b : 1u;
            HIP_TRY(hipModuleLaunchKernel(inplace ? m.f1i : m.fn, g13, 1, 1, bx, by, 1, bytes, stream, a1, nullptr));
            int cap = geneig_host::ql_log_entries();
            void* a2a[] = {&inf, &ws, &P, &cap, &ovf};
            HIP_TRY(hipModuleLaunchKernel(m.fn2a, (unsigned)((P + 63) / 64), 1, 1, 64, 1, 1, m.k2a_lds, stream, a2a,
                                          nullptr));
            void* a2[] = {&inf, &ws, &P};
            HIP_TRY(hipModuleLaunchKernel(m.fn2b, (unsigned)P, 1, 1, m.k2_threads, 1, 1, 0, stream, a2, nullptr));
            void* a2x[] = {&inf, &ws, &P, &ovf};
            HIP_TRY(hipModuleLaunchKernel(m.fn2x, (unsigned)std::min(P, kListGridK2x), 1, 1, m.k2_threads, 1, 1, 0,
                                          stream, a2x, nullptr));
            void* a3[] = {&sp, &w, &u, &inf, &ws, &P};
            HIP_TRY(hipModuleLaunchKernel(inplace ? m.f3i : m.fn3, g13, 1, 1, bx, by, 1, m.k3_lds * (unsigned)b, stream,
                                          a3, nullptr));
        }
        return GENEIG_E_OK;
    }

    // Precision form over the batch (the sequence of geneig_mixed.hip with
    // the module's kernels): FP64 Cholesky, complex64 K1/K2/K3 per chunk,
    // FP64 refinement, then the FP64 pipeline over the pairs it lists, whose
    // number the kernels read on the device.  Waits for the device only when
    // h_fallback asks for that number.
    int launch_mixed(const Module& m, const void* H, const void* S, double* W, void* U, int* info, int batch,
                     int maxit, double tol, int* iters, int* h_fallback, hipStream_t stream) {
        if (h_fallback) *h_fallback = 0;
        if (batch <= 0) return GENEIG_E_OK;
        if (maxit <= 0) maxit = 4;
        if (!(tol > 0.0)) tol = 1e-7;
        const geneig_host::MixedLayout L = mixed_layout_of(m, batch);
        void* ws = nullptr;
        HIP_TRY(geneig_host::workspace(stream, L.bytes, &ws));
        int* fb_count = static_cast<int*>(ws);
        int* fb_list = fb_count + 1;
        void* solver = static_cast<char*>(ws) + L.list_bytes;
        const size_t mat = (size_t)n * n * 16;
        int one = 1;
        void* az[] = {&fb_count, &one};
        HIP_TRY(hipModuleLaunchKernel(m.fz, 1, 1, 1, 64, 1, 1, 0, stream, az, nullptr));
        void* ac[] = {&S, &W, &U, &info, &batch};
        const unsigned tx = m.k13_threads;
        HIP_TRY(hipModuleLaunchKernel(m.fc, (unsigned)batch, 1, 1, tx, 1, 1, m.chol64_lds, stream, ac, nullptr));
        for (int off = 0; off < batch; off += L.chunk_f) {
            int P = std::min(L.chunk_f, batch - off);
            const void* h = static_cast<const char*>(H) + mat * off;
            const void* sp = static_cast<const char*>(S) + mat * off;
            double* w = W + (size_t)n * off;
            void* u = static_cast<char*>(U) + mat * off;
            int* inf = info + off;
            int* ovf = geneig_host::region_ctl(solver, P, m.f_ws_pair_bytes);
            void* azo[] = {&ovf, &one};
            HIP_TRY(hipModuleLaunchKernel(m.fz, 1, 1, 1, 64, 1, 1, 0, stream, azo, nullptr));
            void* a1[] = {&h, &sp, &w, &u, &inf, &solver, &P};
            HIP_TRY(hipModuleLaunchKernel(m.f1f, (unsigned)P, 1, 1, tx, 1, 1, m.f_lds_per_pair, stream, a1, nullptr));
            int cap = geneig_host::ql_log_entries();
            void* a2a[] = {&inf, &solver, &P, &cap, &ovf};
            HIP_TRY(hipModuleLaunchKernel(m.f2af, (unsigned)((P + 63) / 64), 1, 1, 64, 1, 1, m.f_k2a_lds, stream, a2a,
                                          nullptr));
            void* a2[] = {&inf, &solver, &P};
            HIP_TRY(hipModuleLaunchKernel(m.f2bf, (unsigned)P, 1, 1, m.f_k2_threads, 1, 1, 0, stream, a2, nullptr));
            void* a2x[] = {&inf, &solver, &P, &ovf};
            HIP_TRY(hipModuleLaunchKernel(m.f2xf, (unsigned)std::min(P, kListGridK2x), 1, 1, m.f_k2_threads, 1, 1, 0,
                                          stream, a2x, nullptr));
            void* a3[] = {&sp, &w, &u, &inf, &solver, &P};
            HIP_TRY(hipModuleLaunchKernel(m.f3f, (unsigned)P, 1, 1, tx, 1, 1, m.f_k3_lds, stream, a3, nullptr));
        }
        void* ar[] = {&H, &S, &W, &U, &info, &fb_list, &fb_count, &iters, &batch, &maxit, &tol};
        HIP_TRY(hipModuleLaunchKernel(m.fr, (unsigned)batch, 1, 1, m.refine_threads, 1, 1, m.refine_lds, stream, ar,
                                      nullptr));
        int cd = L.chunk_d;
        int* fovf = geneig_host::region_ctl(solver, cd, m.ws_pair_bytes);
        const unsigned g1 = (unsigned)std::min(cd, kListGridK1);
        const unsigned g2a = (unsigned)std::min((cd + 63) / 64, kListGridK2a);
        const unsigned g2x = (unsigned)std::min(cd, kListGridK2x);
        for (int r = 0; r < L.rounds; ++r) {
            int off = r * cd;
            void* azf[] = {&fovf, &one};
            HIP_TRY(hipModuleLaunchKernel(m.fz, 1, 1, 1, 64, 1, 1, 0, stream, azf, nullptr));
            void* a1[] = {&H, &S, &W, &U, &info, &solver, &cd, &fb_list, &fb_count, &off};
            HIP_TRY(hipModuleLaunchKernel(m.f1l, g1, 1, 1, tx, 1, 1, m.lds_per_pair, stream, a1, nullptr));
            int cap = geneig_host::ql_log_entries();
            void* a2a[] = {&info, &solver, &cd, &fb_list, &fb_count, &off, &cap, &fovf};
            HIP_TRY(hipModuleLaunchKernel(m.f2al, g2a, 1, 1, 64, 1, 1, m.k2a_lds, stream, a2a, nullptr));
            void* a2[] = {&info, &solver, &cd, &fb_list, &fb_count, &off};
            HIP_TRY(hipModuleLaunchKernel(m.f2bl, g1, 1, 1, m.k2_threads, 1, 1, 0, stream, a2, nullptr));
            void* a2x[] = {&info, &solver, &cd, &fb_list, &off, &fovf};
            HIP_TRY(hipModuleLaunchKernel(m.f2xl, g2x, 1, 1, m.k2_threads, 1, 1, 0, stream, a2x, nullptr));
            void* a3[] = {&S, &W, &U, &info, &solver, &cd, &fb_list, &fb_count, &off};
            HIP_TRY(hipModuleLaunchKernel(m.f3l, g1, 1, 1, tx, 1, 1, m.k3_lds, stream, a3, nullptr));
        }
        if (h_fallback) {
            int count = 0;
            HIP_TRY(hipMemcpyAsync(&count, fb_count, sizeof(int), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            if (count < 0 || count > batch) return fail(GENEIG_E_HIP, "fallback count %d out of range", count);
            *h_fallback = count;
        }
        return GENEIG_E_OK;
    }

    void set_attrs(const Module& m, int b) {
        if (m.variant == 1) {
            auto set = [](hipFunction_t f, unsigned bytes) {
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)bytes);
            };
            set(m.fc, m.chol64_lds);
            set(m.f1l, m.lds_per_pair);
            set(m.f3l, m.k3_lds);
            set(m.f1f, m.f_lds_per_pair);
            set(m.f3f, m.f_k3_lds);
            set(m.fr, m.refine_lds);
            set(m.f2af, m.f_k2a_lds);
            set(m.f2al, m.k2a_lds);
            return;
        }
        const int bytes = (int)(m.lds_per_pair * (unsigned)b);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(m.fn2a), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)m.k2a_lds);
        for (hipFunction_t f : {m.fn, m.f1i})
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        for (hipFunction_t f : {m.fn3, m.f3i})
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(f), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(m.k3_lds * (unsigned)b));
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(m.fci), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)m.chol64_lds);
    }

    int autotune(int cache_mode, const std::vector<TuneCand>& cands, TuneWinner& win);
};

int geneig_solver::autotune(int cache_mode, const std::vector<TuneCand>& cands, TuneWinner& win) {
    // Workload and validation follow src/nvrtc/nvrtc_solver.cpp:649-925: 1024
    // pairs, 5-sample median timing, 8 spot pairs checked against the first
    // candidate at 1e-10 (eigenvalues scaled by max(|w|,1), and 1 - |u^H B u_ref|).
    const int T = 1024, SPOT = 8;
    const double tol = 1e-10;
    std::vector<std::complex<double>> A, B;
    make_tune_pairs(n, T, A, B);
    const size_t nn = (size_t)n * n;
    void *dA = nullptr, *dB = nullptr, *dU = nullptr;
    double* dW = nullptr;
    int* dI = nullptr;
    auto cleanup = [&] {
        (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dU); (void)hipFree(dW); (void)hipFree(dI);
    };
    if (hipMalloc(&dA, nn * T * 16) != hipSuccess || hipMalloc(&dB, nn * T * 16) != hipSuccess ||
        hipMalloc(&dU, nn * T * 16) != hipSuccess || hipMalloc((void**)&dW, (size_t)n * T * 8) != hipSuccess ||
        hipMalloc((void**)&dI, (size_t)T * 4) != hipSuccess) {
        cleanup();
        return fail(GENEIG_E_HIP, "autotune: device allocation failed");
    }
    (void)hipMemcpy(dA, A.data(), nn * T * 16, hipMemcpyHostToDevice);
    (void)hipMemcpy(dB, B.data(), nn * T * 16, hipMemcpyHostToDevice);
    std::vector<double> Wref, Wc((size_t)n * T);
    std::vector<std::complex<double>> Uref, Uc(nn * T);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    bool have = false;
    for (size_t ci = 0; ci < cands.size(); ++ci) {
        const int b = cands[ci].bpb, v = cands[ci].variant, rows = cands[ci].k1_rows, bx = cands[ci].bdx;
        std::shared_ptr<Module> m;
        const char* lay = nullptr;
        if (acquire(b, v, cache_mode, m, &lay, rows, bx) != GENEIG_E_OK) {
            std::fprintf(stderr, "[autotune] bpb=%d variant=%d k1_rows=%d bdx=%d: rejected at compile/load\n", b, v,
                         rows, bx);
            continue;
        }
        const unsigned bytes = m->lds_total * (unsigned)b;
        if (!fits(b, m->lds_total)) {
            std::fprintf(stderr, "[autotune] bpb=%d variant=%d: %u B LDS does not fit\n", b, v, bytes);
            continue;
        }
        set_attrs(*m, b);
        if (launch_raw(*m, b, dA, dB, dW, dU, dI, T, nullptr) != GENEIG_E_OK) {  // warm-up
            std::fprintf(stderr, "[autotune] bpb=%d variant=%d: launch failed\n", b, v);
            continue;
        }
        std::vector<double> samples;
        for (int s = 0; s < 5; ++s) {
            (void)hipEventRecord(e0, nullptr);
            (void)launch_raw(*m, b, dA, dB, dW, dU, dI, T, nullptr);
            (void)hipEventRecord(e1, nullptr);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            samples.push_back(T / (ms * 1e-3));
        }
        if (hipDeviceSynchronize() != hipSuccess) {
            cleanup();
            return fail(GENEIG_E_HIP, "autotune: kernel failed for bpb=%d", b);
        }
        (void)hipMemcpy(Wc.data(), dW, Wc.size() * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(Uc.data(), dU, Uc.size() * 16, hipMemcpyDeviceToHost);
        const double mps = median(samples);
        bool ok = true;
        double eig = 0.0, phase = 0.0;
        if (!have) {
            Wref = Wc;
            Uref = Uc;
            have = true;
        } else {
            for (int k = 0; k < SPOT; ++k) {
                const size_t p = (size_t)k * (T / SPOT);
                for (int i = 0; i < n; ++i) {
                    const double wr = Wref[p * n + i];
                    eig = std::max(eig, std::fabs(Wc[p * n + i] - wr) / std::max(std::fabs(wr), 1.0));
                    std::complex<double> inner = 0.0;
                    for (int r = 0; r < n; ++r) {
                        std::complex<double> bu = 0.0;
                        for (int c = 0; c < n; ++c) bu += B[nn * p + r + (size_t)c * n] * Uref[nn * p + c + (size_t)i * n];
                        inner += std::conj(Uc[nn * p + r + (size_t)i * n]) * bu;
                    }
                    phase = std::max(phase, 1.0 - std::abs(inner));
                }
            }
            ok = eig < tol && phase < tol;
        }
        std::fprintf(stderr,
                     "[autotune] n=%d bpb=%d variant=%d k1_rows=%d bdx=%d: median %.0f pairs/s eig=%.2e phase=%.2e %s\n",
                     n, b, v, rows, bx, mps, eig, phase, ok ? "[OK]" : "[REJECT correctness]");
        if (ok && mps > win.pairs_per_s) {
            win.bpb = b;
            win.variant = v;
            win.k1_rows = rows;
            win.bdx = bx;
            win.pairs_per_s = mps;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cleanup();
    if (win.bpb == 0) return fail(GENEIG_E_TUNE, "autotune: no correctness-validated candidate for n=%d", n);
    sweep_ran = true;
    return GENEIG_E_OK;
}

extern "C" {

int geneig_solver_create(geneig_solver_t* out, int n, int device_id, int bpb_request,
                         int force_block_dim_x, int tuning_mode, int cache_mode) {
    if (!out) return fail(GENEIG_E_ARG, "null output handle");
    *out = nullptr;
    if (n < 1 || n > 128) return fail(GENEIG_E_ARG, "n=%d outside the supported range 1..128", n);
    if (force_block_dim_x > 0 && force_block_dim_x != 64 && force_block_dim_x != kTileThreadsHost)
        return fail(GENEIG_E_FIT, "block_dim_x=%d: 64 (one wave per pair) and %d (four waves per pair) are built",
                    force_block_dim_x, kTileThreadsHost);
    if (force_block_dim_x == kTileThreadsHost && bpb_request > 1)
        return fail(GENEIG_E_FIT, "block_dim_x=%d takes one pair per workgroup, not %d", force_block_dim_x,
                    bpb_request);
    if (bpb_request > 16) return fail(GENEIG_E_ARG, "pairs per workgroup %d > 16", bpb_request);
    if (cache_mode < 0 || cache_mode > 2) return fail(GENEIG_E_ARG, "bad cache mode %d", cache_mode);
    std::unique_ptr<geneig_solver> s(new geneig_solver);
    s->n = n;
    s->device = device_id;
    HIP_TRY(hipSetDevice(device_id));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device_id));
    s->arch = prop.gcnArchName;
    int v = 0;
    HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device_id));
    s->lds_per_cu = (size_t)v;
    HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device_id));
    s->lds_per_block = std::max((size_t)v, (size_t)0);
    if (s->lds_per_cu < s->lds_per_block) s->lds_per_cu = s->lds_per_block;

    const auto t0 = std::chrono::steady_clock::now();
    int bpb = bpb_request, variant = 0;
    // The layer reported is the deepest one any acquisition of this
    // construction needed (a cold start reports "compile", as the reference).
    int deepest = 0;
    auto note = [&](const char* lay) {
        const int rank = (std::strcmp(lay, "compile") == 0) ? 2 : (std::strcmp(lay, "disk") == 0) ? 1 : 0;
        if (rank > deepest) deepest = rank;
    };
    std::shared_ptr<Module> keep;   // probe module, kept loaded for the final acquisition
    // Threads per pair of K1 / K3 unless the caller or the autotuner picks:
    // four waves above n = 64 (the one-wave kernels keep the matrix in LDS
    // there, and do not fit it from n = 100), one wave up to it.
    int bdx = force_block_dim_x > 0 ? force_block_dim_x : (n > 64 ? kTileThreadsHost : 64);
    if (tuning_mode == GENEIG_TUNING_AUTOTUNE && force_block_dim_x <= 0) {
        TuneWinner win;
        bool hit = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune.find({n, s->arch});
            if (it != g_tune.end()) { win = it->second; hit = true; }
        }
        if (!hit && cache_mode == GENEIG_CACHE_AUTO) {
            std::ifstream f(tune_path_str(n, s->arch));
            if (f >> win.bpb >> win.variant >> win.pairs_per_s >> win.k1_rows >> win.bdx && win.bpb >= 1 &&
                win.bpb <= 16 && (win.k1_rows == 0 || win.k1_rows == 4 || win.k1_rows == 8 || win.k1_rows == 16) &&
                (win.bdx == 64 || (win.bdx == kTileThreadsHost && win.bpb == 1)))
                hit = true;
            else
                win = TuneWinner();
        }
        if (bpb_request > 0 && hit && win.bpb != bpb_request) hit = false;
        if (!hit) {
            // stage 1: threads per pair (the reference's knob) and, for one
            // wave per pair, pairs per workgroup, the one-wave candidates
            // sized by a probe bpb = 1 module (n < 100: above, their LDS
            // matrix does not fit); stage 2: the K1 row-chunk width (the
            // tile shape of its register walks) at the winning pairs per
            // workgroup, the stage-1 winner first as the correctness
            // reference (one wave per pair, n <= 64: the walks exist only
            // there)
            std::vector<TuneCand> cands;
            if (n < 100) {
                std::shared_ptr<Module> m1;
                const char* lay = nullptr;
                int rc = s->acquire(1, 0, cache_mode, m1, &lay);
                if (rc) return rc;
                note(lay);
                keep = m1;
                for (int b : {1, 2, 4, 8})
                    if ((bpb_request <= 0 || b == bpb_request) && s->fits(b, m1->lds_total))
                        cands.push_back({b, 0, 0, 64});
            }
            if (bpb_request <= 1) cands.push_back({1, 0, 0, kTileThreadsHost});
            if (cands.empty()) return fail(GENEIG_E_FIT, "n=%d: no pairs-per-workgroup candidate fits", n);
            int rc = s->autotune(cache_mode, cands, win);
            if (rc) return rc;
            if (n <= 64 && win.bdx == 64) {
                const TuneWinner first = win;
                std::vector<TuneCand> tiles = {{first.bpb, 0, 0, 64}, {first.bpb, 0, 4, 64}, {first.bpb, 0, 16, 64}};
                win = TuneWinner();
                rc = s->autotune(cache_mode, tiles, win);
                if (rc) return rc;
            }
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[{n, s->arch}] = win;
            if (cache_mode == GENEIG_CACHE_AUTO) {
                char line[128];
                int len = std::snprintf(line, sizeof line, "%d %d %.1f %d %d\n", win.bpb, win.variant,
                                        win.pairs_per_s, win.k1_rows, win.bdx);
                write_atomic(tune_path_str(n, s->arch), line, (size_t)len);
            }
        }
        bpb = win.bpb;
        variant = win.variant;
        s->k1_rows = win.k1_rows;
        s->was_autotuned = true;
        s->autotuned_bdx = win.bdx;
        bdx = win.bdx;
    }
    if (bpb <= 0 && bdx == kTileThreadsHost) bpb = 1;
    if (bpb <= 0) {
        std::shared_ptr<Module> m1;
        const char* lay = nullptr;
        int rc = s->acquire(1, 0, cache_mode, m1, &lay);
        if (rc) return rc;
        note(lay);
        keep = m1;
        bpb = s->suggest_bpb(m1->lds_total);
        if (bpb == 0)
            return fail(GENEIG_E_FIT, "n=%d needs %u B LDS per pair, device allows %zu per workgroup", n,
                        m1->lds_total, s->lds_per_block);
    }
    const char* lay = "compile";
    int rc = s->acquire(bpb, variant, cache_mode, s->module, &lay, s->k1_rows, bdx);
    if (rc) return rc;
    note(lay);
    s->probe = std::move(keep);
    s->bdx = bdx;
    s->acquire_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    s->layer = deepest == 2 ? "compile" : deepest == 1 ? "disk" : "in-process";
    s->bpb = bpb;
    s->variant = variant;
    s->shared_bytes = s->module->lds_total * (unsigned)bpb;
    if (!s->fits(bpb, s->module->lds_total))
        return fail(GENEIG_E_FIT, "bpb=%d needs %u B LDS > %zu per workgroup", bpb, s->shared_bytes,
                    s->lds_per_block);
    s->set_attrs(*s->module, bpb);
    *out = s.release();
    return GENEIG_E_OK;
}

void geneig_solver_destroy(geneig_solver_t s) { delete s; }

int geneig_solver_launch(geneig_solver_t s, const void* d_H, const void* d_S, double* d_W, void* d_U,
                         int* d_info, int batch_size, hipStream_t stream) {
    if (!s) return fail(GENEIG_E_ARG, "null solver");
    if (s->variant == 1)
        return s->launch_mixed(*s->module, d_H, d_S, d_W, d_U, d_info, batch_size, 0, 0.0, nullptr, nullptr, stream);
    return s->launch_raw(*s->module, s->bpb, d_H, d_S, d_W, d_U, d_info, batch_size, stream);
}

int geneig_solver_launch_inplace(geneig_solver_t s, void* d_H, void* d_S, double* d_W, int* d_info, int batch_size,
                                 hipStream_t stream) {
    if (!s) return fail(GENEIG_E_ARG, "null solver");
    if (s->variant == 1)
        return fail(GENEIG_E_ARG, "launch_inplace: the precision form refines against the original H and S, "
                                  "which an in-place solve overwrites; use a solver of the FP64 form");
    return s->launch_raw(*s->module, s->bpb, d_H, d_S, d_W, d_H, d_info, batch_size, stream, true);
}

int geneig_mixed_solver_create(geneig_solver_t* out, int n, int device_id, int cache_mode) {
    return geneig_mixed_solver_create_bdx(out, n, device_id, 0, cache_mode);
}

int geneig_mixed_solver_create_bdx(geneig_solver_t* out, int n, int device_id, int block_dim_x, int cache_mode) {
    if (!out) return fail(GENEIG_E_ARG, "null output handle");
    *out = nullptr;
    if (n < 1 || n > 128) return fail(GENEIG_E_ARG, "n=%d outside the supported range 1..128", n);
    if (block_dim_x > 0 && block_dim_x != 64 && block_dim_x != kTileThreadsHost)
        return fail(GENEIG_E_FIT, "block_dim_x=%d: 64 (one wave per pair) and %d (four waves per pair) are built",
                    block_dim_x, kTileThreadsHost);
    const int bdx = block_dim_x > 0 ? block_dim_x : (n > 64 ? kTileThreadsHost : 64);
    if (cache_mode < 0 || cache_mode > 2) return fail(GENEIG_E_ARG, "bad cache mode %d", cache_mode);
    std::unique_ptr<geneig_solver> s(new geneig_solver);
    s->n = n;
    s->device = device_id;
    HIP_TRY(hipSetDevice(device_id));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device_id));
    s->arch = prop.gcnArchName;
    int v = 0;
    HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device_id));
    s->lds_per_block = (size_t)v;
    const auto t0 = std::chrono::steady_clock::now();
    const char* lay = "compile";
    const int rc = s->acquire(1, 1, cache_mode, s->module, &lay, 0, bdx);
    if (rc) return rc;
    s->acquire_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    s->layer = lay;
    s->bpb = 1;
    s->variant = 1;
    s->bdx = bdx;
    const Module& m = *s->module;
    // every kernel of the launch must fit, static broadcast vectors included:
    // the FP64 ones re-solve the pairs the refinement hands back
    s->shared_bytes = std::max(m.lds_total, std::max(m.f_lds_total, m.refine_lds));
    if (s->shared_bytes > s->lds_per_block)
        return fail(GENEIG_E_FIT, "n=%d: precision form needs %u B LDS per workgroup > %zu", n, s->shared_bytes,
                    s->lds_per_block);
    s->set_attrs(m, 1);
    *out = s.release();
    return GENEIG_E_OK;
}

int geneig_mixed_solver_launch(geneig_solver_t s, const void* d_H, const void* d_S, double* d_W, void* d_U,
                               int* d_info, int batch_size, int max_iter, double tol, int* d_iters,
                               int* h_fallback, hipStream_t stream) {
    if (!s) return fail(GENEIG_E_ARG, "null solver");
    if (s->variant != 1) return fail(GENEIG_E_ARG, "solver n=%d was not created for the precision form", s->n);
    return s->launch_mixed(*s->module, d_H, d_S, d_W, d_U, d_info, batch_size, max_iter, tol, d_iters, h_fallback,
                           stream);
}

int geneig_solver_launch_chunked(geneig_solver_t s, const void* d_H, const void* d_S, double* d_W,
                                 void* d_U, int* d_info, int total, int chunk, int nstreams,
                                 const hipStream_t* streams) {
    if (!s) return fail(GENEIG_E_ARG, "null solver");
    if (total <= 0) return GENEIG_E_OK;
    if (chunk <= 0) return fail(GENEIG_E_ARG, "launch_chunked: chunk_size must be > 0");
    if (nstreams <= 0 || !streams) return fail(GENEIG_E_ARG, "launch_chunked: need num_streams > 0 streams");
    const size_t mat = (size_t)s->n * s->n * 16;
    int ci = 0;
    for (int off = 0; off < total; off += chunk, ++ci) {
        const int bs = std::min(chunk, total - off);
        int rc = geneig_solver_launch(s, (const char*)d_H + mat * off, (const char*)d_S + mat * off,
                                      d_W + (size_t)s->n * off, (char*)d_U + mat * off, d_info + off, bs,
                                      streams[ci % nstreams]);
        if (rc) return rc;
    }
    return GENEIG_E_OK;
}

int geneig_solver_launch_chunked_host(geneig_solver_t s, const void* h_H, const void* h_S, double* h_W,
                                      void* h_U, int* h_info, int total, int chunk, int nstreams,
                                      const hipStream_t* streams) {
    if (!s) return fail(GENEIG_E_ARG, "null solver");
    if (total <= 0) return GENEIG_E_OK;
    if (!h_H || !h_S || !h_W || !h_U || !h_info) return fail(GENEIG_E_ARG, "launch_chunked_host: null host buffer");
    if (chunk <= 0) return fail(GENEIG_E_ARG, "launch_chunked_host: chunk_size must be > 0");
    if (nstreams <= 0 || !streams) return fail(GENEIG_E_ARG, "launch_chunked_host: need num_streams > 0 streams");
    const size_t n = (size_t)s->n, mat = n * n * 16;
    const size_t cp = (size_t)std::min(chunk, total);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t oS = up(mat * cp), oU = oS + up(mat * cp), oW = oU + up(mat * cp), oI = oW + up(8 * n * cp);
    const size_t bytes = oI + up(4 * cp);
    int ci = 0;
    for (int off = 0; off < total; off += chunk, ++ci) {
        const int bs = std::min(chunk, total - off);
        const hipStream_t st = streams[ci % nstreams];
        void* base = nullptr;
        int rc = s->staging(st, bytes, &base);
        if (rc) return rc;
        char* b = static_cast<char*>(base);
        const size_t m = mat * (size_t)bs;
        HIP_TRY(hipMemcpyAsync(b, (const char*)h_H + mat * off, m, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b + oS, (const char*)h_S + mat * off, m, hipMemcpyHostToDevice, st));
        rc = geneig_solver_launch(s, b, b + oS, reinterpret_cast<double*>(b + oW), b + oU,
                                  reinterpret_cast<int*>(b + oI), bs, st);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(h_W + n * off, b + oW, 8 * n * bs, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync((char*)h_U + mat * off, b + oU, m, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(h_info + off, b + oI, 4 * (size_t)bs, hipMemcpyDeviceToHost, st));
    }
    return GENEIG_E_OK;
}

int geneig_solver_reserve(geneig_solver_t s, int batch_size, hipStream_t stream) {
    if (!s) return fail(GENEIG_E_ARG, "null solver");
    if (batch_size <= 0) return GENEIG_E_OK;
    void* ws = nullptr;
    HIP_TRY(geneig_host::workspace(stream, s->workspace_bytes(batch_size), &ws));
    return GENEIG_E_OK;
}

size_t geneig_solver_workspace_bytes(geneig_solver_t s, int batch_size) {
    return s ? s->workspace_bytes(batch_size) : 0;
}

int geneig_solver_matrix_size(geneig_solver_t s) { return s ? s->n : 0; }
int geneig_solver_batches_per_block(geneig_solver_t s) { return s ? s->bpb : 0; }
void geneig_solver_block_dim(geneig_solver_t s, unsigned* x, unsigned* y, unsigned* z) {
    if (x) *x = s ? (unsigned)s->bdx : 0u;
    if (y) *y = s ? (unsigned)s->bpb : 0u;
    if (z) *z = s ? 1u : 0u;
}
unsigned geneig_solver_shared_mem_bytes(geneig_solver_t s) { return s ? s->shared_bytes : 0; }
int geneig_solver_autotuned_block_dim_x(geneig_solver_t s) { return s ? s->autotuned_bdx : 0; }
int geneig_solver_was_autotuned(geneig_solver_t s) { return s ? (int)s->was_autotuned : 0; }
int geneig_solver_autotune_sweep_ran(geneig_solver_t s) { return s ? (int)s->sweep_ran : 0; }
const char* geneig_solver_cache_layer_used(geneig_solver_t s) { return s ? s->layer : ""; }
double geneig_solver_acquire_module_ms(geneig_solver_t s) { return s ? s->acquire_ms : 0.0; }
int geneig_solver_variant(geneig_solver_t s) { return s ? s->variant : -1; }
int geneig_solver_k1_chunk_rows(geneig_solver_t s) {
    // above n = 64 the one-wave K1 works in LDS, and the four-wave K1 has no
    // register walks: neither has a row chunk
    if (!s || s->n > 64 || s->bdx != 64) return 0;
    if (s->k1_rows > 0) return s->k1_rows;
    return s->variant == 1 ? 16 : 8;   // the defaults of k1_rch<T>(): FP64 walks 8 rows, complex64 16
}

size_t geneig_cache_dir(char* buf, size_t len) {
    const std::string d = cache_dir_str();
    if (buf && len) std::snprintf(buf, len, "%s", d.c_str());
    return d.size();
}

size_t geneig_cache_path(int n, int bpb, const char* arch, int block_dim_x, int variant, char* buf,
                         size_t len) {
    const std::string p = cache_path_str(n, bpb, arch ? arch : "", block_dim_x, variant);
    if (buf && len) std::snprintf(buf, len, "%s", p.c_str());
    return p.size();
}

int geneig_rtc_compile(int n, int bpb, int variant, const char* arch, void* buf, size_t cap, size_t* size) {
    return geneig_rtc_compile_rows(n, bpb, variant, 0, arch, buf, cap, size);
}

int geneig_rtc_compile_rows(int n, int bpb, int variant, int k1_rows, const char* arch, void* buf, size_t cap,
                            size_t* size) {
    return geneig_rtc_compile_bdx(n, bpb, variant, k1_rows, 64, arch, buf, cap, size);
}

int geneig_rtc_compile_bdx(int n, int bpb, int variant, int k1_rows, int block_dim_x, const char* arch, void* buf,
                           size_t cap, size_t* size) {
    std::vector<char> code;
    int rc = rtc_compile(n, bpb, variant, arch ? arch : "gfx950", code, k1_rows, block_dim_x);
    if (rc) return rc;
    if (size) *size = code.size();
    if (buf && cap >= code.size()) std::memcpy(buf, code.data(), code.size());
    return GENEIG_E_OK;
}

uint64_t geneig_source_hash(void) { return source_hash(); }
uint64_t geneig_fnv1a64(const void* data, size_t len, uint64_t seed) {
    return fnv1a(data, len, seed ? seed : kFnvOffset);
}

int geneig_make_streams(int n, hipStream_t* out) {
    for (int i = 0; i < n; ++i) HIP_TRY(hipStreamCreateWithFlags(&out[i], hipStreamNonBlocking));
    return GENEIG_E_OK;
}
void geneig_destroy_streams(int n, hipStream_t* streams) {
    for (int i = 0; i < n; ++i)
        if (streams[i]) (void)hipStreamDestroy(streams[i]);
}

const char* geneig_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
