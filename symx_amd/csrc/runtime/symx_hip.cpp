// Host side of the C ABI declared in include/symx_hip.h.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <thread>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "internal.h"
#include "sha256.h"

using namespace symx_hip;

// The device prelude restates these structs; both sides must agree byte for byte.
static_assert(sizeof(symx_hip_eval_params) == 1088, "symx_hip_eval_params layout");
static_assert(sizeof(symx_hip_adjoint_params) == 1040, "symx_hip_adjoint_params layout");

namespace
{
    thread_local std::string g_thread_error;

    int fail(symx_hip_ctx ctx, int code, const std::string& msg)
    {
        if (ctx) ctx->err = msg;
        g_thread_error = msg;
        return code;
    }

#define HIP_TRY(ctx, expr)                                                                                              \
    do {                                                                                                                \
        hipError_t e_ = (expr);                                                                                         \
        if (e_ != hipSuccess) return fail((ctx), SYMX_HIP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

    template <typename T>
    int dalloc(symx_hip_ctx ctx, T** p, size_t count)
    {
        *p = nullptr;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)p, count * sizeof(T));
        if (e != hipSuccess) {
            *p = nullptr;
            return fail(ctx, SYMX_HIP_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) + " bytes: " + hipGetErrorString(e));
        }
        return SYMX_HIP_OK;
    }
    template <typename T>
    void dfree(T*& p)
    {
        if (p) (void)hipFree((void*)p);
        p = nullptr;
    }

    int ensure_partials(symx_hip_ctx ctx, size_t n)
    {
        if (n <= ctx->partials_cap) return SYMX_HIP_OK;
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);  // queued reductions may still use the old buffer
        dfree(ctx->d_partials);
        const size_t cap = std::max<size_t>(n, 4096);
        int rc = dalloc(ctx, &ctx->d_partials, cap);
        if (rc) return rc;
        ctx->partials_cap = cap;
        return SYMX_HIP_OK;
    }

    const char* kPrelude = R"PRELUDE(// ---- symx_hip element-kernel prelude ----
typedef int symx_i32;
typedef unsigned char symx_u8;
#define SYMX_HIP_MAX_MAPS 128
struct symx_hip_eval_params {
    const symx_i32* conn;
    const symx_u8* active;
    const double* maps[SYMX_HIP_MAX_MAPS];
    double* E_partial;
    double* G;
    double* H;
    symx_u8* C_out;
    double* E_elem;
    symx_i32 n_elem;
    symx_i32 e_mode;  // 0 sum, 1 sum and store E_elem[e], 2 sum of E_e - E_elem[e]
};
static_assert(sizeof(symx_hip_eval_params) == 1088, "symx_hip_eval_params layout");
struct symx_hip_adjoint_params {
    const double* lambda[SYMX_HIP_MAX_MAPS];
    double* out;
    symx_i32 n_out;
    symx_i32 pad_;
};
static_assert(sizeof(symx_hip_adjoint_params) == 1040, "symx_hip_adjoint_params layout");
__device__ __forceinline__ double symx_ln(double x) { return (x <= 0.0) ? -__builtin_inf() : log(x); }
__device__ __forceinline__ double symx_log10(double x) { return (x <= 0.0) ? -__builtin_inf() : log10(x); }
__device__ __forceinline__ float symx_lnf(float x) { return (x <= 0.0f) ? -__builtin_inff() : logf(x); }
__device__ __forceinline__ float symx_log10f(float x) { return (x <= 0.0f) ? -__builtin_inff() : log10f(x); }
__device__ __forceinline__ float symx_powif(float x, int n)
{
    const bool neg = n < 0;
    unsigned int m = neg ? (unsigned int)(-n) : (unsigned int)n;
    float r = 1.0f, b = x;
    while (m) { if (m & 1u) r *= b; b *= b; m >>= 1; }
    return neg ? 1.0f / r : r;
}
__device__ __forceinline__ double symx_powi(double x, int n)
{
    const bool neg = n < 0;
    unsigned int m = neg ? (unsigned int)(-n) : (unsigned int)n;
    double r = 1.0, b = x;
    while (m) { if (m & 1u) r *= b; b *= b; m >>= 1; }
    return neg ? 1.0 / r : r;
}
// Logical workgroup index: the hardware dispatches workgroups round-robin over the
// 8 XCDs; each XCD gets a contiguous range of elements (shared vertices stay in its L2).
__device__ __forceinline__ int symx_block()
{
    const int b = blockIdx.x, nb = gridDim.x;
    const int x = b % 8, i = b / 8, per = nb / 8, extra = nb % 8;
    return x * per + (x < extra ? x : extra) + i;
}
// Energy of element e as it enters the partial sum: stored (e_mode 1) or differenced
// against the stored value (e_mode 2).
__device__ __forceinline__ double symx_element_energy(double v, const symx_hip_eval_params& p, int e)
{
    if (e < p.n_elem) {
        if (p.e_mode == 1) p.E_elem[e] = v;
        else if (p.e_mode == 2) v -= p.E_elem[e];
    }
    return v;
}
// Fixed-order workgroup sum of the per-element energies (256 threads, one element each).
__device__ __forceinline__ void symx_store_energy(double v, const symx_hip_eval_params& p, int e)
{
    __shared__ double lds[4];
    v = symx_element_energy(v, p, e);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) p.E_partial[symx_block()] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
}
// Element-major rows dst[e * W + i] of the 256 elements of this workgroup, written with
// coalesced stores: each wave transposes its 64 rows through LDS in chunks of 16 values, so
// one store instruction writes 64 consecutive doubles of a chunk instead of one double in
// each of 64 rows. e0: first element of the wave; n_valid: its rows below n_elem. Every
// thread of the workgroup must call it (it synchronises the workgroup).
__device__ __forceinline__ double* symx_stage_rows_buffer()
{
    __shared__ double stage[4 * 64 * 17];  // one buffer for all row widths of a kernel
    return stage;
}
template <int W>
__device__ __forceinline__ void symx_store_rows(double* dst, const double (&v)[W], long long e0, int n_valid)
{
    const int lane = threadIdx.x & 63;
    double* st = symx_stage_rows_buffer() + (threadIdx.x >> 6) * (64 * 17);
#pragma unroll
    for (int c0 = 0; c0 < W; c0 += 16) {
        const int cw = (W - c0) < 16 ? (W - c0) : 16;
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (j < cw) st[lane * 17 + j] = v[c0 + j];
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 16; t++) {
            if (t < cw) {
                const int idx = t * 64 + lane, r = idx / cw, j = idx - r * cw;
                if (r < n_valid) dst[(e0 + r) * W + c0 + j] = st[r * 17 + j];
            }
        }
        __syncthreads();
    }
}
// The same for rows stored as floats (FP32 element-Hessian storage): values are rounded
// once, on the way into LDS, and a wave moves 32 values per row per chunk through its own
// part of the same staging buffer (64 rows x 33 floats).
template <int W, typename V>
__device__ __forceinline__ void symx_store_rows_f32(float* dst, const V (&v)[W], long long e0, int n_valid)
{
    const int lane = threadIdx.x & 63;
    float* st = reinterpret_cast<float*>(symx_stage_rows_buffer()) + (threadIdx.x >> 6) * (64 * 34);
#pragma unroll
    for (int c0 = 0; c0 < W; c0 += 32) {
        const int cw = (W - c0) < 32 ? (W - c0) : 32;
#pragma unroll
        for (int j = 0; j < 32; j++)
            if (j < cw) st[lane * 33 + j] = (float)v[c0 + j];
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 32; t++) {
            if (t < cw) {
                const int idx = t * 64 + lane, r = idx / cw, j = idx - r * cw;
                if (r < n_valid) dst[(e0 + r) * W + c0 + j] = st[r * 33 + j];
            }
        }
        __syncthreads();
    }
}
// Chunked output staging of single-point kernels: each output of the chunk [C0, C0 + CW)
// of a W-wide element-major row is put into the wave's part of the staging buffer as soon
// as it is computed (symx_put / symx_put_f32, lane-major with stride CW + 1), and
// symx_flush_rows writes the chunk with coalesced stores, so no output waits in a register
// for the end of the kernel. Every thread of the workgroup must call the flush.
__device__ __forceinline__ double* symx_stage_wave()
{
    return symx_stage_rows_buffer() + (threadIdx.x >> 6) * (64 * 17);
}
__device__ __forceinline__ void symx_put(double* st, int j, double v)
{
    st[(threadIdx.x & 63) * 17 + j] = v;
}
__device__ __forceinline__ void symx_put_f32(double* st, int j, double v)
{
    reinterpret_cast<float*>(st)[(threadIdx.x & 63) * 33 + j] = (float)v;
}
__device__ __forceinline__ void symx_put_f32(double* st, int j, float v)
{
    reinterpret_cast<float*>(st)[(threadIdx.x & 63) * 33 + j] = v;
}
// Lane l stores value t * 64 + l of the chunk, i.e. row r = q0 + t*64/CW (+1 when the
// column wraps) and column j of it; rows at or past n_valid are not stored when GUARD. When
// CW divides 64 every store instruction covers 64/CW whole rows and the offsets advance by
// compile-time steps.
template <int W, int CW, int SS, bool GUARD, typename T>
__device__ __forceinline__ void symx_flush_store(T* base, const T* st, int n_valid)
{
    const int lane = threadIdx.x & 63;
    const int q0 = lane / CW, j0 = lane - q0 * CW;
    if constexpr (64 % CW == 0) {
        constexpr int RPT = 64 / CW;
        const int go = q0 * W + j0, lo = q0 * SS + j0;
#pragma unroll
        for (int t = 0; t < CW; t++)
            if (!GUARD || q0 + RPT * t < n_valid) base[go + RPT * W * t] = st[lo + RPT * SS * t];
    }
    else {
#pragma unroll
        for (int t = 0; t < CW; t++) {
            const int a = (t * 64) / CW, b = (t * 64) % CW;
            const int jj = j0 + b;
            const bool wrap = jj >= CW;
            const int r = q0 + a + (wrap ? 1 : 0), j = wrap ? jj - CW : jj;
            if (!GUARD || r < n_valid) base[r * W + j] = st[r * SS + j];
        }
    }
}
// e0 is the wave's first element (wave-uniform); rows of inactive elements are zeroed in
// the staging first.
template <int W, int C0, int CW, typename T>
__device__ __forceinline__ void symx_flush_rows(T* dst, double* st_d, long long e0, int n_valid, bool act)
{
    constexpr int SS = sizeof(T) == 4 ? 33 : 17;
    T* st = reinterpret_cast<T*>(st_d);
    if (!act) {
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int j = 0; j < CW; j++) st[lane * SS + j] = (T)0;
    }
    __syncthreads();
    T* const base = dst + (e0 * W + C0);
    if (n_valid == 64) symx_flush_store<W, CW, SS, false>(base, st, n_valid);
    else symx_flush_store<W, CW, SS, true>(base, st, n_valid);
    __syncthreads();
}
// Zero rows [e0, e0 + n_valid) of width W (a workgroup with no active element).
template <int W, typename T>
__device__ __forceinline__ void symx_zero_rows(T* dst, long long e0, int n_valid)
{
    const int lane = threadIdx.x & 63;
    for (int i = lane; i < n_valid * W; i += 64) dst[e0 * W + i] = (T)0;
}
// ---- end of prelude ----
)PRELUDE";

    std::string hiprtc_log(hiprtcProgram p)
    {
        size_t n = 0;
        hiprtcGetProgramLogSize(p, &n);
        std::string log(n + 1, '\0');
        if (n) hiprtcGetProgramLog(p, &log[0]);
        return log;
    }

    bool read_file(const std::string& path, std::vector<char>& out)
    {
        std::ifstream f(path, std::ios::binary);
        if (!f) return false;
        out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        return !out.empty();
    }

    void mkdirs(const std::string& dir)
    {
        std::string cur;
        for (size_t i = 0; i < dir.size(); i++) {
            cur += dir[i];
            if (dir[i] == '/' || i + 1 == dir.size()) mkdir(cur.c_str(), 0755);
        }
    }

    std::vector<std::string> jit_options(const std::string& arch)
    {
        return { "--offload-arch=" + arch, "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops" };
    }

    // Cached code objects: "SYMXHSA1", the code size as 16 hex digits, the SHA-256 of the
    // code (64 hex digits), then the code. A file whose header does not match its contents
    // (truncated, foreign or corrupted) is not used; the source is compiled again.
    constexpr char kCacheMagic[] = "SYMXHSA1";
    constexpr size_t kCacheHeader = 8 + 16 + 64;

    bool read_cached_object(const std::string& path, std::vector<char>& code)
    {
        std::vector<char> raw;
        if (!read_file(path, raw) || raw.size() < kCacheHeader || std::memcmp(raw.data(), kCacheMagic, 8) != 0) return false;
        const std::string size_hex(raw.data() + 8, 16), sha_hex(raw.data() + 24, 64);
        char* end = nullptr;
        const unsigned long long n = std::strtoull(size_hex.c_str(), &end, 16);
        if (end != size_hex.c_str() + 16 || n != raw.size() - kCacheHeader || n == 0) return false;
        Sha256 sha;
        sha.update(raw.data() + kCacheHeader, (size_t)n);
        if (sha.hex() != sha_hex) return false;
        code.assign(raw.begin() + kCacheHeader, raw.end());
        return true;
    }

    void write_cached_object(const std::string& cache_dir, const std::string& path, const std::vector<char>& code)
    {
        mkdirs(cache_dir);
        char size_hex[17];
        std::snprintf(size_hex, sizeof(size_hex), "%016llx", (unsigned long long)code.size());
        Sha256 sha;
        sha.update(code.data(), code.size());
        const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
        std::ofstream f(tmp, std::ios::binary);
        f.write(kCacheMagic, 8);
        f.write(size_hex, 16);
        f.write(sha.hex().data(), 64);
        f.write(code.data(), (std::streamsize)code.size());
        f.close();
        if (f) std::rename(tmp.c_str(), path.c_str());
        else std::remove(tmp.c_str());
    }

    int jit_compile_impl(symx_hip_ctx ctx, const char* source, const char* arch_c, const char* cache_dir, std::vector<char>& code,
                         std::string& key, int& was_cached)
    {
        const std::string arch = arch_c ? arch_c : "gfx950";
        const std::vector<std::string> opts = jit_options(arch);
        // key: the source, the compiler (hiprtc version), the target and the options
        int major = 0, minor = 0;
        hiprtcVersion(&major, &minor);
        Sha256 sha;
        sha.update(source, std::strlen(source));
        sha.update("hiprtc " + std::to_string(major) + "." + std::to_string(minor) + " " + arch);
        for (const auto& o : opts) sha.update(o);
        key = sha.hex();
        was_cached = 0;
        std::string path;
        if (cache_dir && cache_dir[0]) {
            path = std::string(cache_dir) + "/" + key + ".hsaco";
            if (read_cached_object(path, code)) {
                was_cached = 1;
                return SYMX_HIP_OK;
            }
        }
        hiprtcProgram prog;
        if (hiprtcCreateProgram(&prog, source, "symx_generated.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
            return fail(ctx, SYMX_HIP_ERR_JIT, "hiprtcCreateProgram failed");
        std::vector<const char*> copts;
        for (const auto& o : opts) copts.push_back(o.c_str());
        hiprtcResult r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
        if (r != HIPRTC_SUCCESS) {
            std::string log = hiprtc_log(prog);
            hiprtcDestroyProgram(&prog);
            return fail(ctx, SYMX_HIP_ERR_JIT, std::string("hiprtc compilation failed: ") + hiprtcGetErrorString(r) + "\n" + log);
        }
        size_t n = 0;
        hiprtcGetCodeSize(prog, &n);
        code.resize(n);
        hiprtcGetCode(prog, code.data());
        hiprtcDestroyProgram(&prog);
        if (!path.empty()) {
            write_cached_object(cache_dir, path, code);
            // the generated source beside its object, as SymX keeps <name>.cpp in its codegen
            // folder (Compilation.cpp:329); it is not read back
            std::ofstream src(std::string(cache_dir) + "/" + key + ".hip");
            src << source;
        }
        return SYMX_HIP_OK;
    }

    double* vec_ptr(symx_hip_problem p, int32_t v)
    {
        if (v < 0 || v > 4) return nullptr;
        return p->d_vec[v];
    }

    int check_problem(symx_hip_problem p)
    {
        if (!p) return fail(nullptr, SYMX_HIP_ERR_ARG, "null problem");
        return SYMX_HIP_OK;
    }

    int sync_scalar(symx_hip_ctx ctx, const double* d, double* out)
    {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        *out = ctx->h_pinned[0];
        return SYMX_HIP_OK;
    }
}

extern "C" {

int symx_hip_device_count(int* n)
{
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    if (n) *n = c;
    return SYMX_HIP_OK;
}

int symx_hip_ctx_create(int device, symx_hip_ctx* out)
{
    if (!out) return fail(nullptr, SYMX_HIP_ERR_ARG, "null output pointer");
    *out = nullptr;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0) return fail(nullptr, SYMX_HIP_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= c) return fail(nullptr, SYMX_HIP_ERR_ARG, "device index out of range");
    auto* ctx = new symx_hip_ctx_s();
    ctx->device = device;
    HIP_TRY(ctx, hipSetDevice(device));
    HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    int rc = dalloc(ctx, &ctx->d_counters, 4);
    if (!rc) rc = dalloc(ctx, &ctx->d_scalars, 64);
    if (!rc) rc = dalloc(ctx, &ctx->d_pcg, 1);
    if (!rc) rc = ensure_partials(ctx, 16384);
    if (rc) {
        symx_hip_ctx_destroy(ctx);
        return rc;
    }
    HIP_TRY(ctx, hipMemset(ctx->d_counters, 0, 4 * sizeof(unsigned int)));
    if (loop_scratch_address(device, &ctx->d_loop_scratch) != 0) {
        symx_hip_ctx_destroy(ctx);
        return fail(nullptr, SYMX_HIP_ERR_HIP, "hipGetSymbolAddress of the loop scratch array failed");
    }
    HIP_TRY(ctx, hipHostMalloc((void**)&ctx->h_pinned, 64 * sizeof(double), hipHostMallocDefault));
    HIP_TRY(ctx, hipHostMalloc((void**)&ctx->h_pcg, 2 * sizeof(PcgState), hipHostMallocDefault));
    for (auto& ev : ctx->pcg_ev) HIP_TRY(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    *out = ctx;
    return SYMX_HIP_OK;
}

void symx_hip_ctx_destroy(symx_hip_ctx ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    dfree(ctx->d_partials);
    dfree(ctx->d_counters);
    dfree(ctx->d_scalars);
    dfree(ctx->d_pcg);
    if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
    if (ctx->h_pcg) (void)hipHostFree(ctx->h_pcg);
    for (auto& ev : ctx->pcg_ev)
        if (ev) (void)hipEventDestroy(ev);
    for (int b = 0; b < 2; b++) {
        if (ctx->h_stage[b]) (void)hipHostFree(ctx->h_stage[b]);
        if (ctx->stage_ev[b]) (void)hipEventDestroy(ctx->stage_ev[b]);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* symx_hip_last_error(symx_hip_ctx ctx)
{
    if (ctx) return ctx->err.c_str();
    return g_thread_error.c_str();
}

const char* symx_hip_eval_prelude(void) { return kPrelude; }

int symx_hip_jit_compile(const char* source, const char* arch, const char* cache_dir, char* key_out, int* was_cached, size_t* code_size)
{
    if (!source) return fail(nullptr, SYMX_HIP_ERR_ARG, "null source");
    std::vector<char> code;
    std::string key;
    int cached = 0;
    int rc = jit_compile_impl(nullptr, source, arch, cache_dir, code, key, cached);
    if (rc) return rc;
    if (key_out) std::memcpy(key_out, key.c_str(), key.size() + 1);
    if (was_cached) *was_cached = cached;
    if (code_size) *code_size = code.size();
    return SYMX_HIP_OK;
}

int symx_hip_module_load(symx_hip_ctx ctx, const char* source, const char* cache_dir, symx_hip_module* out)
{
    if (!ctx || !source || !out) return fail(ctx, SYMX_HIP_ERR_ARG, "null argument");
    *out = nullptr;
    std::vector<char> code;
    std::string key;
    int cached = 0;
    int rc = jit_compile_impl(ctx, source, "gfx950", cache_dir, code, key, cached);
    if (rc) return rc;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    auto* m = new symx_hip_module_s();
    m->ctx = ctx;
    hipError_t e = hipModuleLoadData(&m->mod, code.data());
    if (e != hipSuccess) {
        delete m;
        return fail(ctx, SYMX_HIP_ERR_JIT, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    }
    *out = m;
    return SYMX_HIP_OK;
}

void symx_hip_module_unload(symx_hip_module mod)
{
    if (!mod) return;
    if (mod->mod) (void)hipModuleUnload(mod->mod);
    delete mod;
}

static int module_function(symx_hip_module mod, const char* name, hipFunction_t* f)
{
    auto it = mod->fns.find(name);
    if (it != mod->fns.end()) {
        *f = it->second;
        return SYMX_HIP_OK;
    }
    hipError_t e = hipModuleGetFunction(f, mod->mod, name);
    if (e != hipSuccess) return fail(mod->ctx, SYMX_HIP_ERR_JIT, std::string("kernel not found in module: ") + name);
    mod->fns[name] = *f;
    return SYMX_HIP_OK;
}

static int launch_eval(symx_hip_ctx ctx, symx_hip_module mod, const std::string& name, const symx_hip_eval_params& params,
                       int32_t elems_per_group = SYMX_HIP_BLOCK)
{
    if (params.n_elem <= 0) return SYMX_HIP_OK;
    hipFunction_t f;
    int rc = module_function(mod, name.c_str(), &f);
    if (rc) return rc;
    symx_hip_eval_params p = params;
    size_t sz = sizeof(p);
    void* config[] = { HIP_LAUNCH_PARAM_BUFFER_POINTER, &p, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END };
    const unsigned grid = (unsigned)((params.n_elem + elems_per_group - 1) / elems_per_group);
    HIP_TRY(ctx, hipModuleLaunchKernel(f, grid, 1, 1, SYMX_HIP_BLOCK, 1, 1, 0, ctx->stream, nullptr, config));
    return SYMX_HIP_OK;
}

int symx_hip_module_launch_eval(symx_hip_ctx ctx, symx_hip_module mod, const char* name, const symx_hip_eval_params* params)
{
    if (!ctx || !mod || !name || !params) return fail(ctx, SYMX_HIP_ERR_ARG, "null argument");
    int rc = launch_eval(ctx, mod, name, *params);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SYMX_HIP_OK;
}

// ---- problem ------------------------------------------------------------------

int symx_hip_problem_create(symx_hip_ctx ctx, int32_t n_dofs, symx_hip_problem* out)
{
    if (!ctx || !out) return fail(ctx, SYMX_HIP_ERR_ARG, "null argument");
    if (n_dofs <= 0 || n_dofs % 3 != 0) return fail(ctx, SYMX_HIP_ERR_ARG, "n_dofs must be a positive multiple of 3");
    auto* p = new symx_hip_problem_s();
    p->ctx = ctx;
    p->n_dofs = n_dofs;
    p->nb = n_dofs / 3;
    for (int v = 0; v < 5; v++) {
        int rc = dalloc(ctx, &p->d_vec[v], (size_t)n_dofs);
        if (rc) {
            symx_hip_problem_destroy(p);
            return rc;
        }
        HIP_TRY(ctx, hipMemsetAsync(p->d_vec[v], 0, sizeof(double) * n_dofs, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    *out = p;
    return SYMX_HIP_OK;
}

void symx_hip_problem_destroy(symx_hip_problem p)
{
    if (!p) return;
    (void)hipStreamSynchronize(p->ctx->stream);
    for (auto& v : p->d_vec) dfree(v);
    for (auto& pot : p->pots) {
        dfree(pot.d_conn);
        dfree(pot.d_active);
        for (auto& m : pot.maps)
            if (m.owner) dfree(m.d_data);
    }
    dfree(p->d_G);
    dfree(p->d_H);
    dfree(p->d_Epart);
    dfree(p->d_Eelem);
    dfree(p->d_flags);
    dfree(p->d_plist);
    dfree(p->d_projected);
    dfree(p->d_select);
    dfree(p->d_filter_blocks);
    dfree(p->d_goff);
    dfree(p->d_glist);
    dfree(p->d_grow);
    dfree(p->d_aslot);
    dfree(p->d_mirror);
    dfree(p->sell.d_slice_ptr);
    dfree(p->sell.d_col);
    dfree(p->sell.d_dcol);
    dfree(p->sell.d_row_at);
    dfree(p->d_coff);
    dfree(p->d_clist);
    dfree(p->d_val64);
    dfree(p->d_val32);
    dfree(p->d_dinv64);
    dfree(p->d_dinv32);
    dfree(p->d_x); dfree(p->d_r); dfree(p->d_z); dfree(p->d_p0); dfree(p->d_p1); dfree(p->d_q);
    dfree(p->f_x); dfree(p->f_r); dfree(p->f_z); dfree(p->f_p0); dfree(p->f_p1); dfree(p->f_q); dfree(p->f_b);
    delete p;
}

int symx_hip_problem_add_potential(symx_hip_problem p, const symx_hip_potential_desc* d, int32_t* pot_id)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!d || !d->module || !d->kernel_P || !d->kernel_PG || !d->kernel_PGH) return fail(ctx, SYMX_HIP_ERR_ARG, "incomplete potential descriptor");
    if (p->finalized) return fail(ctx, SYMX_HIP_ERR_ARG, "problem already finalized");
    if (d->n_elem < 0 || d->conn_stride <= 0 || (d->n_elem > 0 && !d->conn)) return fail(ctx, SYMX_HIP_ERR_ARG, "bad connectivity");
    if (d->n_maps < 0 || d->n_maps > SYMX_HIP_MAX_MAPS) return fail(ctx, SYMX_HIP_ERR_ARG, "too many data maps");
    if (d->n_dof_blocks <= 0) return fail(ctx, SYMX_HIP_ERR_ARG, "potential without DoFs");
    DevPotential pot;
    pot.module = d->module;
    pot.kP = d->kernel_P;
    pot.kPG = d->kernel_PG;
    pot.kPGH = d->kernel_PGH;
    pot.kC = d->kernel_C ? d->kernel_C : "";
    pot.kPGH32 = d->kernel_PGH_f32 ? d->kernel_PGH_f32 : "";
    pot.n_elem = d->n_elem;
    pot.stride = d->conn_stride;
    pot.elems_per_group = d->elems_per_group == 0 ? SYMX_HIP_BLOCK : d->elems_per_group;
    if (pot.elems_per_group != SYMX_HIP_BLOCK && pot.elems_per_group != 4) return fail(ctx, SYMX_HIP_ERR_ARG, "elems_per_group must be 256 or 4");
    if (pot.elems_per_group != SYMX_HIP_BLOCK && d->kernel_C) return fail(ctx, SYMX_HIP_ERR_ARG, "conditions need one element per thread");
    pot.nb = d->n_dof_blocks;
    pot.n_ub = pot.nb * (pot.nb + 1) / 2;
    pot.h_stride = d->h_stride == 0 ? pot.n_ub * 9 : d->h_stride;
    if (pot.h_stride < pot.n_ub * 9 || pot.h_stride % 3 != 0)
        return fail(ctx, SYMX_HIP_ERR_ARG, "h_stride must be a multiple of 3 and at least 9 values per upper block");
    pot.dof_col.assign(d->dof_conn_col, d->dof_conn_col + pot.nb);
    pot.dof_off.assign(d->dof_set_offset, d->dof_set_offset + pot.nb);
    const int64_t ne = d->n_elem;
    pot.h_conn.assign(d->conn, d->conn + ne * pot.stride);
    for (int64_t i = 0; i < ne * pot.stride; i++) {
        if (pot.h_conn[i] < 0) return fail(ctx, SYMX_HIP_ERR_ARG, "negative connectivity entry");
    }
    for (int a = 0; a < pot.nb; a++) {
        if (pot.dof_col[a] < 0 || pot.dof_col[a] >= pot.stride) return fail(ctx, SYMX_HIP_ERR_ARG, "DoF connectivity column out of range");
        if (pot.dof_off[a] < 0 || pot.dof_off[a] % 3 != 0) return fail(ctx, SYMX_HIP_ERR_ARG, "DoF set offset must be a multiple of 3");
        for (int64_t e = 0; e < ne; e++) {
            const int64_t dof = pot.dof_off[a] + 3 * (int64_t)pot.h_conn[e * pot.stride + pot.dof_col[a]];
            if (dof + 3 > p->n_dofs) return fail(ctx, SYMX_HIP_ERR_ARG, "DoF index out of range");
        }
    }
    // Device connectivity, transposed to [stride][n_elem] so that lane e reads conn[c*n+e].
    {
        std::vector<int32_t> t((size_t)ne * pot.stride);
        for (int64_t e = 0; e < ne; e++)
            for (int c = 0; c < pot.stride; c++) t[(size_t)c * ne + e] = pot.h_conn[(size_t)e * pot.stride + c];
        int rc = dalloc(ctx, &pot.d_conn, t.size());
        if (rc) return rc;
        HIP_TRY(ctx, hipMemcpy(pot.d_conn, t.data(), t.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    for (int m = 0; m < d->n_maps; m++) {
        DevMap dm;
        dm.is_dof = d->maps[m].is_dof;
        dm.dof_offset = d->maps[m].dof_offset;
        dm.n_values = d->maps[m].n_values;
        if (!dm.is_dof) {
            if (dm.n_values <= 0 || !d->maps[m].host) return fail(ctx, SYMX_HIP_ERR_ARG, "data map without data");
            dm.host = d->maps[m].host;
            // Maps of the same host array (one per element node, e.g. the rest positions of a
            // 27-node element) share one device copy.
            const DevMap* same = nullptr;
            auto find = [&](const std::vector<DevMap>& ms) {
                for (const DevMap& o : ms)
                    if (!same && !o.is_dof && o.owner && o.host == dm.host && o.n_values == dm.n_values) same = &o;
            };
            find(pot.maps);
            for (const auto& q : p->pots) find(q.maps);
            if (same) {
                dm.d_data = same->d_data;
                dm.owner = false;
            }
            else {
                int rc = dalloc(ctx, &dm.d_data, (size_t)dm.n_values);
                if (rc) return rc;
                HIP_TRY(ctx, hipMemcpy(dm.d_data, d->maps[m].host, dm.n_values * sizeof(double), hipMemcpyHostToDevice));
            }
        }
        else if (dm.dof_offset < 0 || dm.dof_offset >= p->n_dofs) {
            return fail(ctx, SYMX_HIP_ERR_ARG, "DoF map offset out of range");
        }
        pot.maps.push_back(dm);
    }
    if (!pot.kC.empty()) {
        int rc = dalloc(ctx, &pot.d_active, (size_t)std::max<int64_t>(ne, 1));
        if (rc) return rc;
    }
    p->pots.push_back(std::move(pot));
    if (pot_id) *pot_id = (int32_t)p->pots.size() - 1;
    return SYMX_HIP_OK;
}

// ---- host transfers of user arrays ---------------------------------------------------
// User arrays live in pageable host memory, which the runtime copies through its own staging
// with one host thread. Large transfers here go through two pinned staging buffers of
// kStageBytes instead: each chunk is copied on the host by several threads while the DMA of
// the previous chunk runs.
namespace
{
    constexpr size_t kStageBytes = size_t(16) << 20;
    constexpr size_t kStageMin = size_t(4) << 20;  // smaller transfers go straight through

    void parallel_memcpy(void* dst, const void* src, size_t n)
    {
        const unsigned hw = std::max(1u, std::min(std::thread::hardware_concurrency(), 8u));
        const size_t T = std::min<size_t>(hw, (n + (size_t(1) << 20) - 1) >> 20);  // >= 1 MiB per thread
        if (T <= 1) {
            std::memcpy(dst, src, n);
            return;
        }
        const size_t per = (n + T - 1) / T;
        std::vector<std::thread> th;
        for (size_t t = 1; t < T; t++) {
            const size_t a = t * per;
            if (a >= n) break;
            const size_t b = std::min(n, a + per);
            th.emplace_back([=]() { std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a); });
        }
        std::memcpy(dst, src, std::min(per, n));
        for (auto& x : th) x.join();
    }

    int ensure_staging(symx_hip_ctx ctx)
    {
        for (int b = 0; b < 2; b++) {
            if (!ctx->h_stage[b]) HIP_TRY(ctx, hipHostMalloc((void**)&ctx->h_stage[b], kStageBytes, hipHostMallocDefault));
            if (!ctx->stage_ev[b]) HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->stage_ev[b], hipEventDisableTiming));
        }
        return SYMX_HIP_OK;
    }

    // Host -> device on the context's stream (after the work already on it); returns when
    // the data is on the device.
    int host_to_device(symx_hip_ctx ctx, void* d, const void* host, size_t bytes)
    {
        if (bytes < kStageMin) {
            HIP_TRY(ctx, hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            return SYMX_HIP_OK;
        }
        int rc = ensure_staging(ctx);
        if (rc) return rc;
        bool used[2] = { false, false };
        int b = 0;
        for (size_t off = 0; off < bytes; off += kStageBytes, b ^= 1) {
            const size_t n = std::min(kStageBytes, bytes - off);
            if (used[b]) HIP_TRY(ctx, hipEventSynchronize(ctx->stage_ev[b]));  // its previous chunk has left
            parallel_memcpy(ctx->h_stage[b], static_cast<const char*>(host) + off, n);
            HIP_TRY(ctx, hipMemcpyAsync(static_cast<char*>(d) + off, ctx->h_stage[b], n, hipMemcpyHostToDevice, ctx->stream));
            HIP_TRY(ctx, hipEventRecord(ctx->stage_ev[b], ctx->stream));
            used[b] = true;
        }
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return SYMX_HIP_OK;
    }

    // Device -> host after the work already on the context's stream.
    int device_to_host(symx_hip_ctx ctx, void* host, const void* d, size_t bytes)
    {
        if (bytes < kStageMin) {
            HIP_TRY(ctx, hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            return SYMX_HIP_OK;
        }
        int rc = ensure_staging(ctx);
        if (rc) return rc;
        const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
        auto enqueue = [&](size_t i) -> int {
            const size_t off = i * kStageBytes, n = std::min(kStageBytes, bytes - off);
            HIP_TRY(ctx, hipMemcpyAsync(ctx->h_stage[i & 1], static_cast<const char*>(d) + off, n, hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipEventRecord(ctx->stage_ev[i & 1], ctx->stream));
            return SYMX_HIP_OK;
        };
        rc = enqueue(0);
        for (size_t i = 0; !rc && i < nch; i++) {
            // chunk i + 1 goes to the buffer whose chunk i - 1 was copied out in the last round
            if (i + 1 < nch) rc = enqueue(i + 1);
            if (rc) break;
            HIP_TRY(ctx, hipEventSynchronize(ctx->stage_ev[i & 1]));
            const size_t off = i * kStageBytes;
            parallel_memcpy(static_cast<char*>(host) + off, ctx->h_stage[i & 1], std::min(kStageBytes, bytes - off));
        }
        return rc;
    }
}

int symx_hip_problem_set_map_data(symx_hip_problem p, int32_t pot_id, int32_t map_id, const double* host, int64_t n)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (pot_id < 0 || pot_id >= (int32_t)p->pots.size()) return fail(ctx, SYMX_HIP_ERR_ARG, "bad potential id");
    auto& pot = p->pots[pot_id];
    if (map_id < 0 || map_id >= (int32_t)pot.maps.size()) return fail(ctx, SYMX_HIP_ERR_ARG, "bad map id");
    auto& m = pot.maps[map_id];
    if (m.is_dof) return fail(ctx, SYMX_HIP_ERR_ARG, "map is bound to the DoF vector");
    if (n != m.n_values) return fail(ctx, SYMX_HIP_ERR_ARG, "data map size changed");
    if (!m.owner) {
        if (host == m.host) return SYMX_HIP_OK;  // the owning map uploads this array
        // the array moved away from the one it shared: give this map its own copy
        int rc = dalloc(ctx, &m.d_data, (size_t)n);
        if (rc) return rc;
        m.owner = true;
    }
    m.host = host;
    return host_to_device(ctx, m.d_data, host, (size_t)n * sizeof(double));
}

// Topology: gradient scatter lists, block pattern (sliced, diagonal first) and
// slot-sorted Hessian contribution lists. All orders are stable in
// (potential, element, local block row, local block col), which fixes the
// summation order of every output entry.
// Rows sorted by length give slices of very different widths (27-node hexahedra: 27 to 125
// blocks), and the SpMV's fixed schedule (wave v of a workgroup takes slices v, v + 8, ... of
// its chunk) then loads some waves with two wide slices while others hold one narrow one:
// the kernel lasts as long as its most loaded wave (2.3x the mean on a 40^3 Hex27 cube).
// Within each workgroup chunk the whole slices are reordered so that, taken largest first,
// each goes to the least loaded wave that still has a slot, ties to waves with fewer slots
// (1.7x the mean there). Only the order of slices inside a chunk changes.
static void balance_slices(std::vector<int32_t>& order, const std::vector<int32_t>& row_len, int32_t nb)
{
    const int G = pcg_grid(nb);
    const int64_t per = (((int64_t)nb + G - 1) / G + 63) / 64 * 64;
    const int nw = kSpmvThreads / 64;
    const int64_t full = nb / 64;  // slices wholly inside the matrix (a last partial one stays last)
    std::vector<int32_t> width(full);
    for (int64_t sl = 0; sl < full; sl++) {
        int32_t w = 0;
        for (int64_t q = 64 * sl; q < 64 * sl + 64; q++) w = std::max(w, row_len[order[q]]);
        width[sl] = w;
    }
    std::vector<int32_t> tmp;
    for (int64_t c0 = 0; c0 < nb; c0 += per) {
        const int64_t s0 = c0 / 64, s1 = std::min<int64_t>((c0 + per) / 64, full);
        const int64_t m = s1 - s0;
        if (m <= 1) continue;
        std::vector<int64_t> by_width(m);
        for (int64_t i = 0; i < m; i++) by_width[i] = s0 + i;
        std::stable_sort(by_width.begin(), by_width.end(), [&](int64_t a, int64_t b) { return width[a] > width[b]; });
        std::vector<int64_t> load(nw, 0), next(nw), assign(m, -1);
        for (int v = 0; v < nw; v++) next[v] = v;  // next free slot of wave v: v, v + nw, ...
        auto slots_left = [&](int v) { return next[v] < m ? (m - 1 - next[v]) / nw + 1 : 0; };
        for (int64_t sl : by_width) {
            int best = -1;
            for (int v = 0; v < nw; v++) {
                if (next[v] >= m) continue;
                if (best < 0 || load[v] < load[best] || (load[v] == load[best] && slots_left(v) < slots_left(best))) best = v;
            }
            assign[next[best]] = sl;
            next[best] += nw;
            load[best] += width[sl];
        }
        tmp.assign(order.begin() + 64 * s0, order.begin() + 64 * s1);
        for (int64_t i = 0; i < m; i++)
            std::copy(tmp.begin() + 64 * (assign[i] - s0), tmp.begin() + 64 * (assign[i] - s0) + 64, order.begin() + 64 * (s0 + i));
    }
}

int symx_hip_problem_finalize(symx_hip_problem p)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (p->finalized) return SYMX_HIP_OK;
    if (p->pots.empty()) return fail(ctx, SYMX_HIP_ERR_ARG, "problem without potentials");
    const int32_t nb = p->nb;

    int64_t g = 0, h = 0, parts = 0, elems = 0;  // h: values of all Hessian records
    uint64_t n_contrib = 0;
    for (auto& pot : p->pots) {
        pot.g_base = g;
        pot.h_off = h;
        pot.part_base = parts;
        pot.e_base = elems;
        elems += pot.n_elem;
        pot.n_groups = (pot.n_elem + pot.elems_per_group - 1) / pot.elems_per_group;
        g += (int64_t)pot.n_elem * pot.nb;
        h += (int64_t)pot.n_elem * pot.h_stride;
        parts += pot.n_groups;
        n_contrib += (uint64_t)pot.n_elem * pot.nb * pot.nb;
    }
    if ((h / 3) * 2 + 2 >= (int64_t)UINT32_MAX || g >= (int64_t)UINT32_MAX || n_contrib >= (uint64_t)UINT32_MAX)
        return fail(ctx, SYMX_HIP_ERR_ARG, "problem too large for 32-bit assembly maps");
    p->total_g = g;
    p->total_h = h;
    p->total_parts = parts;
    p->total_elem = elems;

    auto row_of = [&](const DevPotential& pot, int64_t e, int a) -> int32_t {
        return pot.dof_off[a] / 3 + pot.h_conn[(size_t)e * pot.stride + pot.dof_col[a]];
    };

    // Gradient lists.
    std::vector<uint32_t> goff((size_t)nb + 1, 0);
    for (auto& pot : p->pots)
        for (int64_t e = 0; e < pot.n_elem; e++)
            for (int a = 0; a < pot.nb; a++) goff[row_of(pot, e, a) + 1]++;
    for (int32_t r = 0; r < nb; r++) goff[r + 1] += goff[r];
    std::vector<uint32_t> glist(goff[nb]);
    {
        std::vector<uint32_t> cur(goff.begin(), goff.end() - 1);
        for (auto& pot : p->pots)
            for (int64_t e = 0; e < pot.n_elem; e++)
                for (int a = 0; a < pot.nb; a++) glist[cur[row_of(pot, e, a)]++] = (uint32_t)(pot.g_base + e * pot.nb + a);
    }

    // Block pattern: candidate columns per row, sorted and unique, diagonal first.
    std::vector<uint64_t> coff((size_t)nb + 1, 0);
    for (int32_t r = 0; r < nb; r++) coff[r + 1] = 1;  // every row keeps its diagonal block
    for (auto& pot : p->pots)
        for (int64_t e = 0; e < pot.n_elem; e++)
            for (int a = 0; a < pot.nb; a++) coff[row_of(pot, e, a) + 1] += pot.nb;
    for (int32_t r = 0; r < nb; r++) coff[r + 1] += coff[r];
    std::vector<int32_t> ccol(coff[nb]);
    {
        std::vector<uint64_t> cur(coff.begin(), coff.end() - 1);
        for (int32_t r = 0; r < nb; r++) ccol[cur[r]++] = r;
        for (auto& pot : p->pots)
            for (int64_t e = 0; e < pot.n_elem; e++)
                for (int a = 0; a < pot.nb; a++) {
                    const int32_t r = row_of(pot, e, a);
                    for (int b = 0; b < pot.nb; b++) ccol[cur[r]++] = row_of(pot, e, b);
                }
    }
    std::vector<int32_t> row_len(nb);
    std::vector<uint64_t> row_begin(nb);
    for (int32_t r = 0; r < nb; r++) {
        int32_t* b = ccol.data() + coff[r];
        int32_t* e = ccol.data() + coff[r + 1];
        std::sort(b, e);
        int32_t* u = std::unique(b, e);
        const int32_t n = (int32_t)(u - b);
        // diagonal first (always present), remaining columns ascending
        int32_t* d = std::lower_bound(b, u, r);
        std::rotate(b, d, d + 1);
        row_len[r] = n;
        row_begin[r] = coff[r];
    }
    SellPattern& S = p->sell;
    S.n_rows = nb;
    S.n_slices = (nb + 63) / 64;
    S.h_slice_ptr.assign((size_t)S.n_slices + 1, 0);
    S.h_row_len = row_len;
    S.nnzb = 0;
    for (int32_t r = 0; r < nb; r++) S.nnzb += row_len[r];
    // Slot count with rows in natural order and with rows sorted by decreasing length
    // (stable) inside windows of kSellSigma rows; keep the sort only if it removes at
    // least 5% of the slots (rows of mixed length interleave in higher-order meshes:
    // a 27-node hex mesh fills 67% of a natural SELL-64 and 97% of a sorted one).
    auto slots_of = [&](const int32_t* order) -> int64_t {
        int64_t t = 0;
        for (int32_t s = 0; s < S.n_slices; s++) {
            int32_t w = 0;
            for (int32_t q = 64 * s; q < std::min(nb, 64 * s + 64); q++) w = std::max(w, row_len[order ? order[q] : q]);
            t += 64 * (int64_t)w;
        }
        return t;
    };
    {
        std::vector<int32_t> order(nb);
        for (int32_t r = 0; r < nb; r++) order[r] = r;
        for (int32_t w0 = 0; w0 < nb; w0 += kSellSigma)
            std::stable_sort(order.begin() + w0, order.begin() + std::min(nb, w0 + kSellSigma),
                             [&](int32_t a, int32_t b) { return row_len[a] > row_len[b]; });
        if (20 * slots_of(order.data()) <= 19 * slots_of(nullptr)) {
            balance_slices(order, row_len, nb);
            S.h_pos.assign(nb, 0);
            for (int32_t q = 0; q < nb; q++) S.h_pos[order[q]] = q;
            S.h_row_at = std::move(order);
        }
    }
    for (int32_t s = 0; s < S.n_slices; s++) {
        int32_t w = 0;
        for (int32_t q = 64 * s; q < std::min(nb, 64 * s + 64); q++) w = std::max(w, row_len[S.h_row_at.empty() ? q : S.h_row_at[q]]);
        S.h_slice_ptr[s + 1] = S.h_slice_ptr[s] + 64 * (int64_t)w;
    }
    S.n_slots = S.h_slice_ptr[S.n_slices];
    if (S.n_slots >= (int64_t)UINT32_MAX) return fail(ctx, SYMX_HIP_ERR_ARG, "block matrix too large");
    S.h_col.assign(S.n_slots, 0);
    for (int32_t r = 0; r < nb; r++) {
        const int32_t q = S.pos_of(r);
        const int64_t base = S.h_slice_ptr[q >> 6] + (q & 63);
        const int32_t s = q >> 6;
        const int32_t w = (int32_t)((S.h_slice_ptr[s + 1] - S.h_slice_ptr[s]) >> 6);
        for (int32_t j = 0; j < w; j++) S.h_col[base + 64 * (int64_t)j] = (j < row_len[r]) ? ccol[row_begin[r] + j] : r;
    }
    auto find_slot = [&](int32_t r, int32_t c) -> int64_t {
        const int32_t* b = ccol.data() + row_begin[r];
        int32_t j;
        if (b[0] == c) j = 0;
        else j = (int32_t)(std::lower_bound(b + 1, b + row_len[r], c) - b);
        const int32_t q = S.pos_of(r);
        return S.h_slice_ptr[q >> 6] + 64 * (int64_t)j + (q & 63);
    };

    // Contributions, counting-sorted by slot (stable).
    std::vector<uint32_t> slot_of(n_contrib);
    std::vector<uint32_t> soff((size_t)S.n_slots + 1, 0);
    {
        uint64_t k = 0;
        for (auto& pot : p->pots) {
            std::vector<int32_t> rows(pot.nb);
            for (int64_t e = 0; e < pot.n_elem; e++) {
                for (int a = 0; a < pot.nb; a++) rows[a] = row_of(pot, e, a);
                for (int a = 0; a < pot.nb; a++)
                    for (int b = 0; b < pot.nb; b++) {
                        const uint32_t sl = (uint32_t)find_slot(rows[a], rows[b]);
                        slot_of[k++] = sl;
                        soff[sl + 1]++;
                    }
            }
        }
    }
    for (int64_t t = 0; t < S.n_slots; t++) soff[t + 1] += soff[t];
    std::vector<uint32_t> clist(n_contrib);
    {
        std::vector<uint32_t> cur(soff.begin(), soff.end() - 1);
        uint64_t k = 0;
        for (auto& pot : p->pots) {
            for (int64_t e = 0; e < pot.n_elem; e++) {
                for (int a = 0; a < pot.nb; a++)
                    for (int b = 0; b < pot.nb; b++) {
                        const int lo = std::min(a, b), hi = std::max(a, b);
                        const int ub = lo * pot.nb - (lo * (lo - 1)) / 2 + (hi - lo);
                        const uint64_t off3 = ((uint64_t)pot.h_off + (uint64_t)e * pot.h_stride + 9u * ub) / 3;
                        clist[cur[slot_of[k++]]++] = (uint32_t)(off3 * 2 + (a > b ? 1u : 0u));
                    }
            }
        }
    }

    // Traversal order of the two assembly kernels: SELL slices (64 block rows) in patches of
    // kPatchSlices grown breadth-first over the block pattern, seeds in slice order. The
    // workgroups in flight then gather the element blocks of one compact region of the mesh,
    // and neighbouring rows that share elements are assembled at about the same time (in
    // slice order a lexicographically numbered grid meets every element again one grid
    // layer later, after the L2 has dropped it). Within a slice the slot order is unchanged,
    // so the value stores stay contiguous; the per-slot and per-row summation order is
    // unchanged.
    std::vector<int32_t> strav;
    strav.reserve(S.n_slices);
    {
        constexpr size_t kPatchSlices = 64;
        std::vector<uint8_t> seen(S.n_slices, 0);
        auto row_at_pos = [&](int32_t q) { return S.h_row_at.empty() ? q : S.h_row_at[q]; };
        for (int32_t seed = 0; seed < S.n_slices; seed++) {
            if (seen[seed]) continue;
            const size_t head = strav.size();
            strav.push_back(seed);
            seen[seed] = 1;
            for (size_t qi = head; qi < strav.size() && strav.size() - head < kPatchSlices; qi++) {
                const int32_t sl = strav[qi];
                for (int32_t q = 64 * sl; q < std::min(nb, 64 * sl + 64) && strav.size() - head < kPatchSlices; q++) {
                    const int32_t r = row_at_pos(q);
                    for (int32_t j = 0; j < row_len[r] && strav.size() - head < kPatchSlices; j++) {
                        const int32_t sc = S.pos_of(ccol[row_begin[r] + j]) >> 6;
                        if (!seen[sc]) {
                            seen[sc] = 1;
                            strav.push_back(sc);
                        }
                    }
                }
            }
        }
    }
    // gradient lists and slots in traversal order
    std::vector<int32_t> trav;
    trav.reserve(nb);
    std::vector<uint32_t> aslot;
    aslot.reserve(S.n_slots);
    for (int32_t sl : strav) {
        for (int32_t q = 64 * sl; q < std::min(nb, 64 * sl + 64); q++) trav.push_back(S.h_row_at.empty() ? q : S.h_row_at[q]);
        for (int64_t t = S.h_slice_ptr[sl]; t < S.h_slice_ptr[sl + 1]; t++) aslot.push_back((uint32_t)t);
    }
    if ((int64_t)trav.size() != nb || (int64_t)aslot.size() != S.n_slots)
        return fail(ctx, SYMX_HIP_ERR_ARG, "internal: assembly traversal does not cover every row and slot");
    std::vector<uint32_t> goff_t((size_t)nb + 1, 0), glist_t(glist.size());
    for (int32_t t = 0; t < nb; t++) {
        const int32_t r = trav[t];
        goff_t[t + 1] = goff_t[t] + (goff[r + 1] - goff[r]);
        std::copy(glist.begin() + goff[r], glist.begin() + goff[r + 1], glist_t.begin() + goff_t[t]);
    }
    // Mirrored assembly: the slot of block (c, r) lists the element blocks of (r, c) in the
    // same (potential, element) order with the transpose flag flipped, so its sum is the
    // transposed sum of (r, c) bit for bit. The upper slot (c > r) then forms the block once
    // and stores it to both, and each element block is read once instead of twice. Every
    // pair of lists is checked; if one does not mirror exactly, every slot sums its own list.
    std::vector<uint32_t> mirror;
    {
        bool ok = true;
        std::vector<uint32_t> m_of(S.n_slots, ~0u);
        std::vector<uint8_t> lower(S.n_slots, 0);
        for (int32_t sl = 0; sl < S.n_slices && ok; sl++) {
            for (int64_t t = S.h_slice_ptr[sl]; t < S.h_slice_ptr[sl + 1] && ok; t++) {
                const int64_t o = t - S.h_slice_ptr[sl];
                const int32_t q = 64 * sl + (int32_t)(o & 63);
                const int32_t j = (int32_t)(o >> 6);
                if (q >= nb) continue;
                const int32_t r = S.h_row_at.empty() ? q : S.h_row_at[q];
                if (j == 0 || j >= row_len[r]) continue;  // diagonal or padding: its own list
                const int32_t c = S.h_col[t];
                if (c < r) {
                    lower[t] = 1;
                    continue;
                }
                const int64_t m = find_slot(c, r);
                const uint32_t n0 = soff[t + 1] - soff[t];
                if (m < 0 || m >= S.n_slots || S.h_col[m] != r || soff[m + 1] - soff[m] != n0) {
                    ok = false;
                    break;
                }
                for (uint32_t k = 0; k < n0 && ok; k++)
                    if (clist[soff[m] + k] != (clist[soff[t] + k] ^ 1u)) ok = false;
                m_of[t] = (uint32_t)m;
            }
        }
        if (ok) {
            std::vector<uint32_t> kept;
            kept.reserve(aslot.size());
            for (uint32_t sl : aslot)
                if (!lower[sl]) {
                    kept.push_back(sl);
                    mirror.push_back(m_of[sl]);
                }
            aslot.swap(kept);
        }
    }
    p->n_assemble = (int64_t)aslot.size();
    std::vector<uint32_t> coff_t(aslot.size() + 1, 0), clist_t;
    for (size_t t = 0; t < aslot.size(); t++) coff_t[t + 1] = coff_t[t] + (soff[aslot[t] + 1] - soff[aslot[t]]);
    clist_t.resize(coff_t.back());
    for (size_t t = 0; t < aslot.size(); t++) {
        const uint32_t sl = aslot[t];
        coff_t[t + 1] = coff_t[t] + (soff[sl + 1] - soff[sl]);
        std::copy(clist.begin() + soff[sl], clist.begin() + soff[sl + 1], clist_t.begin() + coff_t[t]);
    }
    std::vector<uint32_t>().swap(clist);  // the slot-order lists are not needed any more
    std::vector<uint32_t>().swap(slot_of);
    std::vector<uint32_t>().swap(glist);

    // Upload.
    int rc = 0;
    rc = rc ? rc : dalloc(ctx, &p->d_G, (size_t)std::max<int64_t>(g * 3, 1));
    rc = rc ? rc : dalloc(ctx, &p->d_H, (size_t)std::max<int64_t>(h, 1));
    rc = rc ? rc : dalloc(ctx, &p->d_Epart, (size_t)std::max<int64_t>(parts, 1));
    rc = rc ? rc : dalloc(ctx, &p->d_Eelem, (size_t)std::max<int64_t>(p->total_elem, 1));
    {
        int64_t max_elem = 1;
        for (auto& pot : p->pots) max_elem = std::max<int64_t>(max_elem, pot.n_elem);
        rc = rc ? rc : dalloc(ctx, &p->d_flags, (size_t)max_elem);
        rc = rc ? rc : dalloc(ctx, &p->d_plist, (size_t)max_elem);
    }
    rc = rc ? rc : dalloc(ctx, &p->d_goff, goff_t.size());
    rc = rc ? rc : dalloc(ctx, &p->d_glist, std::max<size_t>(glist_t.size(), 1));
    rc = rc ? rc : dalloc(ctx, &p->d_grow, trav.size());
    rc = rc ? rc : dalloc(ctx, &p->d_aslot, std::max<size_t>(aslot.size(), 1));
    if (!mirror.empty()) rc = rc ? rc : dalloc(ctx, &p->d_mirror, mirror.size());
    rc = rc ? rc : dalloc(ctx, &S.d_slice_ptr, S.h_slice_ptr.size());
    // Column offsets from the row fit 16 bits for meshes numbered with some locality
    // (structured grids up to ~180^2 vertices per layer); the SpMV then reads 2 bytes
    // per block for its column instead of 4.
    std::vector<int16_t> dcol;
    {
        bool fits = true;
        dcol.assign(S.n_slots, 0);
        for (int32_t r = 0; r < nb && fits; r++) {
            const int32_t q = S.pos_of(r);
            const int64_t base = S.h_slice_ptr[q >> 6] + (q & 63);
            const int32_t w = (int32_t)((S.h_slice_ptr[(q >> 6) + 1] - S.h_slice_ptr[q >> 6]) >> 6);
            for (int32_t j = 0; j < w; j++) {
                const int64_t d = (int64_t)S.h_col[base + 64 * (int64_t)j] - r;
                if (d < INT16_MIN || d > INT16_MAX) {
                    fits = false;
                    break;
                }
                dcol[base + 64 * (int64_t)j] = (int16_t)d;
            }
        }
        if (!fits) dcol.clear();
    }
    if (dcol.empty()) rc = rc ? rc : dalloc(ctx, &S.d_col, (size_t)S.n_slots);
    else rc = rc ? rc : dalloc(ctx, &S.d_dcol, (size_t)S.n_slots);
    if (!S.h_row_at.empty()) rc = rc ? rc : dalloc(ctx, &S.d_row_at, S.h_row_at.size());
    rc = rc ? rc : dalloc(ctx, &p->d_coff, coff_t.size());
    rc = rc ? rc : dalloc(ctx, &p->d_clist, std::max<size_t>(clist_t.size(), 1));
    const size_t nd = (size_t)p->n_dofs;
    rc = rc ? rc : dalloc(ctx, &p->d_x, nd);
    rc = rc ? rc : dalloc(ctx, &p->d_r, nd);
    rc = rc ? rc : dalloc(ctx, &p->d_z, nd);
    rc = rc ? rc : dalloc(ctx, &p->d_p0, nd);
    rc = rc ? rc : dalloc(ctx, &p->d_p1, nd);
    rc = rc ? rc : dalloc(ctx, &p->d_q, nd);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpy(p->d_goff, goff_t.data(), goff_t.size() * 4, hipMemcpyHostToDevice));
    if (!glist_t.empty()) HIP_TRY(ctx, hipMemcpy(p->d_glist, glist_t.data(), glist_t.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(p->d_grow, trav.data(), trav.size() * 4, hipMemcpyHostToDevice));
    if (!aslot.empty()) HIP_TRY(ctx, hipMemcpy(p->d_aslot, aslot.data(), aslot.size() * 4, hipMemcpyHostToDevice));
    if (p->d_mirror) HIP_TRY(ctx, hipMemcpy(p->d_mirror, mirror.data(), mirror.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(S.d_slice_ptr, S.h_slice_ptr.data(), S.h_slice_ptr.size() * 8, hipMemcpyHostToDevice));
    if (S.d_col) HIP_TRY(ctx, hipMemcpy(S.d_col, S.h_col.data(), S.h_col.size() * 4, hipMemcpyHostToDevice));
    if (S.d_dcol) HIP_TRY(ctx, hipMemcpy(S.d_dcol, dcol.data(), dcol.size() * 2, hipMemcpyHostToDevice));
    if (S.d_row_at) HIP_TRY(ctx, hipMemcpy(S.d_row_at, S.h_row_at.data(), S.h_row_at.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(p->d_coff, coff_t.data(), coff_t.size() * 4, hipMemcpyHostToDevice));
    if (!clist_t.empty()) HIP_TRY(ctx, hipMemcpy(p->d_clist, clist_t.data(), clist_t.size() * 4, hipMemcpyHostToDevice));
    // PCG: 5 partials per workgroup (launch_pcg_iter*); vector reductions: 2 per workgroup + slack
    rc = ensure_partials(ctx, (size_t)std::max<int64_t>({ 2 * (int64_t)reduction_groups(p->n_dofs) + 2 * ((nb + 255) / 256) + 16,
                                                          5 * (int64_t)reduction_groups(nb) + 16, 4096 }));
    if (rc) return rc;
    p->finalized = true;
    return SYMX_HIP_OK;
}

int symx_hip_problem_set_dofs(symx_hip_problem p, const double* host)
{
    return symx_hip_problem_upload(p, SYMX_HIP_VEC_DOFS, host);
}
int symx_hip_problem_get_dofs(symx_hip_problem p, double* host)
{
    return symx_hip_problem_download(p, SYMX_HIP_VEC_DOFS, host);
}
int symx_hip_problem_download(symx_hip_problem p, int32_t vec, double* host)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    double* d = vec_ptr(p, vec);
    if (!d || !host) return fail(p->ctx, SYMX_HIP_ERR_ARG, "bad vector id or null pointer");
    return device_to_host(p->ctx, host, d, sizeof(double) * (size_t)p->n_dofs);
}
int symx_hip_problem_upload(symx_hip_problem p, int32_t vec, const double* host)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    double* d = vec_ptr(p, vec);
    if (!d || !host) return fail(p->ctx, SYMX_HIP_ERR_ARG, "bad vector id or null pointer");
    if (vec == SYMX_HIP_VEC_GRAD) p->grad_max_valid = false;
    return host_to_device(p->ctx, d, host, sizeof(double) * (size_t)p->n_dofs);
}
int symx_hip_problem_device_ptr(symx_hip_problem p, int32_t vec, double** dptr)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    double* d = vec_ptr(p, vec);
    if (!d || !dptr) return fail(p->ctx, SYMX_HIP_ERR_ARG, "bad vector id or null pointer");
    if (vec == SYMX_HIP_VEC_GRAD) {
        p->grad_ptr_exposed = true;  // the caller may write it: no cached max|grad| from here on
        p->grad_max_valid = false;
    }
    *dptr = d;
    return SYMX_HIP_OK;
}

// FP32 element-Hessian storage reuses the FP64 buffer (it needs half of it).
static float* h32_of(symx_hip_problem p, const DevPotential& pot)
{
    return reinterpret_cast<float*>(p->d_H) + pot.h_off;
}

static void fill_params(symx_hip_problem p, DevPotential& pot, symx_hip_eval_params& prm)
{
    std::memset(&prm, 0, sizeof(prm));
    prm.conn = pot.d_conn;
    prm.active = pot.kC.empty() ? nullptr : pot.d_active;
    for (size_t m = 0; m < pot.maps.size(); m++)
        prm.maps[m] = pot.maps[m].is_dof ? p->d_vec[SYMX_HIP_VEC_DOFS] + pot.maps[m].dof_offset : pot.maps[m].d_data;
    prm.E_partial = p->d_Epart + pot.part_base;
    prm.G = p->d_G + pot.g_base * 3;
    prm.H = p->d_H + pot.h_off;
    prm.C_out = pot.d_active;
    prm.E_elem = p->d_Eelem + pot.e_base;
    prm.n_elem = pot.n_elem;
    prm.e_mode = SYMX_HIP_EMODE_SUM;
}

int symx_hip_problem_eval(symx_hip_problem p, int32_t mode, double* E)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->finalized) return fail(ctx, SYMX_HIP_ERR_ARG, "problem not finalized");
    if (mode < SYMX_HIP_EVAL_P || mode > SYMX_HIP_EVAL_PGH_F32) return fail(ctx, SYMX_HIP_ERR_ARG, "bad evaluation mode");
    const bool f32 = mode == SYMX_HIP_EVAL_PGH_F32;
    if (f32)
        for (auto& pot : p->pots)
            if (pot.n_elem > 0 && pot.kPGH32.empty()) return fail(ctx, SYMX_HIP_ERR_ARG, "a potential has no FP32-storage Hessian kernel");
    for (auto& pot : p->pots) {
        if (pot.n_elem == 0) continue;
        symx_hip_eval_params prm;
        fill_params(p, pot, prm);
        if (f32) prm.H = reinterpret_cast<double*>(h32_of(p, pot));  // the _f32 kernels store floats
        if (!pot.kC.empty()) {
            int rc = launch_eval(ctx, pot.module, pot.kC, prm);
            if (rc) return rc;
        }
        const std::string& k = mode == SYMX_HIP_EVAL_P ? pot.kP : (mode == SYMX_HIP_EVAL_PG ? pot.kPG : (f32 ? pot.kPGH32 : pot.kPGH));
        int rc = launch_eval(ctx, pot.module, k, prm, pot.elems_per_group);
        if (rc) return rc;
        if (mode != SYMX_HIP_EVAL_P) {
            // PG / PGH keep the element energies, the reference point of
            // symx_hip_problem_energy_change, and take them from the energy kernel that
            // evaluates the trial points there: the same code gives bitwise the same E_e at
            // the same DoFs, and the differences carry no rounding of the derivative
            // kernels' different instruction mix (its partial sums replace theirs)
            prm.e_mode = SYMX_HIP_EMODE_STORE;
            rc = launch_eval(ctx, pot.module, pot.kP, prm, pot.elems_per_group);
            if (rc) return rc;
        }
    }
    if (mode != SYMX_HIP_EVAL_P) p->e_ref_valid = true;
    const bool with_h = mode >= SYMX_HIP_EVAL_PGH;
    if (with_h && p->d_projected) launch_fill_u8(ctx->stream, p->d_projected, p->total_elem);
    launch_sum(ctx->stream, p->d_Epart, p->total_parts, ctx->d_scalars);
    const bool with_grad = mode >= SYMX_HIP_EVAL_PG;
    if (with_grad) {
        launch_grad_assemble(ctx->stream, p->d_goff, p->d_glist, p->d_grow, p->d_G, p->d_vec[SYMX_HIP_VEC_GRAD], p->nb);
        // max|grad| (the Newton residual) comes back with the energy: one host round trip
        launch_max_abs(ctx->stream, p->n_dofs, p->d_vec[SYMX_HIP_VEC_GRAD], ctx->d_scalars + 1, ctx->d_partials);
    }
    if (with_h) {
        p->h_valid = true;
        p->h_f32 = f32;
    }
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, ctx->d_scalars, sizeof(double) * (with_grad ? 2 : 1), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (with_grad) {
        p->grad_max = ctx->h_pinned[1];
        p->grad_max_valid = !p->grad_ptr_exposed;
    }
    if (E) *E = ctx->h_pinned[0];
    return SYMX_HIP_OK;
}

int symx_hip_problem_energy_change(symx_hip_problem p, double* dE)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->finalized) return fail(ctx, SYMX_HIP_ERR_ARG, "problem not finalized");
    if (!p->e_ref_valid) return fail(ctx, SYMX_HIP_ERR_ARG, "no reference energies (evaluate with EVAL_PG or EVAL_PGH first)");
    if (!dE) return fail(ctx, SYMX_HIP_ERR_ARG, "null argument");
    for (auto& pot : p->pots) {
        if (pot.n_elem == 0) continue;
        symx_hip_eval_params prm;
        fill_params(p, pot, prm);
        prm.e_mode = SYMX_HIP_EMODE_DELTA;
        if (!pot.kC.empty()) {
            int rc = launch_eval(ctx, pot.module, pot.kC, prm);
            if (rc) return rc;
        }
        int rc = launch_eval(ctx, pot.module, pot.kP, prm, pot.elems_per_group);
        if (rc) return rc;
    }
    launch_sum(ctx->stream, p->d_Epart, p->total_parts, ctx->d_scalars);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, ctx->d_scalars, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    *dE = ctx->h_pinned[0];
    return SYMX_HIP_OK;
}

int symx_hip_problem_gradient_dot(symx_hip_problem p, int32_t vec, double* out)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->finalized) return fail(ctx, SYMX_HIP_ERR_ARG, "problem not finalized");
    double* dv = vec_ptr(p, vec);
    if (!dv || !out) return fail(ctx, SYMX_HIP_ERR_ARG, "bad vector id or null pointer");
    for (auto& pot : p->pots) {
        if (pot.n_elem == 0) continue;
        symx_hip_eval_params prm;
        fill_params(p, pot, prm);  // e_mode SUM: the stored element energies stay
        if (!pot.kC.empty()) {
            int rc = launch_eval(ctx, pot.module, pot.kC, prm);
            if (rc) return rc;
        }
        int rc = launch_eval(ctx, pot.module, pot.kPG, prm, pot.elems_per_group);
        if (rc) return rc;
    }
    // the element Hessians still belong to the last EVAL_PGH; the assembled gradient goes to
    // the FP64 PCG residual vector (scratch outside a solve), not to VEC_GRAD
    launch_grad_assemble(ctx->stream, p->d_goff, p->d_glist, p->d_grow, p->d_G, p->d_r, p->nb);
    launch_dot(ctx->stream, p->n_dofs, p->d_r, dv, ctx->d_scalars, ctx->d_partials);
    HIP_TRY(ctx, hipGetLastError());
    return sync_scalar(ctx, ctx->d_scalars, out);
}

int symx_hip_problem_n_element_hessians(symx_hip_problem p, int64_t* n)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    int64_t c = 0;
    for (auto& pot : p->pots) c += pot.n_elem;
    if (n) *n = c;
    return SYMX_HIP_OK;
}

int symx_hip_problem_element_outputs(symx_hip_problem p, int32_t pot_id, double* G, double* H)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->finalized || pot_id < 0 || pot_id >= (int32_t)p->pots.size()) return fail(ctx, SYMX_HIP_ERR_ARG, "bad potential id");
    auto& pot = p->pots[pot_id];
    // the context's stream is non-blocking: a plain hipMemcpy does not wait for its kernels
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (G) HIP_TRY(ctx, hipMemcpy(G, p->d_G + pot.g_base * 3, sizeof(double) * pot.n_elem * pot.nb * 3, hipMemcpyDeviceToHost));
    if (H) {
        // records of h_stride values on the device, n_ub * 9 here (the pad is dropped)
        const size_t w = (size_t)pot.n_ub * 9, rs = (size_t)pot.h_stride, n = (size_t)pot.n_elem * rs;
        std::vector<double> d;
        std::vector<float> f;
        if (p->h_f32) {
            f.resize(n);
            HIP_TRY(ctx, hipMemcpy(f.data(), h32_of(p, pot), sizeof(float) * n, hipMemcpyDeviceToHost));
        }
        else {
            d.resize(n);
            HIP_TRY(ctx, hipMemcpy(d.data(), p->d_H + pot.h_off, sizeof(double) * n, hipMemcpyDeviceToHost));
        }
        for (size_t e = 0; e < (size_t)pot.n_elem; e++)
            for (size_t i = 0; i < w; i++) H[e * w + i] = p->h_f32 ? (double)f[e * rs + i] : d[e * rs + i];
    }
    return SYMX_HIP_OK;
}

int symx_hip_problem_project(symx_hip_problem p, double eps, int32_t mirror, int64_t* n_changed)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->h_valid) return fail(ctx, SYMX_HIP_ERR_ARG, "no element Hessians (evaluate with EVAL_PGH first)");
    // without n_changed nothing is counted and the call returns without waiting for the device
    unsigned long long* d_count = n_changed ? (unsigned long long*)(ctx->d_scalars + 8) : nullptr;
    if (d_count) HIP_TRY(ctx, hipMemsetAsync(d_count, 0, sizeof(unsigned long long), ctx->stream));
    for (auto& pot : p->pots) {
        // elements of a conditional potential that are inactive have no Hessian to project
        const uint8_t* only = pot.kC.empty() ? nullptr : pot.d_active;
        const int bad = p->h_f32 ? launch_project_ub(ctx->stream, h32_of(p, pot), pot.n_elem, pot.nb, pot.h_stride, eps, mirror, p->d_flags, p->d_plist, ctx->d_counters,
                                                     only, d_count)
                                 : launch_project_ub(ctx->stream, p->d_H + pot.h_off, pot.n_elem, pot.nb, pot.h_stride, eps, mirror, p->d_flags, p->d_plist,
                                                     ctx->d_counters, only, d_count);
        if (bad)
            return fail(ctx, SYMX_HIP_ERR_ARG, "PD projection supports 1-4, 8, 10 and 27 DoF blocks per element");
    }
    HIP_TRY(ctx, hipGetLastError());
    if (!n_changed) return SYMX_HIP_OK;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, d_count, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long c;
    std::memcpy(&c, ctx->h_pinned, 8);
    *n_changed = (int64_t)c;
    return SYMX_HIP_OK;
}

int symx_hip_problem_project_selective(symx_hip_problem p, double eps, int32_t mirror, double threshold, int64_t* n_changed,
                                       int64_t* n_newly_projected, int64_t* n_active_blocks)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->h_valid) return fail(ctx, SYMX_HIP_ERR_ARG, "no element Hessians (evaluate with EVAL_PGH first)");
    if (!(threshold >= 0.0)) return fail(ctx, SYMX_HIP_ERR_ARG, "threshold must be >= 0");
    if (p->pots.size() > 16) return fail(ctx, SYMX_HIP_ERR_ARG, "selective projection supports up to 16 potentials");
    if (!p->d_projected) {
        int rc = dalloc(ctx, &p->d_projected, (size_t)std::max<int64_t>(p->total_elem, 1));
        if (!rc) rc = dalloc(ctx, &p->d_select, (size_t)std::max<int64_t>(p->total_elem, 1));
        if (rc) return rc;
        // the Hessians of the current evaluation are all unprojected so far
        launch_fill_u8(ctx->stream, p->d_projected, p->total_elem);
        launch_fill_u8(ctx->stream, p->d_select, p->total_elem);
    }
    PotRanges pr;
    pr.n = (int32_t)p->pots.size();
    for (int32_t q = 0; q < pr.n; q++) {
        const DevPotential& pot = p->pots[q];
        if (pot.nb > 4 && pot.nb != 8 && pot.nb != 10 && pot.nb != 27)
            return fail(ctx, SYMX_HIP_ERR_ARG, "PD projection supports 1-4, 8, 10 and 27 DoF blocks per element");
        pr.nb[q] = pot.nb;
        pr.g_base[q] = pot.g_base;
        pr.g_end[q] = pot.g_base + (int64_t)pot.n_elem * pot.nb;
        pr.e_base[q] = pot.e_base;
        pr.active[q] = pot.kC.empty() ? nullptr : pot.d_active;
    }
    unsigned long long* cnt = (unsigned long long*)(ctx->d_scalars + 8);  // [changed, newly projected, active blocks]
    HIP_TRY(ctx, hipMemsetAsync(cnt, 0, 3 * sizeof(unsigned long long), ctx->stream));
    launch_ppn_select(ctx->stream, p->d_goff, p->d_glist, p->d_grow, p->d_vec[SYMX_HIP_VEC_GRAD], p->nb, threshold, pr, p->d_projected, p->d_select, cnt + 2);
    for (auto& pot : p->pots) {
        if (p->h_f32)
            launch_project_ub(ctx->stream, h32_of(p, pot), pot.n_elem, pot.nb, pot.h_stride, eps, mirror, p->d_flags, p->d_plist, ctx->d_counters, p->d_select + pot.e_base, cnt);
        else
            launch_project_ub(ctx->stream, p->d_H + pot.h_off, pot.n_elem, pot.nb, pot.h_stride, eps, mirror, p->d_flags, p->d_plist, ctx->d_counters,
                              p->d_select + pot.e_base, cnt);
    }
    launch_ppn_commit(ctx->stream, p->d_projected, p->d_select, p->total_elem, cnt + 1);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, cnt, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long c[3];
    std::memcpy(c, ctx->h_pinned, sizeof(c));
    if (n_changed) *n_changed = (int64_t)c[0];
    if (n_newly_projected) *n_newly_projected = (int64_t)c[1];
    if (n_active_blocks) *n_active_blocks = (int64_t)c[2];
    return SYMX_HIP_OK;
}

int symx_hip_problem_assemble(symx_hip_problem p, int32_t precision)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->h_valid) return fail(ctx, SYMX_HIP_ERR_ARG, "no element Hessians (evaluate with EVAL_PGH first)");
    const SellPattern& S = p->sell;
    const size_t nval = (size_t)S.n_slots * 9;
    const float* h32 = p->h_f32 ? reinterpret_cast<const float*>(p->d_H) : nullptr;
    if (precision == SYMX_HIP_FP64) {
        if (!p->d_val64) {
            int rc = dalloc(ctx, &p->d_val64, nval);
            if (!rc) rc = dalloc(ctx, &p->d_dinv64, (size_t)S.n_rows * 6);
            if (rc) return rc;
        }
        launch_bsr_assemble(ctx->stream, p->n_assemble, p->d_aslot, p->d_mirror, p->d_coff, p->d_clist, p->d_H, h32, p->d_val64, nullptr);
        launch_dinv(ctx->stream, S, p->d_val64, nullptr, p->d_dinv64, nullptr);
    }
    else if (precision == SYMX_HIP_MIXED) {
        if (!p->d_val32) {
            int rc = dalloc(ctx, &p->d_val32, nval);
            if (!rc) rc = dalloc(ctx, &p->d_dinv32, (size_t)S.n_rows * 6);
            const size_t nd = (size_t)p->n_dofs;
            if (!rc) rc = dalloc(ctx, &p->f_x, nd);
            if (!rc) rc = dalloc(ctx, &p->f_r, nd);
            if (!rc) rc = dalloc(ctx, &p->f_z, nd);
            if (!rc) rc = dalloc(ctx, &p->f_p0, nd);
            if (!rc) rc = dalloc(ctx, &p->f_p1, nd);
            if (!rc) rc = dalloc(ctx, &p->f_q, nd);
            if (!rc) rc = dalloc(ctx, &p->f_b, nd);
            if (rc) return rc;
        }
        launch_bsr_assemble(ctx->stream, p->n_assemble, p->d_aslot, p->d_mirror, p->d_coff, p->d_clist, p->d_H, h32, nullptr, p->d_val32);
        launch_dinv(ctx->stream, S, nullptr, p->d_val32, nullptr, p->d_dinv32);
    }
    else {
        return fail(ctx, SYMX_HIP_ERR_ARG, "bad precision");
    }
    HIP_TRY(ctx, hipGetLastError());  // stream-ordered: the solve that follows does not wait for the host
    p->assembled_precision = precision;
    return SYMX_HIP_OK;
}

int symx_hip_problem_export_bsr(symx_hip_problem p, int64_t* nnzb, int64_t* row_ptr, int32_t* cols, double* vals)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->finalized) return fail(ctx, SYMX_HIP_ERR_ARG, "problem not finalized");
    const SellPattern& S = p->sell;
    if (nnzb) *nnzb = S.nnzb;
    if (!row_ptr && !cols && !vals) return SYMX_HIP_OK;
    // assembly is enqueued without waiting; the copies below run outside the context's stream
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<double> v;
    const bool f64 = p->assembled_precision == SYMX_HIP_FP64;
    if (vals) {
        if (p->assembled_precision < 0) return fail(ctx, SYMX_HIP_ERR_ARG, "matrix not assembled");
        v.resize((size_t)S.n_slots * 9);
        if (f64) {
            HIP_TRY(ctx, hipMemcpy(v.data(), p->d_val64, v.size() * 8, hipMemcpyDeviceToHost));
        }
        else {
            std::vector<float> f(v.size());
            HIP_TRY(ctx, hipMemcpy(f.data(), p->d_val32, f.size() * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < f.size(); i++) v[i] = f[i];
        }
    }
    // Output rows with ascending columns (standard block-CSR).
    int64_t k = 0;
    for (int32_t r = 0; r < S.n_rows; r++) {
        if (row_ptr) row_ptr[r] = k;
        const int32_t pos = S.pos_of(r);
        const int64_t base = S.h_slice_ptr[pos >> 6] + (pos & 63);
        std::vector<std::pair<int32_t, int32_t>> order;
        for (int32_t j = 0; j < S.h_row_len[r]; j++) order.push_back({ S.h_col[base + 64 * (int64_t)j], j });
        std::sort(order.begin(), order.end());
        for (auto& cj : order) {
            if (cols) cols[k] = cj.first;
            if (vals) {
                const int64_t t = base + 64 * (int64_t)cj.second;
                for (int q = 0; q < 9; q++) vals[k * 9 + q] = v[f64 ? sell_val_index<double>(t, q) : sell_val_index<float>(t, q)];
            }
            k++;
        }
    }
    if (row_ptr) row_ptr[S.n_rows] = k;
    return SYMX_HIP_OK;
}

// Iterations to enqueue before the next look at the PCG state. The remaining count is
// estimated from the rate since the previous look (err per iteration, geometric; CG speeds
// up as it goes, so the mean rate since the start overestimates it) and from the mean rate,
// taking the smaller; 85% of it is enqueued, at least two and at most 256 iterations. A look
// costs one host round trip, an iteration enqueued past convergence one no-op launch per
// kernel (each about 30 us at the 6M-tet block), so the guess aims just short of the end.
// Iterations past convergence exit at once on the device: the guess only trades launches
// against round trips, and the iterates do not depend on it.
static int next_pcg_batch(const PcgState& h, int batch, double abs_tol, double rel_tol, int32_t prev_it, double prev_err)
{
    double next = 2.0 * batch;
    if (h.it > 0 && h.err0 > 0.0 && h.err > 0.0 && std::isfinite(h.err)) {
        const double ratio = h.err / h.err0;
        const double target = std::max(rel_tol, abs_tol / h.err0);
        if (ratio < 1.0 && target < ratio) {
            double n = h.it * std::log(target / ratio) / std::log(ratio);
            if (h.it > prev_it && prev_err > h.err && h.err > 0.0) {
                const double per_it = std::log(h.err / prev_err) / (h.it - prev_it);
                const double n_recent = std::log(target / ratio) / per_it;
                if (std::isfinite(n_recent) && n_recent > 0.0) n = std::min(n, n_recent);
            }
            if (std::isfinite(n)) next = n < 4.0 ? std::ceil(n) : std::ceil(0.85 * n);
        }
    }
    next = std::min(next, (double)std::max(h.max_iter - h.it, 1));
    return (int)std::max(2.0, std::min(256.0, next));
}

static int run_pcg(symx_hip_problem p, bool fp32, const double* b64, double* x64, double abs_tol, double rel_tol, int32_t max_iter,
                   int32_t stop_on_indef, PcgState& out)
{
    symx_hip_ctx ctx = p->ctx;
    const SellPattern& S = p->sell;
    PcgState st{};
    st.abs_tol = abs_tol;
    st.rel_tol = rel_tol;
    st.max_iter = max_iter;
    st.stop_on_indef = stop_on_indef;
    *ctx->h_pcg = st;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->d_pcg, ctx->h_pcg, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
    if (fp32) launch_pcg_init32(ctx->stream, S, p->d_dinv32, p->f_b, p->f_x, p->f_r, p->f_z, p->f_p0, ctx->d_partials);
    else launch_pcg_init64(ctx->stream, S, p->d_dinv64, b64, x64, p->d_r, p->d_z, p->d_p0, ctx->d_partials);
    // Batches of iterations in a two-deep pipeline: while the device runs one batch, the
    // host reads the state at the end of the previous one (a copy into its own pinned slot,
    // waited on by event) and enqueues the next, so the device does not idle for the round
    // trip; the next batch holds the predicted remainder beyond the batch in flight, and
    // none is enqueued while the one in flight is expected to finish the solve. Iterations
    // past convergence exit at once (done is sticky).
    int32_t k = 1;
    int pending[2] = { 0, 0 };  // iterations enqueued ahead of each slot's state copy
    auto enqueue = [&](int n, int slot) -> int {
        for (int i = 0; i < n; i++, k++) {
            if (fp32)
                launch_pcg_iter32(ctx->stream, S, k, p->d_val32, p->d_dinv32, p->f_x, p->f_r, p->f_z, (k & 1) ? p->f_p0 : p->f_p1,
                                  (k & 1) ? p->f_p1 : p->f_p0, p->f_q, ctx->d_pcg, ctx->d_partials);
            else
                launch_pcg_iter64(ctx->stream, S, k, p->d_val64, p->d_dinv64, x64, p->d_r, p->d_z, (k & 1) ? p->d_p0 : p->d_p1,
                                  (k & 1) ? p->d_p1 : p->d_p0, p->d_q, ctx->d_pcg, ctx->d_partials);
        }
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pcg + slot, ctx->d_pcg, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->pcg_ev[slot], ctx->stream));
        pending[slot] = n;
        return SYMX_HIP_OK;
    };
    int rc = enqueue(8, 0);
    if (!rc) rc = enqueue(8, 1);
    if (rc) return rc;
    int cur = 0, batch = 8;
    int32_t prev_it = 0;
    double prev_err = 0.0;
    while (true) {
        HIP_TRY(ctx, hipEventSynchronize(ctx->pcg_ev[cur]));
        const PcgState h = ctx->h_pcg[cur];
        pending[cur] = 0;
        if (h.done) break;
        const int other = cur ^ 1;
        if (pending[other] == 0) {
            // nothing in flight: the solve is not done, so enqueue the predicted rest now
            batch = next_pcg_batch(h, batch, abs_tol, rel_tol, prev_it, prev_err);
            rc = enqueue(batch, cur);
            if (rc) return rc;
        }
        else {
            const int rest = next_pcg_batch(h, batch, abs_tol, rel_tol, prev_it, prev_err) - pending[other];
            if (rest >= 2) {
                batch = rest;
                rc = enqueue(batch, cur);
                if (rc) return rc;
            }
            cur = other;
        }
        prev_it = h.it;
        prev_err = h.err;
    }
    // the state of the finished solve (later batches leave it unchanged)
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipMemcpy(ctx->h_pcg, ctx->d_pcg, sizeof(PcgState), hipMemcpyDeviceToHost));
    out = *ctx->h_pcg;
    return SYMX_HIP_OK;
}

int symx_hip_problem_solve(symx_hip_problem p, int32_t precision, double abs_tol, double rel_tol, int32_t max_iter, int32_t stop_on_indef,
                           symx_hip_pcg_info* info)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (p->assembled_precision != precision) return fail(ctx, SYMX_HIP_ERR_ARG, "matrix not assembled in the requested precision");
    symx_hip_pcg_info res{};
    double* b = p->d_vec[SYMX_HIP_VEC_RHS];
    double* x = p->d_vec[SYMX_HIP_VEC_DU];
    if (precision == SYMX_HIP_FP64) {
        PcgState st;
        int rc = run_pcg(p, false, b, x, abs_tol, rel_tol, max_iter, stop_on_indef, st);
        if (rc) return rc;
        res.n_iterations = st.it;
        res.converged = st.converged;
        res.found_indefiniteness = st.indef;
        res.error = st.err;
        res.n_refinements = 0;
    }
    else {
        // FP64 iterative refinement around FP32 PCG.
        const SellPattern& S = p->sell;
        double* rr_dev = ctx->d_scalars + 16;
        HIP_TRY(ctx, hipMemsetAsync(x, 0, sizeof(double) * p->n_dofs, ctx->stream));
        // r = b - A*0 = b exactly, r32 = (float) b, rr = ||b||^2
        launch_residual_mixed(ctx->stream, S, p->d_val32, b, x, p->d_r, p->f_b, rr_dev, ctx->d_partials);
        double bb = 0.0;
        int rc = sync_scalar(ctx, rr_dev, &bb);
        if (rc) return rc;
        res.converged = 0;
        if (bb < abs_tol * abs_tol) {
            res.converged = 1;
            res.error = 0.0;
        }
        else {
            const double err0 = std::sqrt(bb / bb);
            double err = err0;
            int total = 0;
            if (err < abs_tol) res.converged = 1;
            const double target = std::max(abs_tol, rel_tol * err0);
            // Each refinement shrinks the FP64 residual by what the FP32 inner solve really
            // achieved (about 0.2 on stiff penalty systems), so allow many refinements and
            // stop when two in a row gain less than 10%.
            int stalled = 0;
            for (int outer = 0; !res.converged && outer < 64 && total < max_iter && stalled < 2; outer++) {
                const double inner_rel = std::min(0.5, std::max(0.5 * target / err, 2e-6));
                PcgState st;
                rc = run_pcg(p, true, nullptr, nullptr, 0.0, inner_rel, max_iter - total, stop_on_indef, st);
                if (rc) return rc;
                total += st.it;
                res.n_refinements++;
                if (st.indef) res.found_indefiniteness = 1;
                if (st.indef && stop_on_indef) {
                    // Keep x from before this inner solve, like the FP64 path keeps its last iterate.
                    break;
                }
                launch_x_update_mixed(ctx->stream, p->n_dofs, x, p->f_x);
                launch_residual_mixed(ctx->stream, S, p->d_val32, b, x, p->d_r, p->f_b, rr_dev, ctx->d_partials);
                double rr = 0.0;
                rc = sync_scalar(ctx, rr_dev, &rr);
                if (rc) return rc;
                const double err_new = std::sqrt(rr / bb);
                stalled = err_new > 0.9 * err ? stalled + 1 : 0;
                err = err_new;
                if (err < abs_tol || err / err0 < rel_tol) res.converged = 1;
            }
            res.n_iterations = total;
            res.error = err;
        }
    }
    if (info) *info = res;
    return SYMX_HIP_OK;
}

int symx_hip_problem_max_abs(symx_hip_problem p, int32_t vec, double* out)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    double* d = vec_ptr(p, vec);
    if (!d) return fail(p->ctx, SYMX_HIP_ERR_ARG, "bad vector id");
    if (vec == SYMX_HIP_VEC_GRAD && p->grad_max_valid) {
        *out = p->grad_max;  // computed by the evaluation that wrote the gradient
        return SYMX_HIP_OK;
    }
    launch_max_abs(p->ctx->stream, p->n_dofs, d, p->ctx->d_scalars, p->ctx->d_partials);
    return sync_scalar(p->ctx, p->ctx->d_scalars, out);
}
int symx_hip_problem_dot(symx_hip_problem p, int32_t a, int32_t b, double* out)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    double* da = vec_ptr(p, a);
    double* db = vec_ptr(p, b);
    if (!da || !db) return fail(p->ctx, SYMX_HIP_ERR_ARG, "bad vector id");
    launch_dot(p->ctx->stream, p->n_dofs, da, db, p->ctx->d_scalars, p->ctx->d_partials);
    return sync_scalar(p->ctx, p->ctx->d_scalars, out);
}
int symx_hip_problem_axpy(symx_hip_problem p, int32_t dst, int32_t a, double s, int32_t b)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    double* dd = vec_ptr(p, dst);
    double* da = vec_ptr(p, a);
    double* db = vec_ptr(p, b);
    if (!dd || !da || !db) return fail(p->ctx, SYMX_HIP_ERR_ARG, "bad vector id");
    if (dst == SYMX_HIP_VEC_GRAD) p->grad_max_valid = false;
    launch_axpy(p->ctx->stream, p->n_dofs, dd, da, s, db);
    HIP_TRY(p->ctx, hipGetLastError());
    return SYMX_HIP_OK;
}
int symx_hip_problem_scale(symx_hip_problem p, int32_t dst, int32_t a, double s)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    double* dd = vec_ptr(p, dst);
    double* da = vec_ptr(p, a);
    if (!dd || !da) return fail(p->ctx, SYMX_HIP_ERR_ARG, "bad vector id");
    if (dst == SYMX_HIP_VEC_GRAD) p->grad_max_valid = false;
    launch_scale(p->ctx->stream, p->n_dofs, dd, da, s);
    HIP_TRY(p->ctx, hipGetLastError());
    return SYMX_HIP_OK;
}

int symx_hip_problem_adjoint_solve(symx_hip_problem p, int32_t precision, double rel_tol, int32_t max_iter, const double* host_rhs,
                                   double* host_lambda, symx_hip_pcg_info* info)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!host_rhs) return fail(ctx, SYMX_HIP_ERR_ARG, "null rhs");
    // assemble the unprojected Hessian of the last evaluation
    int rc = symx_hip_problem_assemble(p, precision);
    if (rc) return rc;
    rc = symx_hip_problem_upload(p, SYMX_HIP_VEC_RHS, host_rhs);
    if (rc) return rc;
    rc = symx_hip_problem_solve(p, precision, 0.0, rel_tol, max_iter, 1, info);
    if (rc) return rc;
    if (host_lambda) return symx_hip_problem_download(p, SYMX_HIP_VEC_DU, host_lambda);
    return SYMX_HIP_OK;
}

int symx_hip_problem_launch_adjoint(symx_hip_problem p, int32_t pot_id, const char* name, int32_t n_out, double* host_out)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->finalized || pot_id < 0 || pot_id >= (int32_t)p->pots.size() || !name || n_out <= 0 || !host_out)
        return fail(ctx, SYMX_HIP_ERR_ARG, "bad adjoint launch arguments");
    DevPotential& pot = p->pots[pot_id];
    if (pot.n_elem == 0) return SYMX_HIP_OK;
    hipFunction_t f;
    int rc = module_function(pot.module, name, &f);
    if (rc) return rc;
    double* d_out = nullptr;
    rc = dalloc(ctx, &d_out, (size_t)pot.n_elem * n_out);
    if (rc) return rc;
    struct
    {
        symx_hip_eval_params prm;
        symx_hip_adjoint_params adj;
    } args;
    static_assert(sizeof(args) == sizeof(symx_hip_eval_params) + sizeof(symx_hip_adjoint_params), "kernel argument layout");
    fill_params(p, pot, args.prm);
    std::memset(&args.adj, 0, sizeof(args.adj));
    for (size_t m = 0; m < pot.maps.size(); m++)
        if (pot.maps[m].is_dof) args.adj.lambda[m] = p->d_vec[SYMX_HIP_VEC_DU] + pot.maps[m].dof_offset;
    args.adj.out = d_out;
    args.adj.n_out = n_out;
    size_t sz = sizeof(args);
    void* config[] = { HIP_LAUNCH_PARAM_BUFFER_POINTER, &args, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END };
    const unsigned grid = (unsigned)((pot.n_elem + SYMX_HIP_BLOCK - 1) / SYMX_HIP_BLOCK);
    hipError_t e = hipModuleLaunchKernel(f, grid, 1, 1, SYMX_HIP_BLOCK, 1, 1, 0, ctx->stream, nullptr, config);
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, d_out, sizeof(double) * pot.n_elem * n_out, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(d_out);
    if (e != hipSuccess) return fail(ctx, SYMX_HIP_ERR_HIP, std::string("adjoint kernel: ") + hipGetErrorString(e));
    return SYMX_HIP_OK;
}

int symx_hip_problem_set_plane_filter(symx_hip_problem p, const int32_t* blocks, int32_t n, int32_t axis, double plane)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (n < 0 || (n > 0 && !blocks) || axis < 0 || axis > 2) return fail(ctx, SYMX_HIP_ERR_ARG, "bad plane filter");
    for (int32_t i = 0; i < n; i++)
        if (blocks[i] < 0 || blocks[i] >= p->nb) return fail(ctx, SYMX_HIP_ERR_ARG, "plane filter block out of range");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // no queued step-filter kernel may still read the old list
    dfree(p->d_filter_blocks);
    p->n_filter_blocks = 0;
    if (n > 0) {
        int rc = dalloc(ctx, &p->d_filter_blocks, (size_t)n);
        if (rc) return rc;
        HIP_TRY(ctx, hipMemcpy(p->d_filter_blocks, blocks, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    }
    p->n_filter_blocks = n;
    p->filter_axis = axis;
    p->filter_plane = plane;
    return SYMX_HIP_OK;
}

int symx_hip_problem_plane_max_step(symx_hip_problem p, double eta, double* t_max)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!t_max || !(eta > 0.0)) return fail(ctx, SYMX_HIP_ERR_ARG, "bad plane step arguments");
    *t_max = 1.0;
    if (p->n_filter_blocks == 0) return SYMX_HIP_OK;
    unsigned long long* d = (unsigned long long*)(ctx->d_scalars + 12);
    launch_plane_max_step(ctx->stream, p->d_filter_blocks, p->n_filter_blocks, p->d_vec[SYMX_HIP_VEC_DOFS], p->d_vec[SYMX_HIP_VEC_DU],
                          p->filter_axis, p->filter_plane, eta, d);
    HIP_TRY(ctx, hipGetLastError());
    return sync_scalar(ctx, (const double*)d, t_max);
}

int symx_hip_loop_eval(symx_hip_ctx ctx, symx_hip_module mod, const char* name, const int32_t* conn, int32_t n_elem, int32_t conn_stride,
                       const symx_hip_map_desc* maps, int32_t n_maps, const uint8_t* active, int32_t n_out, double* host_out)
{
    if (!ctx || !mod || !name || n_elem < 0 || conn_stride <= 0 || n_maps < 0 || n_maps > SYMX_HIP_MAX_MAPS || n_out <= 0 || !host_out ||
        (n_elem > 0 && !conn))
        return fail(ctx, SYMX_HIP_ERR_ARG, "bad loop arguments");
    if (n_elem == 0) return SYMX_HIP_OK;
    hipFunction_t f;
    int rc = module_function(mod, name, &f);
    if (rc) return rc;
    // transposed connectivity [col][elem], as the element kernels read it
    std::vector<int32_t> t((size_t)n_elem * conn_stride);
    for (int64_t e = 0; e < n_elem; e++)
        for (int32_t c = 0; c < conn_stride; c++) t[(size_t)c * n_elem + e] = conn[e * conn_stride + c];
    int32_t* d_conn = nullptr;
    uint8_t* d_active = nullptr;
    double* d_out = nullptr;
    std::vector<double*> d_maps(n_maps, nullptr);
    auto cleanup = [&]() {
        dfree(d_conn);
        dfree(d_active);
        dfree(d_out);
        for (auto& m : d_maps) dfree(m);
    };
    rc = dalloc(ctx, &d_conn, t.size());
    if (!rc && active) rc = dalloc(ctx, &d_active, (size_t)n_elem);
    if (!rc) rc = dalloc(ctx, &d_out, (size_t)n_elem * n_out);
    for (int32_t m = 0; !rc && m < n_maps; m++) {
        if (maps[m].n_values < 0 || (maps[m].n_values > 0 && !maps[m].host)) rc = fail(ctx, SYMX_HIP_ERR_ARG, "bad data map");
        else rc = dalloc(ctx, &d_maps[m], (size_t)std::max<int64_t>(maps[m].n_values, 1));
    }
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpyAsync(d_conn, t.data(), sizeof(int32_t) * t.size(), hipMemcpyHostToDevice, ctx->stream);
    if (!rc && e == hipSuccess && active) e = hipMemcpyAsync(d_active, active, (size_t)n_elem, hipMemcpyHostToDevice, ctx->stream);
    for (int32_t m = 0; !rc && e == hipSuccess && m < n_maps; m++)
        if (maps[m].n_values > 0)
            e = hipMemcpyAsync(d_maps[m], maps[m].host, sizeof(double) * maps[m].n_values, hipMemcpyHostToDevice, ctx->stream);
    if (!rc && e == hipSuccess) {
        struct
        {
            symx_hip_eval_params prm;
            symx_hip_adjoint_params adj;
        } args;
        std::memset(&args, 0, sizeof(args));
        args.prm.conn = d_conn;
        args.prm.active = d_active;
        for (int32_t m = 0; m < n_maps; m++) args.prm.maps[m] = d_maps[m];
        args.prm.n_elem = n_elem;
        args.adj.out = d_out;
        args.adj.n_out = n_out;
        size_t sz = sizeof(args);
        void* config[] = { HIP_LAUNCH_PARAM_BUFFER_POINTER, &args, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END };
        const unsigned grid = (unsigned)((n_elem + SYMX_HIP_BLOCK - 1) / SYMX_HIP_BLOCK);
        e = hipModuleLaunchKernel(f, grid, 1, 1, SYMX_HIP_BLOCK, 1, 1, 0, ctx->stream, nullptr, config);
        if (e == hipSuccess) e = hipMemcpyAsync(host_out, d_out, sizeof(double) * (size_t)n_elem * n_out, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    if (!rc && e != hipSuccess) {
        (void)hipStreamSynchronize(ctx->stream);
        rc = fail(ctx, SYMX_HIP_ERR_HIP, std::string("loop evaluation: ") + hipGetErrorString(e));
    }
    cleanup();
    return rc;
}

int symx_hip_loop_eval_noalloc(symx_hip_ctx ctx, symx_hip_module mod, const char* name, const int32_t* d_conn, int32_t n_elem,
                               int32_t conn_stride, const double* const* d_maps, int32_t n_maps, const uint8_t* d_active, int32_t n_out,
                               double* d_out)
{
    if (!ctx || !mod || !name || n_elem < 0 || conn_stride <= 0 || n_maps < 0 || n_maps > SYMX_HIP_MAX_MAPS || n_out <= 0 ||
        (n_elem > 0 && (!d_conn || !d_out)) || (n_maps > 0 && !d_maps))
        return fail(ctx, SYMX_HIP_ERR_ARG, "bad loop arguments");
    if ((int64_t)n_elem * conn_stride > (int64_t)SYMX_HIP_LOOP_SCRATCH_ENTRIES)
        return fail(ctx, SYMX_HIP_ERR_ARG, "loop too large for the allocation-free path: n_elem * conn_stride exceeds "
                                           "SYMX_HIP_LOOP_SCRATCH_ENTRIES (use symx_hip_loop_eval)");
    if (!ctx->d_loop_scratch) return fail(ctx, SYMX_HIP_ERR_ARG, "context has no loop scratch address");
    if (n_elem == 0) return SYMX_HIP_OK;
    hipFunction_t f;
    int rc = module_function(mod, name, &f);
    if (rc) return rc;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    // the whole range the kernel reads, [0, n_elem * conn_stride), is written first
    launch_transpose_conn(ctx->stream, d_conn, n_elem, conn_stride, ctx->d_loop_scratch);
    struct
    {
        symx_hip_eval_params prm;
        symx_hip_adjoint_params adj;
    } args;
    std::memset(&args, 0, sizeof(args));
    args.prm.conn = ctx->d_loop_scratch;
    args.prm.active = d_active;
    for (int32_t m = 0; m < n_maps; m++) args.prm.maps[m] = d_maps[m];
    args.prm.n_elem = n_elem;
    args.adj.out = d_out;
    args.adj.n_out = n_out;
    size_t sz = sizeof(args);
    void* config[] = { HIP_LAUNCH_PARAM_BUFFER_POINTER, &args, HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz, HIP_LAUNCH_PARAM_END };
    const unsigned grid = (unsigned)((n_elem + SYMX_HIP_BLOCK - 1) / SYMX_HIP_BLOCK);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipModuleLaunchKernel(f, grid, 1, 1, SYMX_HIP_BLOCK, 1, 1, 0, ctx->stream, nullptr, config));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SYMX_HIP_OK;
}

int symx_hip_problem_profile(symx_hip_problem p, int32_t phase, int32_t pot_id, int32_t arg, int32_t reps, double* avg_ms)
{
    if (check_problem(p)) return SYMX_HIP_ERR_ARG;
    symx_hip_ctx ctx = p->ctx;
    if (!p->finalized || reps <= 0) return fail(ctx, SYMX_HIP_ERR_ARG, "problem not finalized or reps <= 0");
    if (pot_id < 0 || pot_id >= (int32_t)p->pots.size()) return fail(ctx, SYMX_HIP_ERR_ARG, "bad potential id");
    DevPotential& pot = p->pots[pot_id];
    const SellPattern& S = p->sell;
    double* scratch = nullptr;
    // phase 1 projects a copy of the element Hessians in their storage precision
    const size_t h_vals = (size_t)pot.n_elem * pot.h_stride;
    const size_t h_bytes = h_vals * (p->h_f32 ? sizeof(float) : sizeof(double));
    const void* h_src = p->h_f32 ? (const void*)h32_of(p, pot) : (const void*)(p->d_H + pot.h_off);
    auto project_scratch = [&]() {
        return p->h_f32 ? launch_project_ub(ctx->stream, reinterpret_cast<float*>(scratch), pot.n_elem, pot.nb, pot.h_stride, 1e-10, 0, p->d_flags, p->d_plist,
                                            ctx->d_counters, nullptr, nullptr)
                        : launch_project_ub(ctx->stream, scratch, pot.n_elem, pot.nb, pot.h_stride, 1e-10, 0, p->d_flags, p->d_plist, ctx->d_counters, nullptr,
                                            nullptr);
    };
    if (phase == 1) {
        if (!p->h_valid) return fail(ctx, SYMX_HIP_ERR_ARG, "no element Hessians");
        int rc = dalloc(ctx, &scratch, std::max<size_t>(h_vals, 1));
        if (rc) return rc;
    }
    if ((phase == 2 || phase == 3) && p->assembled_precision != arg) {
        int rc = symx_hip_problem_assemble(p, arg);
        if (rc) return rc;
    }
    auto launch_once = [&](int i) -> int {
        symx_hip_eval_params prm;
        fill_params(p, pot, prm);
        switch (phase) {
        case 0:
            if (arg == SYMX_HIP_MIXED) {
                if (pot.kPGH32.empty()) return fail(ctx, SYMX_HIP_ERR_ARG, "no FP32-storage Hessian kernel");
                prm.H = reinterpret_cast<double*>(h32_of(p, pot));
                p->h_f32 = true;  // the element Hessians now hold this launch's FP32 values
                return launch_eval(ctx, pot.module, pot.kPGH32, prm, pot.elems_per_group);
            }
            p->h_f32 = false;
            return launch_eval(ctx, pot.module, pot.kPGH, prm, pot.elems_per_group);
        case 4: return launch_eval(ctx, pot.module, pot.kP, prm, pot.elems_per_group);
        case 1: {
            // project a fresh copy of the element Hessians each repetition
            if (hipMemcpyAsync(scratch, h_src, h_bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) return SYMX_HIP_ERR_HIP;
            return project_scratch() ? SYMX_HIP_ERR_ARG : 0;
        }
        case 2:
            launch_bsr_assemble(ctx->stream, p->n_assemble, p->d_aslot, p->d_mirror, p->d_coff, p->d_clist, p->d_H,
                                p->h_f32 ? reinterpret_cast<const float*>(p->d_H) : nullptr, arg == SYMX_HIP_FP64 ? p->d_val64 : nullptr,
                                arg == SYMX_HIP_FP64 ? nullptr : p->d_val32);
            return 0;
        case 3:
            if (arg == SYMX_HIP_FP64)
                launch_pcg_iter64(ctx->stream, S, i + 1, p->d_val64, p->d_dinv64, p->d_x, p->d_r, p->d_z, ((i + 1) & 1) ? p->d_p0 : p->d_p1,
                                  ((i + 1) & 1) ? p->d_p1 : p->d_p0, p->d_q, ctx->d_pcg, ctx->d_partials);
            else
                launch_pcg_iter32(ctx->stream, S, i + 1, p->d_val32, p->d_dinv32, p->f_x, p->f_r, p->f_z, ((i + 1) & 1) ? p->f_p0 : p->f_p1,
                                  ((i + 1) & 1) ? p->f_p1 : p->f_p0, p->f_q, ctx->d_pcg, ctx->d_partials);
            return 0;
        default: return fail(ctx, SYMX_HIP_ERR_ARG, "bad profile phase");
        }
    };
    if (phase == 3) {
        // a PCG run that never converges (zero tolerances) so every launch does work
        PcgState st{};
        st.abs_tol = 0.0;
        st.rel_tol = 0.0;
        st.max_iter = 1 << 30;
        st.stop_on_indef = 0;
        *ctx->h_pcg = st;
        HIP_TRY(ctx, hipMemcpyAsync(ctx->d_pcg, ctx->h_pcg, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
        if (arg == SYMX_HIP_FP64)
            launch_pcg_init64(ctx->stream, S, p->d_dinv64, p->d_vec[SYMX_HIP_VEC_GRAD], p->d_x, p->d_r, p->d_z, p->d_p0, ctx->d_partials);
        else {
            if (!p->f_b) return fail(ctx, SYMX_HIP_ERR_ARG, "mixed buffers missing");
            launch_residual_mixed(ctx->stream, S, p->d_val32, p->d_vec[SYMX_HIP_VEC_GRAD], p->d_vec[SYMX_HIP_VEC_DU], p->d_r, p->f_b,
                                  ctx->d_scalars + 16, ctx->d_partials + 8192);
            launch_pcg_init32(ctx->stream, S, p->d_dinv32, p->f_b, p->f_x, p->f_r, p->f_z, p->f_p0, ctx->d_partials);
        }
    }
    int rc = launch_once(0);  // warm-up
    if (rc) {
        dfree(scratch);
        return rc;
    }
    hipEvent_t e0, e1;
    HIP_TRY(ctx, hipEventCreate(&e0));
    HIP_TRY(ctx, hipEventCreate(&e1));
    auto chk = [&](hipError_t e) {
        if (e != hipSuccess && !rc) rc = fail(ctx, SYMX_HIP_ERR_HIP, std::string("profile: ") + hipGetErrorString(e));
    };
    float total_ms = 0.0f;
    if (phase == 1) {
        // time projection launches only (copies excluded): one event pair per repetition
        for (int i = 0; i < reps && !rc; i++) {
            chk(hipMemcpyAsync(scratch, h_src, h_bytes, hipMemcpyDeviceToDevice, ctx->stream));
            chk(hipEventRecord(e0, ctx->stream));
            project_scratch();
            chk(hipEventRecord(e1, ctx->stream));
            chk(hipEventSynchronize(e1));
            float ms = 0.0f;
            chk(hipEventElapsedTime(&ms, e0, e1));
            total_ms += ms;
        }
    }
    else {
        chk(hipEventRecord(e0, ctx->stream));
        for (int i = 1; i <= reps && !rc; i++) rc = launch_once(i);
        chk(hipEventRecord(e1, ctx->stream));
        chk(hipEventSynchronize(e1));
        chk(hipEventElapsedTime(&total_ms, e0, e1));
    }
    chk(hipEventDestroy(e0));
    chk(hipEventDestroy(e1));
    dfree(scratch);
    HIP_TRY(ctx, hipGetLastError());
    if (rc) return rc;
    if (avg_ms) *avg_ms = (double)total_ms / reps;
    return SYMX_HIP_OK;
}

int symx_hip_batch_eval(symx_hip_ctx ctx, symx_hip_module mod, const char* name, const double* host_in, int64_t n, int32_t n_in,
                        double* host_out, int32_t n_out)
{
    if (!ctx || !mod || !name || n < 0 || n_in < 0 || n_out <= 0) return fail(ctx, SYMX_HIP_ERR_ARG, "bad argument");
    if (n == 0) return SYMX_HIP_OK;
    if (n > INT32_MAX) return fail(ctx, SYMX_HIP_ERR_ARG, "batch too large");
    hipFunction_t f;
    int rc = module_function(mod, name, &f);
    if (rc) return rc;
    double *d_in = nullptr, *d_out = nullptr;
    rc = dalloc(ctx, &d_in, (size_t)std::max<int64_t>(n * n_in, 1));
    if (!rc) rc = dalloc(ctx, &d_out, (size_t)(n * n_out));
    if (rc) {
        dfree(d_in);
        return rc;
    }
    hipError_t e = hipSuccess;
    if (n_in > 0) e = hipMemcpyAsync(d_in, host_in, sizeof(double) * n * n_in, hipMemcpyHostToDevice, ctx->stream);
    int nn = (int)n;
    void* args[] = { &d_in, &d_out, &nn };
    if (e == hipSuccess)
        e = hipModuleLaunchKernel(f, (unsigned)((n + 255) / 256), 1, 1, 256, 1, 1, 0, ctx->stream, args, nullptr);
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, d_out, sizeof(double) * n * n_out, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(d_in);
    dfree(d_out);
    if (e != hipSuccess) return fail(ctx, SYMX_HIP_ERR_HIP, std::string("batch evaluation: ") + hipGetErrorString(e));
    return SYMX_HIP_OK;
}

int symx_hip_project_dense(symx_hip_ctx ctx, double* d_mats, int64_t n_mats, int32_t n, double eps, int32_t mirror, uint8_t* d_changed)
{
    if (!ctx || (!d_mats && n_mats > 0)) return fail(ctx, SYMX_HIP_ERR_ARG, "null argument");
    if (launch_project_dense(ctx->stream, d_mats, n_mats, n, eps, mirror, d_changed))
        return fail(ctx, SYMX_HIP_ERR_ARG, "dense projection supports n in {3, 6, 9, 12, 24, 30, 81}");
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SYMX_HIP_OK;
}

} // extern "C"
