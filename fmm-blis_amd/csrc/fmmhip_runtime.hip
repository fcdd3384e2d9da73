// fmmhip_runtime.hip -- device set-up, memory, streams, events, seeded fills
// and reductions behind the fmmhip_* C ABI (include/fmm_blis_amd.h).
//
// The reference has no device layer; its counterparts are BLIS's own memory
// and thread set-up (blis/frame/base/bli_init.c, bli_obj.c).  Errors are
// returned as fmmhip_status_t codes, never by aborting.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include <thread>
#include <string.h>
#include <stdio.h>

#include "fmm_internal.h"

static std::once_flag g_init_once;
static int g_init_status = FMMHIP_ERR_NODEVICE;
static hipStream_t g_stream = nullptr;
static int g_cus = 0;
static char g_arch[64] = {0};

static void do_init()
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { g_init_status = FMMHIP_ERR_NODEVICE; return; }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) { g_init_status = FMMHIP_ERR_HIP; return; }
    strncpy(g_arch, prop.gcnArchName, sizeof(g_arch) - 1);
    g_cus = prop.multiProcessorCount;
    if (strncmp(g_arch, "gfx950", 6) != 0) {
        fprintf(stderr, "fmm_blis_amd: device arch '%s' is not gfx950\n", g_arch);
        g_init_status = FMMHIP_ERR_NODEVICE;
        return;
    }
    if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
        g_init_status = FMMHIP_ERR_HIP;
        return;
    }
    // the no-allocation entry point's scratch on this device (a failure
    // here is reported by that entry point alone)
    (void)fmmhip_noalloc_setup();
    g_init_status = FMMHIP_SUCCESS;
}

extern "C" int fmmhip_init(void)
{
    std::call_once(g_init_once, do_init);
    return g_init_status;
}

extern "C" void* fmmhip_default_stream(void)
{
    fmmhip_init();
    return (void*)g_stream;
}

extern "C" const char* fmmhip_status_string(int s)
{
    switch (s) {
    case FMMHIP_SUCCESS: return "success";
    case FMMHIP_ERR_INVALID: return "invalid argument";
    case FMMHIP_ERR_NOMEM: return "out of memory";
    case FMMHIP_ERR_HIP: return "HIP runtime error";
    case FMMHIP_ERR_UNSUPPORTED: return "unsupported";
    case FMMHIP_ERR_NODEVICE: return "no gfx950 device";
    case FMMHIP_ERR_IO: return "coefficient file unreadable or malformed";
    case FMMHIP_ERR_TOO_LARGE: return "problem exceeds the no-allocation scratch";
    default: return "unknown status";
    }
}

extern "C" int fmmhip_device_arch(char* buf, size_t len)
{
    int rc = fmmhip_init();
    if (buf && len) { strncpy(buf, g_arch, len - 1); buf[len - 1] = 0; }
    return rc;
}

extern "C" int fmmhip_device_cu_count(int* cus)
{
    int rc = fmmhip_init();
    if (cus) *cus = g_cus;
    return rc;
}

#define HIP_OK(x) do { if ((x) != hipSuccess) return FMMHIP_ERR_HIP; } while (0)
#define INIT_OR_RETURN() do { int _s = fmmhip_init(); if (_s) return _s; } while (0)

static inline hipStream_t S(void* s) { return s ? (hipStream_t)s : g_stream; }

extern "C" int fmmhip_malloc(void** p, size_t bytes)
{
    INIT_OR_RETURN();
    if (!p) return FMMHIP_ERR_INVALID;
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) { *p = nullptr; (void)hipGetLastError(); return FMMHIP_ERR_NOMEM; }
    return FMMHIP_SUCCESS;
}
extern "C" int fmmhip_free(void* p) { INIT_OR_RETURN(); HIP_OK(hipFree(p)); return 0; }
// The library's streams are non-blocking (they do not wait on the null
// stream), and a pageable hipMemcpy may return before its DMA lands, so the
// copies below synchronise the device on both sides: on return the data is
// in place and visible to every stream.
// Large pageable copies go through pinned bounce buffers, in parallel:
// kStageThreads host threads each own a contiguous slice of the copy, two
// pinned chunks and a stream, and overlap their host memcpy of one chunk
// with the DMA of the other.  A single pageable hipMemcpy runs one staging
// pipeline (about 24 GB/s measured); the copies of a call are what the
// drop-in on host objects waits for.
namespace {

constexpr int kStageThreads = 8;
constexpr size_t kStageChunk = (size_t)16 << 20;
constexpr size_t kParallelMin = (size_t)64 << 20;

struct StageWorker {
    hipStream_t st = nullptr;
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
};

std::mutex g_sw_mu;                          // one parallel copy at a time
StageWorker g_sw[kStageThreads];
int g_sw_state = 0;                          // 0 not tried, 1 ready, -1 unavailable

int sw_ready()
{
    if (g_sw_state) return g_sw_state > 0;
    g_sw_state = -1;
    for (StageWorker& w : g_sw) {
        if (hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking) != hipSuccess) return 0;
        for (int i = 0; i < 2; i++) {
            if (hipHostMalloc(&w.buf[i], kStageChunk, hipHostMallocDefault) != hipSuccess) return 0;
            if (hipEventCreateWithFlags(&w.ev[i], hipEventDisableTiming) != hipSuccess) return 0;
        }
    }
    g_sw_state = 1;
    return 1;
}

// host -> device: chunk c is copied into pinned buffer c&1 once that buffer's
// previous DMA has completed, then its DMA is queued.
bool slice_h2d(StageWorker& w, char* d, const char* s, size_t bytes)
{
    int i = 0;
    for (size_t off = 0; off < bytes; off += kStageChunk, i ^= 1) {
        const size_t n = std::min(kStageChunk, bytes - off);
        if (hipEventSynchronize(w.ev[i]) != hipSuccess) return false;
        memcpy(w.buf[i], s + off, n);
        if (hipMemcpyAsync(d + off, w.buf[i], n, hipMemcpyHostToDevice, w.st) != hipSuccess) return false;
        if (hipEventRecord(w.ev[i], w.st) != hipSuccess) return false;
    }
    return hipStreamSynchronize(w.st) == hipSuccess;
}

// device -> host: the DMA of chunk c+1 is queued before chunk c is copied out.
bool slice_d2h(StageWorker& w, char* d, const char* s, size_t bytes)
{
    const size_t nc = (bytes + kStageChunk - 1) / kStageChunk;
    auto issue = [&](size_t c) {
        const size_t off = c * kStageChunk, n = std::min(kStageChunk, bytes - off);
        return hipMemcpyAsync(w.buf[c & 1], s + off, n, hipMemcpyDeviceToHost, w.st) == hipSuccess &&
               hipEventRecord(w.ev[c & 1], w.st) == hipSuccess;
    };
    if (nc && !issue(0)) return false;
    for (size_t c = 0; c < nc; c++) {
        if (c + 1 < nc && !issue(c + 1)) return false;
        if (hipEventSynchronize(w.ev[c & 1]) != hipSuccess) return false;
        const size_t off = c * kStageChunk;
        memcpy(d + off, w.buf[c & 1], std::min(kStageChunk, bytes - off));
    }
    return hipStreamSynchronize(w.st) == hipSuccess;
}

// Splits the copy into kStageThreads page-aligned slices; false if the
// pinned path is unavailable (the caller then uses a plain hipMemcpy).
bool parallel_copy(void* d, const void* s, size_t b, bool h2d, int* rc)
{
    std::lock_guard<std::mutex> lk(g_sw_mu);
    if (!sw_ready()) return false;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const size_t per = ((b + kStageThreads - 1) / kStageThreads + 4095) & ~(size_t)4095;
    bool ok[kStageThreads];
    std::thread th[kStageThreads];
    auto work = [&](int t) {
        const size_t off = std::min(b, per * t), n = std::min(per, b - off);
        if (hipSetDevice(dev) != hipSuccess) { ok[t] = false; return; }
        ok[t] = h2d ? slice_h2d(g_sw[t], (char*)d + off, (const char*)s + off, n)
                    : slice_d2h(g_sw[t], (char*)d + off, (const char*)s + off, n);
    };
    for (int t = 0; t < kStageThreads; t++) {
        ok[t] = true;
        try {
            th[t] = std::thread(work, t);
        } catch (...) {
            work(t);                 // no thread available: this slice on the caller's thread
        }
    }
    for (std::thread& x : th)
        if (x.joinable()) x.join();
    *rc = FMMHIP_SUCCESS;
    for (int t = 0; t < kStageThreads; t++)
        if (!ok[t]) *rc = FMMHIP_ERR_HIP;
    return true;
}

}  // namespace

extern "C" int fmmhip_memcpy_h2d(void* d, const void* s, size_t b)
{
    INIT_OR_RETURN();
    HIP_OK(hipDeviceSynchronize());
    int rc = 0;
    if (b >= kParallelMin && parallel_copy(d, s, b, true, &rc)) return rc;
    HIP_OK(hipMemcpy(d, s, b, hipMemcpyHostToDevice));
    HIP_OK(hipDeviceSynchronize());
    return 0;
}
extern "C" int fmmhip_memcpy_d2h(void* d, const void* s, size_t b)
{
    INIT_OR_RETURN();
    HIP_OK(hipDeviceSynchronize());
    int rc = 0;
    if (b >= kParallelMin && parallel_copy(d, s, b, false, &rc)) return rc;
    HIP_OK(hipMemcpy(d, s, b, hipMemcpyDeviceToHost));
    HIP_OK(hipDeviceSynchronize());
    return 0;
}
extern "C" int fmmhip_memcpy_d2d(void* d, const void* s, size_t b, void* stream)
{
    INIT_OR_RETURN();
    HIP_OK(hipMemcpyAsync(d, s, b, hipMemcpyDeviceToDevice, S(stream)));
    return 0;
}
// A plain streaming copy kernel (16 bytes per lane per iteration): the
// reference workload of known size for calibrating HBM byte counters.
__global__ __launch_bounds__(256) void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) dst[i] = src[i];
}

extern "C" int fmmhip_copy_kernel(void* d, const void* s, size_t b, void* stream)
{
    INIT_OR_RETURN();
    if (!d || !s || b % 16 || ((uintptr_t)d | (uintptr_t)s) % 16) return FMMHIP_ERR_INVALID;
    const long n = (long)(b / 16);
    if (!n) return 0;
    const long blocks = std::min<long>((n + 255) / 256, 256L * 8 * 4);
    hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, S(stream), (const uint4*)s, (uint4*)d, n);
    HIP_OK(hipGetLastError());
    return 0;
}

// ---- measured peaks -----------------------------------------------------------

// Streaming kernels for the device's practical HBM ceilings: UNR independent
// 16-byte non-temporal loads in flight per thread before their stores
// (copy), loads only (read: XOR-folded, one conditional store keeps them),
// stores only (write).  The copy16 kernel above stays the byte-counter
// calibration workload; these measure bandwidth.
typedef unsigned int hbm_u4 __attribute__((ext_vector_type(4)));

template <int UNR>
__global__ __launch_bounds__(256) void hbm_copy_kernel(const hbm_u4* __restrict__ src, hbm_u4* __restrict__ dst, long n)
{
    const long stride = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (UNR - 1) * stride < n; i += UNR * stride) {
        hbm_u4 v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < UNR; u++) __builtin_nontemporal_store(v[u], dst + i + u * stride);
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

template <int UNR>
__global__ __launch_bounds__(256) void hbm_read_kernel(const hbm_u4* __restrict__ src, hbm_u4* __restrict__ out, long n)
{
    const long stride = (long)gridDim.x * blockDim.x;
    hbm_u4 acc = {0u, 0u, 0u, 0u};
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (UNR - 1) * stride < n; i += UNR * stride) {
        hbm_u4 v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < UNR; u++) acc ^= v[u];
    }
    for (; i < n; i += stride) acc ^= src[i];
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B9u) out[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void hbm_write_kernel(hbm_u4* __restrict__ dst, long n)
{
    const long stride = (long)gridDim.x * blockDim.x;
    const hbm_u4 v = {1u, 2u, 3u, (unsigned)threadIdx.x};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        __builtin_nontemporal_store(v, dst + i);
}

extern "C" int fmmhip_measure_hbm(double* copy_gbs, double* read_gbs, double* write_gbs)
{
    INIT_OR_RETURN();
    hipStream_t st = g_stream;
    const size_t cb = (size_t)4 << 30;          // 4 GiB per buffer: far beyond the caches
    const long n = (long)(cb / 16);
    void *x = nullptr, *y = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FMMHIP_ERR_HIP;
    float best[3] = {1e30f, 1e30f, 1e30f};
    if (hipMalloc(&x, cb) != hipSuccess || hipMalloc(&y, cb) != hipSuccess) { rc = FMMHIP_ERR_NOMEM; goto done; }
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
    if (hipMemsetAsync(x, 1, cb, st) != hipSuccess) goto done;
    // grids of 1, 2 and 4 blocks per CU (the best of them is reported), each
    // timed three times after a warm-up
    for (int mult = 1; mult <= 4; mult *= 2) {
        const unsigned blocks = (unsigned)(g_cus * mult);
        for (int kind = 0; kind < 3; kind++)
            for (int rep = 0; rep < 4; rep++) {
                if (hipEventRecord(e0, st) != hipSuccess) goto done;
                if (kind == 0)
                    hipLaunchKernelGGL(hbm_copy_kernel<4>, dim3(blocks), dim3(256), 0, st, (const hbm_u4*)x,
                                       (hbm_u4*)y, n);
                else if (kind == 1)
                    hipLaunchKernelGGL(hbm_read_kernel<8>, dim3(blocks), dim3(256), 0, st, (const hbm_u4*)x,
                                       (hbm_u4*)y, n);
                else
                    hipLaunchKernelGGL(hbm_write_kernel, dim3(blocks), dim3(256), 0, st, (hbm_u4*)y, n);
                if (hipGetLastError() != hipSuccess) goto done;
                float ms = 0.f;
                if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
                if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
                if (rep) best[kind] = std::min(best[kind], ms);
            }
    }
    if (copy_gbs) *copy_gbs = 2.0 * (double)cb / (best[0] * 1e-3) / 1e9;
    if (read_gbs) *read_gbs = (double)cb / (best[1] * 1e-3) / 1e9;
    if (write_gbs) *write_gbs = (double)cb / (best[2] * 1e-3) / 1e9;
    rc = FMMHIP_SUCCESS;
done:
    (void)hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    return rc;
}

typedef double peak_d4 __attribute__((ext_vector_type(4)));

// Eight independent accumulator chains of v_mfma_f64_16x16x4_f64 per wave,
// no memory traffic: the FP64 matrix-core rate the device sustains (four
// waves per SIMD, accumulators kept in the architectural registers).  The
// sum is stored (one double per block) so that the work is not removed.
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(4, 4)))
void mfma_f64_peak_kernel(double* __restrict__ out, int iters)
{
    peak_d4 acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = (peak_d4){0.0, 0.0, 0.0, 0.0};
    const double a = 1.0 + 1e-9 * threadIdx.x, b = 1.0 - 1e-9 * threadIdx.x;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int j = 0; j < 8; j++) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
    }
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < 8; j++) t += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

extern "C" int fmmhip_measure_peaks(double* fp64_mfma_tflops, double* copy_gbs)
{
    INIT_OR_RETURN();
    hipStream_t st = g_stream;
    const int blocks = g_cus * 4;               // 4 waves per block: 4 waves per SIMD
    const int iters = 40000;
    double* out = nullptr;
    void *x = nullptr, *y = nullptr;
    const size_t cb = (size_t)4 << 30;          // 4 GiB copied
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FMMHIP_ERR_HIP;
    float ms = 0.f, best_mfma = 1e30f, best_copy = 1e30f;
    if (hipMalloc(&out, sizeof(double) * blocks) != hipSuccess) return FMMHIP_ERR_NOMEM;
    if (hipMalloc(&x, cb) != hipSuccess || hipMalloc(&y, cb) != hipSuccess) { rc = FMMHIP_ERR_NOMEM; goto done; }
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
    if (hipMemsetAsync(x, 0, cb, st) != hipSuccess) goto done;
    for (int rep = 0; rep < 4; rep++) {         // the first of each is a warm-up
        if (hipEventRecord(e0, st) != hipSuccess) goto done;
        hipLaunchKernelGGL(mfma_f64_peak_kernel, dim3(blocks), dim3(256), 0, st, out, iters);
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
        if (rep) best_mfma = std::min(best_mfma, ms);
        if (hipEventRecord(e0, st) != hipSuccess) goto done;
        if (fmmhip_copy_kernel(y, x, cb, st)) goto done;
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
        if (rep) best_copy = std::min(best_copy, ms);
    }
    if (fp64_mfma_tflops) *fp64_mfma_tflops = (double)blocks * 4 * iters * 8 * 2048.0 / (best_mfma * 1e-3) / 1e12;
    if (copy_gbs) *copy_gbs = 2.0 * (double)cb / (best_copy * 1e-3) / 1e9;
    rc = FMMHIP_SUCCESS;
    if (copy_gbs) {
        // the streaming copy (several loads in flight) where it beats the plain one
        double c = 0.0;
        if (fmmhip_measure_hbm(&c, nullptr, nullptr) == FMMHIP_SUCCESS && c > *copy_gbs) *copy_gbs = c;
    }
done:
    (void)hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    if (out) (void)hipFree(out);
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    return rc;
}

extern "C" int fmmhip_memset(void* p, int v, size_t b)
{
    INIT_OR_RETURN();
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemset(p, v, b));
    HIP_OK(hipDeviceSynchronize());
    return 0;
}
extern "C" int fmmhip_stream_create(void** s)
{
    INIT_OR_RETURN();
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = (void*)st;
    return 0;
}
extern "C" int fmmhip_stream_destroy(void* s) { INIT_OR_RETURN(); HIP_OK(hipStreamDestroy((hipStream_t)s)); return 0; }
extern "C" int fmmhip_stream_sync(void* s) { INIT_OR_RETURN(); HIP_OK(hipStreamSynchronize(S(s))); return 0; }
extern "C" int fmmhip_device_sync(void) { INIT_OR_RETURN(); HIP_OK(hipDeviceSynchronize()); return 0; }
extern "C" int fmmhip_event_create(void** e)
{
    INIT_OR_RETURN();
    hipEvent_t ev;
    HIP_OK(hipEventCreate(&ev));
    *e = (void*)ev;
    return 0;
}
extern "C" int fmmhip_event_destroy(void* e) { INIT_OR_RETURN(); HIP_OK(hipEventDestroy((hipEvent_t)e)); return 0; }
extern "C" int fmmhip_event_record(void* e, void* s) { INIT_OR_RETURN(); HIP_OK(hipEventRecord((hipEvent_t)e, S(s))); return 0; }
extern "C" int fmmhip_event_elapsed_ms(float* ms, void* a, void* b)
{
    INIT_OR_RETURN();
    HIP_OK(hipEventSynchronize((hipEvent_t)b));
    HIP_OK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return 0;
}
extern "C" int fmmhip_mem_info(size_t* f, size_t* t)
{
    INIT_OR_RETURN();
    HIP_OK(hipMemGetInfo(f, t));
    return 0;
}

extern "C" double fmmhip_ws_bound(void)
{
    static std::once_flag once;
    static double bound = 0.0;
    std::call_once(once, [] {
        size_t f = 0, t = 0;
        if (fmmhip_init() == 0 && hipMemGetInfo(&f, &t) == hipSuccess) bound = 0.5 * (double)t;
    });
    return bound;
}

extern "C" int fmmhip_pointer_is_device(const void* p)
{
    if (fmmhip_init()) return -1;
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) { (void)hipGetLastError(); return 0; }
    return (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) ? 1 : 0;
}

// ---- seeded fills (bit-identical to the CPU rule) --------------------------

__device__ __forceinline__ uint64_t smix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t counter_bits(uint64_t base, uint64_t idx)
{
    return smix(base + (idx + 1) * 0x9E3779B97F4A7C15ull);
}

static uint64_t stream_base(uint64_t seed, uint64_t sid)
{
    uint64_t z = seed ^ (sid * 0xD1B54A32D192ED03ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <int MODE, typename T>
__global__ void fill_kernel(T* X, long m, long n, long ld, uint64_t base, int amp)
{
    long total = m * n;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        long j = idx / m, i = idx - j * m;
        uint64_t z = counter_bits(base, (uint64_t)idx) >> 11;
        double v;
        if (MODE == 0) v = 2.0 * ((double)z * 0x1.0p-53) - 1.0;
        else v = (double)((long)(z % (uint64_t)(2 * amp + 1)) - amp);
        X[i + j * ld] = (T)v;
    }
}

static int grid_for(long total)
{
    long g = (total + 255) / 256;
    return (int)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

extern "C" int fmmhip_fill_uniform(double* X, long m, long n, long ld, uint64_t seed, uint64_t sid, void* s)
{
    INIT_OR_RETURN();
    if (m < 0 || n < 0 || ld < m || (m * n && !X)) return FMMHIP_ERR_INVALID;
    if (!m || !n) return 0;
    hipLaunchKernelGGL((fill_kernel<0, double>), dim3(grid_for(m * n)), dim3(256), 0, S(s), X, m, n, ld, stream_base(seed, sid), 0);
    HIP_OK(hipGetLastError());
    return 0;
}
extern "C" int fmmhip_fill_uniform_f32(float* X, long m, long n, long ld, uint64_t seed, uint64_t sid, void* s)
{
    INIT_OR_RETURN();
    if (m < 0 || n < 0 || ld < m || (m * n && !X)) return FMMHIP_ERR_INVALID;
    if (!m || !n) return 0;
    hipLaunchKernelGGL((fill_kernel<0, float>), dim3(grid_for(m * n)), dim3(256), 0, S(s), X, m, n, ld, stream_base(seed, sid), 0);
    HIP_OK(hipGetLastError());
    return 0;
}
extern "C" int fmmhip_fill_int(double* X, long m, long n, long ld, uint64_t seed, uint64_t sid, int amp, void* s)
{
    INIT_OR_RETURN();
    if (m < 0 || n < 0 || ld < m || amp < 0 || (m * n && !X)) return FMMHIP_ERR_INVALID;
    if (!m || !n) return 0;
    hipLaunchKernelGGL((fill_kernel<1, double>), dim3(grid_for(m * n)), dim3(256), 0, S(s), X, m, n, ld, stream_base(seed, sid), amp);
    HIP_OK(hipGetLastError());
    return 0;
}

__global__ void scale_rc_kernel(double* X, long m, long n, long ld, const double* rs, const double* cs)
{
    long total = m * n;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        long j = idx / m, i = idx - j * m;
        double v = X[i + j * ld];
        if (rs) v *= rs[i];
        if (cs) v *= cs[j];
        X[i + j * ld] = v;
    }
}

extern "C" int fmmhip_scale_rows_cols(double* X, long m, long n, long ld, const double* rs, const double* cs, void* s)
{
    INIT_OR_RETURN();
    if (m < 0 || n < 0 || ld < m) return FMMHIP_ERR_INVALID;
    if (!m || !n) return 0;
    hipLaunchKernelGGL(scale_rc_kernel, dim3(grid_for(m * n)), dim3(256), 0, S(s), X, m, n, ld, rs, cs);
    HIP_OK(hipGetLastError());
    return 0;
}

// C <- beta*C over a general-strided m x n matrix; beta == 0 writes zeros.
template <typename T>
__global__ void scale_kernel(T* C, long m, long n, long rs, long cs, T beta)
{
    long total = m * n;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        long j = idx / m, i = idx - j * m;
        T* p = C + i * rs + j * cs;
        *p = beta == (T)0 ? (T)0 : beta * *p;
    }
}

extern "C" int fmmhip_scale_d(double* C, long m, long n, long rs, long cs, double beta, void* s)
{
    INIT_OR_RETURN();
    if (!m || !n || beta == 1.0) return 0;
    hipLaunchKernelGGL(scale_kernel<double>, dim3(grid_for(m * n)), dim3(256), 0, S(s), C, m, n, rs, cs, beta);
    HIP_OK(hipGetLastError());
    return 0;
}
extern "C" int fmmhip_scale_s(float* C, long m, long n, long rs, long cs, float beta, void* s)
{
    INIT_OR_RETURN();
    if (!m || !n || beta == 1.0f) return 0;
    hipLaunchKernelGGL(scale_kernel<float>, dim3(grid_for(m * n)), dim3(256), 0, S(s), C, m, n, rs, cs, beta);
    HIP_OK(hipGetLastError());
    return 0;
}

// Y := X for a general-strided m x n matrix X (element (i,j) at
// X[i*rs + j*cs]) into column-major Y (ld = m).  Consecutive threads write
// consecutive elements of Y; the reads follow X's strides.
template <typename T>
__global__ void gather_kernel(const T* __restrict__ X, long m, long n, long rs, long cs, T* __restrict__ Y)
{
    const long total = m * n;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const long j = idx / m, i = idx - j * m;
        Y[idx] = X[i * rs + j * cs];
    }
}

extern "C" int fmmhip_gather(int is_double, const void* X, long m, long n, long rs, long cs, void* Y, void* s)
{
    INIT_OR_RETURN();
    if (!m || !n) return 0;
    if (is_double)
        hipLaunchKernelGGL(gather_kernel<double>, dim3(grid_for(m * n)), dim3(256), 0, S(s), (const double*)X, m, n,
                           rs, cs, (double*)Y);
    else
        hipLaunchKernelGGL(gather_kernel<float>, dim3(grid_for(m * n)), dim3(256), 0, S(s), (const float*)X, m, n,
                           rs, cs, (float*)Y);
    HIP_OK(hipGetLastError());
    return 0;
}

// ---- product split: temporaries and their reduction -------------------------

// Zeroes `n16` 16-byte words (the G temporaries, contiguous and aligned).
__global__ void zero16_kernel(uint4* __restrict__ T, long n16)
{
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long)gridDim.x * blockDim.x) T[i] = z;
}

// C := beta*C + T_0 + T_1 + ... + T_{G-1} element by element, in that order
// (beta == 0 does not read C).  The walk's fast index is the dimension
// with the smaller stride, so both C and the temporaries coalesce.
template <typename T>
__global__ void split_reduce_kernel(T* __restrict__ C, long inner, long outer, long si, long so, T beta,
                                    const T* __restrict__ tmp, long elems, int groups)
{
    const long total = inner * outer;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const long o = idx / inner, i = idx - o * inner;
        const long off = i * si + o * so;
        T v = beta == (T)0 ? (T)0 : beta * C[off];
        for (int g = 0; g < groups; g++) v += tmp[(long)g * elems + off];
        C[off] = v;
    }
}

extern "C" int fmmhip_split_zero(int is_double, void* T, long elems, int groups, void* s)
{
    INIT_OR_RETURN();
    const long bytes = elems * groups * (is_double ? 8 : 4);
    if (bytes <= 0) return 0;
    if ((bytes & 15) || ((uintptr_t)T & 15)) return FMMHIP_ERR_INVALID;
    const long n16 = bytes / 16;
    hipLaunchKernelGGL(zero16_kernel, dim3(grid_for(n16)), dim3(256), 0, S(s), (uint4*)T, n16);
    HIP_OK(hipGetLastError());
    return 0;
}

extern "C" int fmmhip_split_reduce(int is_double, void* C, long m, long n, long rs, long cs, double beta,
                                   const void* T, long elems, int groups, void* s)
{
    INIT_OR_RETURN();
    if (m <= 0 || n <= 0) return 0;
    const bool by_row = rs <= cs;
    const long inner = by_row ? m : n, outer = by_row ? n : m, si = by_row ? rs : cs, so = by_row ? cs : rs;
    if (is_double)
        hipLaunchKernelGGL(split_reduce_kernel<double>, dim3(grid_for(m * n)), dim3(256), 0, S(s), (double*)C, inner,
                           outer, si, so, beta, (const double*)T, elems, groups);
    else
        hipLaunchKernelGGL(split_reduce_kernel<float>, dim3(grid_for(m * n)), dim3(256), 0, S(s), (float*)C, inner,
                           outer, si, so, (float)beta, (const float*)T, elems, groups);
    HIP_OK(hipGetLastError());
    return 0;
}

// ---- Frobenius norms (FP64, deterministic two-stage reduction) -------------

template <bool DIFF>
__global__ void fro_kernel(const double* X, const double* Y, long m, long n, long ldx, long ldy, double* part)
{
    __shared__ double sh[256];
    double s = 0.0;
    long total = m * n;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        long j = idx / m, i = idx - j * m;
        double d = X[i + j * ldx];
        if (DIFF) d -= Y[i + j * ldy];
        s += d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

static int fro_impl(const double* X, const double* Y, long m, long n, long ldx, long ldy, double* out, bool diff)
{
    INIT_OR_RETURN();
    if (!out || m < 0 || n < 0) return FMMHIP_ERR_INVALID;
    *out = 0.0;
    if (!m || !n) return 0;
    HIP_OK(hipDeviceSynchronize());   // X, Y may come from any stream
    const int nb = 1024;
    double* part = nullptr;
    if (hipMalloc(&part, nb * sizeof(double)) != hipSuccess) return FMMHIP_ERR_NOMEM;
    if (diff) hipLaunchKernelGGL(fro_kernel<true>, dim3(nb), dim3(256), 0, g_stream, X, Y, m, n, ldx, ldy, part);
    else hipLaunchKernelGGL(fro_kernel<false>, dim3(nb), dim3(256), 0, g_stream, X, Y, m, n, ldx, ldy, part);
    double h[nb];
    int rc = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
        hipStreamSynchronize(g_stream) != hipSuccess) rc = FMMHIP_ERR_HIP;
    (void)hipFree(part);
    if (rc) return rc;
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += h[i];
    *out = sqrt(s);
    return 0;
}

extern "C" int fmmhip_fro_diff(const double* X, const double* Y, long m, long n, long ldx, long ldy, double* out)
{
    return fro_impl(X, Y, m, n, ldx, ldy, out, true);
}
extern "C" int fmmhip_fro(const double* X, long m, long n, long ld, double* out)
{
    return fro_impl(X, nullptr, m, n, ld, ld, out, false);
}
