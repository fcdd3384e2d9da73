// fmmhip_noalloc.hip -- fmmhip_dgemm_fmm_noalloc, the FP64 FMM entry point
// that allocates nothing, for callers that cannot allocate device memory at
// call time (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).
//
// It runs the pre-summed pack mode without product split (the same
// kernels, lists and arithmetic as fmmhip_dgemm_fmm with pack 1 and split
// 1, so the same C bit for bit).  What that mode keeps in the per-stream
// workspace -- the coefficient lists and the S_r / T_r leaves -- lives here
// in `scratch`, a file-scope device array of FMMHIP_NOALLOC_SCRATCH_BYTES
// that every GPU the library runs on holds until the process ends.  Host
// code never uses the array's name as an address: at set-up (fmmhip_init,
// or the first call on another device) hipGetSymbolAddress gives each
// device's own copy, kept per device, and only that pointer reaches kernel
// arguments and copies.  The lists are built on the host, without malloc,
// in one pinned staging buffer that the init set-up allocates, and copied
// into the scratch on the caller's stream; an event recorded after that
// copy guards the buffer's reuse.
//
// Every call on one GPU shares that GPU's scratch and finds the previous
// call's contents in it (the summation stage writes every leaf element the
// kernel reads, so nothing needs clearing).  Calls on one stream are
// ordered; calls that overlap on one GPU (different streams, or host
// threads with different streams) race on the scratch.
#include <hip/hip_runtime.h>
#include <mutex>
#include <string.h>

#include "fmmhip_host.h"

static __device__ __align__(256) unsigned char scratch[FMMHIP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 16;
// The first part of the scratch holds the lists, the rest the leaves.
constexpr size_t kListBytes = (size_t)8 << 20;
constexpr size_t kLeafBytes = (size_t)FMMHIP_NOALLOC_SCRATCH_BYTES - kListBytes;
static_assert(FMMHIP_NOALLOC_SCRATCH_BYTES > kListBytes, "scratch holds the lists and the leaves");

struct PerDevice {
    void* scratch = nullptr;       // this device's copy of `scratch` (hipGetSymbolAddress)
    hipEvent_t listed = nullptr;   // recorded after this device's last list copy from g_host
};
PerDevice g_dev[kMaxDevices];
char* g_host = nullptr;            // pinned (portable) host image of one call's lists
int g_host_dev = -1;               // device whose copy read g_host last
std::mutex g_mu;                   // set-up and the host image

// Set-up of the current device (g_mu held): its scratch address and event.
int setup_locked(int dev)
{
    PerDevice& d = g_dev[dev];
    if (d.scratch) return FMMHIP_SUCCESS;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return FMMHIP_ERR_HIP;
    }
    if (!d.listed && hipEventCreateWithFlags(&d.listed, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        d.listed = nullptr;
        return FMMHIP_ERR_HIP;
    }
    d.scratch = p;
    return FMMHIP_SUCCESS;
}

// Lists of one call, built into the pinned host image and copied into the
// device's scratch on the call's stream.
struct ScratchSource : DescSource {
    PerDevice* d;
    int dev;
    int get(fmmhip_alg_t alg, const fmm_geom_t& g, bool presum, bool mat, bool a_mn, bool b_mn, hipStream_t st,
            DescView* v) override
    {
        if (!presum || mat) return FMMHIP_ERR_UNSUPPORTED;
        // the previous call's copy out of the host image has completed
        if (g_host_dev >= 0 && hipEventSynchronize(g_dev[g_host_dev].listed) != hipSuccess) return FMMHIP_ERR_HIP;
        size_t used = 0;
        int rc = desc_pack_presum_into(alg, g, a_mn, b_mn, g_host, d->scratch, kListBytes, &used, v);
        if (rc) return FMMHIP_ERR_TOO_LARGE;
        if (hipMemcpyAsync(d->scratch, g_host, used, hipMemcpyHostToDevice, st) != hipSuccess) return FMMHIP_ERR_HIP;
        if (hipEventRecord(d->listed, st) != hipSuccess) return FMMHIP_ERR_HIP;
        g_host_dev = this->dev;
        return FMMHIP_SUCCESS;
    }
};

}  // namespace

// Called once by fmmhip_init for the device current at initialisation: the
// pinned staging buffer, and that device's scratch address and event.
int fmmhip_noalloc_setup(void)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FMMHIP_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_host) {
        void* h = nullptr;
        if (hipHostMalloc(&h, kListBytes, hipHostMallocPortable) != hipSuccess) {
            (void)hipGetLastError();
            return FMMHIP_ERR_NOMEM;
        }
        g_host = (char*)h;
    }
    return setup_locked(dev);
}

extern "C" int fmmhip_dgemm_fmm_noalloc(fmmhip_alg_t alg, long m, long n, long k, double alpha, const double* A,
                                        long rsA, long csA, const double* B, long rsB, long csB, double beta,
                                        double* C, long rsC, long csC, void* stream)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    if (!alg) return FMMHIP_ERR_INVALID;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FMMHIP_ERR_HIP;
    hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)fmmhip_default_stream();
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_host) return FMMHIP_ERR_HIP;      // the init set-up failed
    if ((rc = setup_locked(dev))) return rc;
    PerDevice& d = g_dev[dev];
    // no gather buffer here: operands with neither stride 1 are refused
    const bool work = m > 0 && n > 0 && k > 0 && alpha != 0.0;
    if (work && ((rsA != 1 && csA != 1) || (rsB != 1 && csB != 1))) return FMMHIP_ERR_UNSUPPORTED;
    ScratchSource src;
    src.d = &d;
    src.dev = dev;
    std::unique_lock<std::mutex> elk = enqueue_lock();
    Prepared P;
    rc = prepare_call(alg, Kind::D64, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, 1, 1, st, &P,
                      &src);
    if (rc) return rc;
    if (P.ws_bytes > kLeafBytes) return FMMHIP_ERR_TOO_LARGE;
    if (P.ws_bytes) bind_workspace(&P, (char*)d.scratch + kListBytes);
    P.untimed = true;              // the launch timing creates events
    return launch_prepared(&P);
}
