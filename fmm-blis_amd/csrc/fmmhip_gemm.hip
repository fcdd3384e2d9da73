// fmmhip_gemm.hip -- host side of the fused FMM kernel: per-geometry
// coefficient descriptors on the device, launch-shape selection, and the
// fmmhip_{d,s}gemm{,_fmm} entry points of the C ABI.  The kernel itself is
// in fmmhip_kernel.h, instantiated per datatype in fmmhip_gemm_*.hip.
#include <hip/hip_runtime.h>
#include <mutex>
#include <vector>
#include <string.h>
#include <stdlib.h>

#include "fmmhip_kernel.h"

using fmmk::KArgs;
using fmmk::BM;
using fmmk::BN;

namespace {

// ---- host side -------------------------------------------------------------

struct DevDesc {
    fmm_geom_t geo;
    int32_t* ptr_h = nullptr;
    fmm_entry_t* ent_h = nullptr;
    long nent = 0;
    void* dev = nullptr;         // [ptr (3*(R+1) int32), pad, entries]
    size_t ent_off = 0;
    unsigned long last_use = 0;
};

struct DevCache {
    std::mutex mu;
    std::vector<DevDesc*> items;
    unsigned long clock = 0;
};

constexpr size_t kCacheCap = 64;

bool same_geo(const fmm_geom_t& a, const fmm_geom_t& b)
{
    return a.m == b.m && a.n == b.n && a.k == b.k && a.rsA == b.rsA && a.csA == b.csA &&
           a.rsB == b.rsB && a.csB == b.csB && a.rsC == b.rsC && a.csC == b.csC;
}

void free_desc(DevDesc* d)
{
    if (d->dev) (void)hipFree(d->dev);
    free(d->ptr_h);
    free(d->ent_h);
    delete d;
}

// Returns the device copy of the algorithm's entries for geometry g,
// uploading (stream-ordered) when it is not cached.
int get_desc(fmmhip_alg_t alg, const fmm_geom_t& g, hipStream_t stream, const int** dptr,
             const fmm_entry_t** dent)
{
    if (!alg->dev_cache) {
        static std::mutex create_mu;
        std::lock_guard<std::mutex> lk(create_mu);
        if (!alg->dev_cache) alg->dev_cache = new DevCache();
    }
    DevCache* cache = (DevCache*)alg->dev_cache;
    std::lock_guard<std::mutex> lk(cache->mu);
    cache->clock++;
    for (DevDesc* d : cache->items) {
        if (same_geo(d->geo, g)) {
            d->last_use = cache->clock;
            *dptr = (const int*)d->dev;
            *dent = (const fmm_entry_t*)((char*)d->dev + d->ent_off);
            return FMMHIP_SUCCESS;
        }
    }
    if (cache->items.size() >= kCacheCap) {
        // evict the least recently used descriptor; kernels may still read it
        size_t victim = 0;
        for (size_t i = 1; i < cache->items.size(); i++)
            if (cache->items[i]->last_use < cache->items[victim]->last_use) victim = i;
        if (hipDeviceSynchronize() != hipSuccess) return FMMHIP_ERR_HIP;
        free_desc(cache->items[victim]);
        cache->items.erase(cache->items.begin() + victim);
    }
    DevDesc* d = new DevDesc();
    d->geo = g;
    int rc = fmm_build_entries(alg, &g, &d->ptr_h, &d->ent_h, &d->nent);
    if (rc) { delete d; return rc; }
    size_t ptr_bytes = sizeof(int32_t) * 3 * (size_t)(alg->R + 1);
    d->ent_off = (ptr_bytes + 255) & ~(size_t)255;
    size_t total = d->ent_off + sizeof(fmm_entry_t) * (size_t)(d->nent + 1);
    if (hipMalloc(&d->dev, total) != hipSuccess) { (void)hipGetLastError(); d->dev = nullptr; free_desc(d); return FMMHIP_ERR_NOMEM; }
    if (hipMemcpyAsync(d->dev, d->ptr_h, ptr_bytes, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync((char*)d->dev + d->ent_off, d->ent_h, sizeof(fmm_entry_t) * (size_t)d->nent,
                       hipMemcpyHostToDevice, stream) != hipSuccess) {
        free_desc(d);
        return FMMHIP_ERR_HIP;
    }
    d->last_use = cache->clock;
    cache->items.push_back(d);
    *dptr = (const int*)d->dev;
    *dent = (const fmm_entry_t*)((char*)d->dev + d->ent_off);
    return FMMHIP_SUCCESS;
}

// Pairs along the contiguous dimension can be loaded as one 16-byte vector
// when the base, the leading dimension and every part offset keep them
// 16-byte aligned (pairs start at even indices of the tile).
bool pairs_aligned(const void* base, size_t esz, long ld, const fmm_entry_t* ent, long e0, long e1)
{
    const size_t pair = 2 * esz;
    if (((uintptr_t)base) % pair) return false;
    if (ld % 2) return false;
    for (long e = e0; e < e1; e++)
        if (ent[e].off % 2) return false;
    return true;
}

enum class Kind { D64, D_F32LEAF, S32 };

int gemm_common(fmmhip_alg_t alg, Kind kind, long m, long n, long k, double alpha,
                const void* A, long rsA, long csA, const void* B, long rsB, long csB, double beta,
                void* C, long rsC, long csC, void* stream)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    if (!alg || m < 0 || n < 0 || k < 0) return FMMHIP_ERR_INVALID;
    hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)fmmhip_default_stream();
    if (m == 0 || n == 0) return FMMHIP_SUCCESS;
    if (!C) return FMMHIP_ERR_INVALID;
    if (k == 0 || alpha == 0.0) {
        if (kind == Kind::S32) return fmmhip_scale_s((float*)C, m, n, rsC, csC, (float)beta, st);
        return fmmhip_scale_d((double*)C, m, n, rsC, csC, beta, st);
    }
    if (!A || !B) return FMMHIP_ERR_INVALID;
    // operand layouts: one of the strides must be 1
    bool a_mn, b_mn;
    if (rsA == 1) a_mn = true; else if (csA == 1) a_mn = false; else return FMMHIP_ERR_UNSUPPORTED;
    if (csB == 1) b_mn = true; else if (rsB == 1) b_mn = false; else return FMMHIP_ERR_UNSUPPORTED;
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;
    // per-thread tile offsets are 32-bit: (tile extent) x (stride) must fit
    const long kStrideMax = 1L << 23;
    if ((a_mn ? csA : rsA) > kStrideMax || (b_mn ? rsB : csB) > kStrideMax || rsC > kStrideMax ||
        csC > kStrideMax || rsC < 0 || csC < 0)
        return FMMHIP_ERR_UNSUPPORTED;

    fmm_geom_t g;
    memset(&g, 0, sizeof(g));
    g.m = m; g.n = n; g.k = k;
    g.lm = (m + alg->mt - 1) / alg->mt;
    g.lk = (k + alg->kt - 1) / alg->kt;
    g.ln = (n + alg->nt - 1) / alg->nt;
    g.rsA = rsA; g.csA = csA; g.rsB = rsB; g.csB = csB; g.rsC = rsC; g.csC = csC;

    const int* dptr = nullptr;
    const fmm_entry_t* dent = nullptr;
    rc = get_desc(alg, g, st, &dptr, &dent);
    if (rc) return rc;

    // vector-load eligibility from the host copy of the entries
    DevCache* cache = (DevCache*)alg->dev_cache;
    bool va = false, vb = false;
    // launch shape: the algorithm's setting, else one part in flight for
    // classical-like algorithms and two for those with multi-part products
    int shape = alg->shape;
    if (shape < 0) shape = (alg->max_nnz_u > 1 || alg->max_nnz_v > 1) ? 1 : 0;
    {
        std::lock_guard<std::mutex> lk(cache->mu);
        for (DevDesc* d : cache->items) {
            if (!same_geo(d->geo, g)) continue;
            size_t esz = kind == Kind::S32 ? 4 : 8;
            va = pairs_aligned(A, esz, a_mn ? csA : rsA, d->ent_h, 0, alg->nnz_u);
            vb = pairs_aligned(B, esz, b_mn ? rsB : csB, d->ent_h, alg->nnz_u, alg->nnz_u + alg->nnz_v);
            break;
        }
    }

    KArgs ka;
    ka.A = A; ka.B = B; ka.C = C;
    ka.lda = a_mn ? csA : rsA;
    ka.ldb = b_mn ? rsB : csB;
    ka.rsC = rsC; ka.csC = csC;
    ka.lm = (int)g.lm; ka.lk = (int)g.lk; ka.ln = (int)g.ln;
    ka.min_m = (int)(m / alg->mt); ka.min_k = (int)(k / alg->kt); ka.min_n = (int)(n / alg->nt);
    ka.R = alg->R;
    ka.tiles_m = (int)((g.lm + BM - 1) / BM);
    ka.tiles_n = (int)((g.ln + BN - 1) / BN);
    ka.ptr = dptr;
    ka.ent = dent;
    ka.alpha = alpha;
    ka.beta = beta;
    ka.va = va;
    ka.vb = vb;
    long nwg = (long)ka.tiles_m * ka.tiles_n;
    if (nwg <= 0) return FMMHIP_SUCCESS;
    if (nwg > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;

    const bool c_col = rsC == 1;
    switch (kind) {
    case Kind::D64: return fmmk_launch_d64(ka, shape, a_mn, b_mn, c_col, (int)nwg, st);
    case Kind::D_F32LEAF: return fmmk_launch_d_f32leaf(ka, shape, a_mn, b_mn, c_col, (int)nwg, st);
    case Kind::S32: return fmmk_launch_s32(ka, shape, a_mn, b_mn, c_col, (int)nwg, st);
    }
    return FMMHIP_ERR_INVALID;
}

fmmhip_alg_t classical_alg()
{
    static std::once_flag once;
    static fmmhip_alg_t alg = nullptr;
    std::call_once(once, [] { fmmhip_alg_create_classical(&alg); });
    return alg;
}

}  // namespace

extern "C" int fmmhip_alg_set_shape(fmmhip_alg_t alg, int shape)
{
    if (!alg || shape < -1 || shape > 1) return FMMHIP_ERR_INVALID;
    alg->shape = shape;
    return FMMHIP_SUCCESS;
}

extern "C" void fmmhip_alg_release_cache(struct fmmhip_alg_s* alg)
{
    if (!alg || !alg->dev_cache) return;
    DevCache* cache = (DevCache*)alg->dev_cache;
    if (!cache->items.empty()) (void)hipDeviceSynchronize();
    for (DevDesc* d : cache->items) free_desc(d);
    delete cache;
    alg->dev_cache = nullptr;
}

extern "C" int fmmhip_dgemm_fmm(fmmhip_alg_t alg, fmmhip_leaf_t leaf, long m, long n, long k,
                                double alpha, const double* A, long rsA, long csA, const double* B,
                                long rsB, long csB, double beta, double* C, long rsC, long csC,
                                void* stream)
{
    Kind kind;
    switch (leaf) {
    case FMMHIP_LEAF_NATIVE: kind = Kind::D64; break;
    case FMMHIP_LEAF_F32: kind = Kind::D_F32LEAF; break;
    default: return FMMHIP_ERR_UNSUPPORTED;
    }
    return gemm_common(alg, kind, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, stream);
}

extern "C" int fmmhip_sgemm_fmm(fmmhip_alg_t alg, long m, long n, long k, float alpha, const float* A,
                                long rsA, long csA, const float* B, long rsB, long csB, float beta,
                                float* C, long rsC, long csC, void* stream)
{
    return gemm_common(alg, Kind::S32, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, stream);
}

extern "C" int fmmhip_dgemm(long m, long n, long k, double alpha, const double* A, long rsA, long csA,
                            const double* B, long rsB, long csB, double beta, double* C, long rsC,
                            long csC, void* stream)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    return gemm_common(classical_alg(), Kind::D64, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C,
                       rsC, csC, stream);
}

extern "C" int fmmhip_sgemm(long m, long n, long k, float alpha, const float* A, long rsA, long csA,
                            const float* B, long rsB, long csB, float beta, float* C, long rsC,
                            long csC, void* stream)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    return gemm_common(classical_alg(), Kind::S32, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C,
                       rsC, csC, stream);
}
