// fmmhip_gemm.hip -- the fused FMM kernel for gfx950 and its launchers.
//
// One kernel computes C := beta*C + alpha * sum_r W_r (x) (S_r * T_r) with
//   S_r = sum U[(i,p)][r] A_ip,  T_r = sum V[(p,j)][r] B_pj
// for any composed <m~,k~,n~;R> algorithm; the classical GEMM is the
// 1x1x1, R = 1 algorithm through the same code.
//
// Reference counterparts (dnparikh/fmm-blis):
//   pack stage   plugin/ref_kernels/bli_packm_fmm_ref.c:89-147 -- the
//                weighted sum of the nonzero parts is formed in registers
//                while the tile streams from HBM and is written to LDS once
//   leaf         the native BLIS micro-kernel called at
//                plugin/ref_kernels/bli_gemm_fmm_ref.c:85-97 -- FP64 MFMA
//                (v_mfma_f64_16x16x4_f64) out of LDS
//   W epilogue   plugin/ref_kernels/bli_gemm_fmm_ref.c:101-120 -- each
//                accumulator tile is added into every C part with a nonzero
//                W coefficient, clipped to the part; no M_r is stored
//   R loop       frame/base/bli_strassen_ab.c:248-294 -- inside the kernel
//
// Work decomposition: one workgroup per BM x BN tile position of the leaf,
// looping over all R products.  It is the only writer of that tile position
// in every C part, so the read-modify-write of C is race-free without
// atomics and deterministic.  The first product that writes a part applies
// beta (beta == 0 never reads C).
#include <hip/hip_runtime.h>
#include <mutex>
#include <vector>
#include <string.h>
#include <stdlib.h>

#include "fmm_internal.h"

namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d2_t __attribute__((ext_vector_type(2)));
typedef float f2_t __attribute__((ext_vector_type(2)));

constexpr int BM = 128;          // tile rows (leaf M)
constexpr int BN = 128;          // tile cols (leaf N)
constexpr int BK = 16;           // k per LDS stage
constexpr int NT = 512;          // 8 waves
constexpr int WAVES_PER_EU = 4;  // 2 workgroups per CU
constexpr int WAVES_M = 2;       // wave grid 2 x 4, 64 x 32 per wave
constexpr int WAVES_N = 4;
constexpr int WTM = BM / WAVES_M;
constexpr int WTN = BN / WAVES_N;
constexpr int FI = WTM / 16;     // 16x16 MFMA blocks per wave along M
constexpr int FJ = WTN / 16;     // ... along N
constexpr int LDS_ROW = BM + 16; // padded LDS row (elements)
constexpr int PAIRS = BM * BK / 2 / NT;   // staged pairs per thread per operand
constexpr int MN_PAIRS = BM / 2;          // pairs along mn in one k row
constexpr int K_ROWS = NT / MN_PAIRS;     // k rows per pass (mn-contiguous)
constexpr int K_PAIRS = BK / 2;           // pairs along k in one mn column
constexpr int MN_COLS = NT / K_PAIRS;     // mn columns per pass (k-contiguous)
static_assert(BM == BN, "one staging shape serves both operands");
static_assert(PAIRS * NT * 2 == BM * BK, "staging covers the tile");

struct KArgs {
    const void* A;
    const void* B;
    void* C;
    long lda;                    // the non-unit stride of A (A_MN: A(i,p) at A[i + p*lda])
    long ldb;                    // the non-unit stride of B (B_MN: B(p,j) at B[p*ldb + j])
    long rsC, csC;
    int lm, lk, ln;              // leaf dims
    int min_m, min_k, min_n;     // smallest part extents: tiles inside them need no bounds checks
    int R;
    int tiles_m, tiles_n;
    const int* ptr;              // u, v, w CSR pointers, 3*(R+1)
    const fmm_entry_t* ent;
    double alpha, beta;
};

// Leaf arithmetic traits.
template <typename TL> struct Leaf;
// Accumulator layout of the 16x16 output block, lane l, register q:
//   FP64 (v_mfma_f64_16x16x4_f64):  row l/16 + 4q,    col l%16
//   FP32 (v_mfma_f32_16x16x4_f32):  row 4(l/16) + q,  col l%16
// (both checked on the device with known operands).
template <> struct Leaf<double> {
    static constexpr int ROW_LANE = 1;   // row step per lane group (l/16)
    static constexpr int ROW_REG = 4;    // row step per register q
    typedef d4_t acc_t;
    static __device__ __forceinline__ acc_t mfma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Leaf<float> {
    static constexpr int ROW_LANE = 4;
    static constexpr int ROW_REG = 1;
    typedef f4_t acc_t;
    static __device__ __forceinline__ acc_t mfma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <typename TI> struct Pair;
template <> struct Pair<double> { typedef d2_t t; };
template <> struct Pair<float> { typedef f2_t t; };

// Staging of one operand tile (BM or BN along "mn", BK along k) for one
// k-step: the weighted sum over the product's nonzero parts, held in
// registers as PAIRS pairs per thread.  The first part's raw values are
// loaded before the MFMA block of the current k-step and scaled after it,
// so those loads are in flight while the matrix cores run; further parts of
// the same product are loaded and accumulated after the block.
//   MN_CONTIG: the operand is contiguous along mn (pairs along mn),
//              otherwise along k (pairs along k).
//   VEC2:      pairs are 16-byte aligned and loaded as one vector.
//   IS_B:      the entry's (rows, cols) are (k, mn) rather than (mn, k).
// Addresses are a uniform tile pointer (scalar registers) plus a per-thread
// 32-bit offset fixed for the whole kernel.
template <typename TI, bool MN_CONTIG, bool VEC2, bool IS_B>
struct Stage {
    double v[PAIRS][2];
    int lmn0, lk0;               // tile-local coordinates of the thread's pair 0
    int off0;                    // its element offset in the tile

    // pair s sits at (lmn0, lk0 + K_ROWS*s) when contiguous along mn,
    // at (lmn0 + MN_COLS*s, lk0) when contiguous along k
    static constexpr int dmn(int s) { return MN_CONTIG ? 0 : MN_COLS * s; }
    static constexpr int dk(int s) { return MN_CONTIG ? K_ROWS * s : 0; }

    __device__ __forceinline__ void init(int tid, long ld)
    {
        if (MN_CONTIG) {
            lmn0 = 2 * (tid % MN_PAIRS);
            lk0 = tid / MN_PAIRS;
            off0 = lmn0 + lk0 * (int)ld;
        } else {
            lk0 = 2 * (tid % K_PAIRS);
            lmn0 = tid / K_PAIRS;
            off0 = lmn0 * (int)ld + lk0;
        }
    }

    // Raw values of part e over the tile at (mn0, k0); zero outside the
    // part's extent.  `interior`: the whole tile lies inside every part.
    __device__ __forceinline__ void load(const TI* __restrict__ base, const fmm_entry_t& e, long ld,
                                         int mn0, int k0, bool interior, double (&x)[PAIRS][2]) const
    {
        const TI* t = base + e.off + (MN_CONTIG ? (long)mn0 + (long)k0 * ld : (long)mn0 * ld + k0) + off0;
        typedef typename Pair<TI>::t pair_t;
        const int ld32 = (int)ld;
        if (interior) {
#pragma unroll
            for (int s = 0; s < PAIRS; s++) {
                const TI* p = t + (MN_CONTIG ? dk(s) * ld32 : dmn(s) * ld32);
                if (VEC2) {
                    pair_t q = *(const pair_t*)p;
                    x[s][0] = (double)q[0];
                    x[s][1] = (double)q[1];
                } else {
                    x[s][0] = (double)p[0];
                    x[s][1] = (double)p[1];
                }
            }
            return;
        }
        const int emn = (IS_B ? e.cols : e.rows) - mn0;   // valid extent in tile coords
        const int ek = (IS_B ? e.rows : e.cols) - k0;
#pragma unroll
        for (int s = 0; s < PAIRS; s++) {
            const TI* p = t + (MN_CONTIG ? dk(s) * ld32 : dmn(s) * ld32);
            const int mn_0 = lmn0 + dmn(s), k_0 = lk0 + dk(s);
            // element 1 is one step along the contiguous dimension
            const bool ok0 = mn_0 < emn && k_0 < ek;
            const bool ok1 = MN_CONTIG ? (mn_0 + 1 < emn && k_0 < ek) : (mn_0 < emn && k_0 + 1 < ek);
            double x0 = 0.0, x1 = 0.0;
            if (VEC2 && ok1) {
                pair_t q = *(const pair_t*)p;
                x0 = (double)q[0];
                x1 = (double)q[1];
            } else {
                if (ok0) x0 = (double)p[0];
                if (ok1) x1 = (double)p[1];
            }
            x[s][0] = x0;
            x[s][1] = x1;
        }
    }

    __device__ __forceinline__ void begin(const TI* base, const fmm_entry_t& e, long ld, int mn0, int k0,
                                          bool interior)
    {
        load(base, e, ld, mn0, k0, interior, v);
    }

    __device__ __forceinline__ void scale(double c)
    {
        if (c != 1.0) {
#pragma unroll
            for (int s = 0; s < PAIRS; s++) { v[s][0] *= c; v[s][1] *= c; }
        }
    }

    // v += coef * part[e]
    __device__ __forceinline__ void add(const TI* base, const fmm_entry_t& e, long ld, int mn0, int k0,
                                        bool interior)
    {
        double x[PAIRS][2];
        load(base, e, ld, mn0, k0, interior, x);
        const double c = e.coef;
#pragma unroll
        for (int s = 0; s < PAIRS; s++) {
            v[s][0] = fma(c, x[s][0], v[s][0]);
            v[s][1] = fma(c, x[s][1], v[s][1]);
        }
    }

    // Writes the staged sums into an LDS tile laid out [BK][LDS_ROW].
    template <typename TL>
    __device__ __forceinline__ void store(TL* lds) const
    {
        TL* q = lds + lk0 * LDS_ROW + lmn0;
#pragma unroll
        for (int s = 0; s < PAIRS; s++) {
            TL* qs = q + dk(s) * LDS_ROW + dmn(s);
            if (MN_CONTIG) {
                qs[0] = (TL)v[s][0];
                qs[1] = (TL)v[s][1];
            } else {
                qs[0] = (TL)v[s][0];
                qs[LDS_ROW] = (TL)v[s][1];
            }
        }
    }
};

// XCD-aware tile order: consecutive workgroup ids are dealt round-robin to
// the 8 XCDs, so ids are first regrouped so that each XCD receives a
// contiguous run of tiles, then the run is walked in groups of 8 tile rows
// (column-major inside a group) so that concurrently running workgroups on
// one XCD share A row panels and B column panels in that XCD's L2.
__device__ __forceinline__ void tile_of(int bid, int tiles_m, int tiles_n, int& tm, int& tn)
{
    const int nwg = tiles_m * tiles_n;
    int lin = bid;
    if ((nwg & 7) == 0) lin = (bid & 7) * (nwg >> 3) + (bid >> 3);
    const int GROUP = 8;
    const int per_group = GROUP * tiles_n;
    const int g = lin / per_group;
    const int first_m = g * GROUP;
    const int gsz = min(tiles_m - first_m, GROUP);
    const int in = lin - g * per_group;
    tm = first_m + in % gsz;
    tn = in / gsz;
}

template <typename TI, typename TL, typename TO, bool A_MN, bool B_MN, bool VA, bool VB, bool C_COL>
__global__ __attribute__((amdgpu_flat_work_group_size(NT, NT), amdgpu_waves_per_eu(WAVES_PER_EU, WAVES_PER_EU)))
void fmm_gemm_kernel(KArgs a)
{
    typedef typename Leaf<TL>::acc_t acc_t;
    __shared__ TL sA[2][BK * LDS_ROW];
    __shared__ TL sB[2][BK * LDS_ROW];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave % WAVES_M) * WTM;
    const int wn = (wave / WAVES_M) * WTN;

    int tm, tn;
    tile_of(blockIdx.x, a.tiles_m, a.tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;

    const TI* __restrict__ Ab = (const TI*)a.A;
    const TI* __restrict__ Bb = (const TI*)a.B;
    TO* __restrict__ Cb = (TO*)a.C;
    const int R = a.R;
    const int* __restrict__ uptr = a.ptr;
    const int* __restrict__ vptr = a.ptr + (R + 1);
    const int* __restrict__ wptr = a.ptr + 2 * (R + 1);
    const fmm_entry_t* __restrict__ ent = a.ent;
    const int nk = (a.lk + BK - 1) / BK;
    const long lda = a.lda, ldb = a.ldb;

    // tiles inside every part need no bounds checks
    const bool a_rows_in = m0 + BM <= a.min_m;
    const bool b_cols_in = n0 + BN <= a.min_n;
    const bool c_in = a_rows_in && b_cols_in;

    Stage<TI, A_MN, VA, false> stA;
    Stage<TI, B_MN, VB, true> stB;
    stA.init(tid, lda);
    stB.init(tid, ldb);

    // epilogue coordinates of this lane's accumulator element (i, j, q):
    // row = m0 + wm + 16 i + ROW_LANE (lane/16) + ROW_REG q, col = n0 + wn + 16 j + lane%16
    const int er0 = wm + Leaf<TL>::ROW_LANE * (lane >> 4);
    const int ec0 = wn + (lane & 15);


    for (int r = 0; r < R; r++) {
        const int u0 = uptr[r], u1 = uptr[r + 1];
        const int v0 = vptr[r], v1 = vptr[r + 1];

        acc_t acc[FI][FJ];
#pragma unroll
        for (int i = 0; i < FI; i++)
#pragma unroll
            for (int j = 0; j < FJ; j++) acc[i][j] = (acc_t){0, 0, 0, 0};

        // Per product, the first nonzero part of each operand is loaded
        // ahead of the MFMA block; the rest are folded in after it.
        const fmm_entry_t ea0 = ent[u0];
        const fmm_entry_t eb0 = ent[v0];

        // prologue: stage k-step 0
        {
            const bool kin = BK <= a.min_k;
            stA.begin(Ab, ea0, lda, m0, 0, a_rows_in && kin);
            stB.begin(Bb, eb0, ldb, n0, 0, b_cols_in && kin);
            stA.scale(ea0.coef);
            for (int e = u0 + 1; e < u1; e++) stA.add(Ab, ent[e], lda, m0, 0, a_rows_in && kin);
            stB.scale(eb0.coef);
            for (int e = v0 + 1; e < v1; e++) stB.add(Bb, ent[e], ldb, n0, 0, b_cols_in && kin);
            stA.template store<TL>(sA[0]);
            stB.template store<TL>(sB[0]);
        }
        __syncthreads();

        for (int kt = 0; kt < nk; kt++) {
            const int cur = kt & 1;
            const bool more = kt + 1 < nk;
            const int k0 = (kt + 1) * BK;
            const bool kin = k0 + BK <= a.min_k;
            if (more) {
                stA.begin(Ab, ea0, lda, m0, k0, a_rows_in && kin);
                stB.begin(Bb, eb0, ldb, n0, k0, b_cols_in && kin);
            }
            const TL* la = sA[cur];
            const TL* lb = sB[cur];
#pragma unroll
            for (int ks = 0; ks < BK / 4; ks++) {
                const int krow = (4 * ks + (lane >> 4)) * LDS_ROW;
                TL fa[FI], fb[FJ];
#pragma unroll
                for (int i = 0; i < FI; i++) fa[i] = la[krow + wm + 16 * i + (lane & 15)];
#pragma unroll
                for (int j = 0; j < FJ; j++) fb[j] = lb[krow + wn + 16 * j + (lane & 15)];
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) acc[i][j] = Leaf<TL>::mfma(fa[i], fb[j], acc[i][j]);
            }
            if (more) {
                stA.scale(ea0.coef);
                for (int e = u0 + 1; e < u1; e++) stA.add(Ab, ent[e], lda, m0, k0, a_rows_in && kin);
                stB.scale(eb0.coef);
                for (int e = v0 + 1; e < v1; e++) stB.add(Bb, ent[e], ldb, n0, k0, b_cols_in && kin);
                stA.template store<TL>(sA[cur ^ 1]);
                stB.template store<TL>(sB[cur ^ 1]);
            }
            __syncthreads();
        }

        // W epilogue: C_part += alpha*w*acc for every nonzero w of product r;
        // the part's first writer applies beta (beta == 0: C is not read).
        const int w0 = wptr[r], w1 = wptr[r + 1];
        for (int e = w0; e < w1; e++) {
            const fmm_entry_t we = ent[e];
            const double lam = a.alpha * we.coef;
            const bool first = we.first != 0;
            const double beta = a.beta;
            const long rsC = C_COL ? 1 : a.rsC;
            const int rs32 = (int)rsC, cs32 = (int)a.csC;
            // lane's element (i, j, q) is at tile row er0 + 16 i + ROW_REG q, col ec0 + 16 j
            TO* ct = Cb + we.off + (long)(m0 + er0) * rsC + (long)(n0 + ec0) * a.csC;
            const int erows = we.rows - m0 - er0, ecols = we.cols - n0 - ec0;
#pragma unroll
            for (int j = 0; j < FJ; j++) {
#pragma unroll
                for (int i = 0; i < FI; i++) {
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int dr = 16 * i + Leaf<TL>::ROW_REG * q, dc = 16 * j;
                        if (c_in || (dr < erows && dc < ecols)) {
                            TO* p = ct + (C_COL ? dr + dc * cs32 : dr * rs32 + dc * cs32);
                            double old = 0.0;
                            if (!first) old = (double)*p;
                            else if (beta != 0.0) old = beta * (double)*p;
                            *p = (TO)fma(lam, (double)acc[i][j][q], old);
                        }
                    }
                }
            }
        }
    }
}

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

// Turns the runtime layout flags into template arguments, one at a time.
template <typename TI, typename TL, typename TO, bool... F>
struct Dispatch {
    static void go(const KArgs& ka, const bool* flags, int nwg, hipStream_t st)
    {
        if constexpr (sizeof...(F) == 5) {
            hipLaunchKernelGGL((fmm_gemm_kernel<TI, TL, TO, F...>), dim3(nwg), dim3(NT), 0, st, ka);
        } else {
            if (flags[sizeof...(F)]) Dispatch<TI, TL, TO, F..., true>::go(ka, flags, nwg, st);
            else Dispatch<TI, TL, TO, F..., false>::go(ka, flags, nwg, st);
        }
    }
};

template <typename TI, typename TL, typename TO>
int launch_typed(const KArgs& ka, bool a_mn, bool b_mn, bool va, bool vb, bool c_col, int nwg, hipStream_t st)
{
    const bool flags[5] = {a_mn, b_mn, va, vb, c_col};
    Dispatch<TI, TL, TO>::go(ka, flags, nwg, st);
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
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
    long nwg = (long)ka.tiles_m * ka.tiles_n;
    if (nwg <= 0) return FMMHIP_SUCCESS;
    if (nwg > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;

    switch (kind) {
    case Kind::D64: return launch_typed<double, double, double>(ka, a_mn, b_mn, va, vb, rsC == 1, (int)nwg, st);
    case Kind::D_F32LEAF: return launch_typed<double, float, double>(ka, a_mn, b_mn, va, vb, rsC == 1, (int)nwg, st);
    case Kind::S32: return launch_typed<float, float, float>(ka, a_mn, b_mn, va, vb, rsC == 1, (int)nwg, st);
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
