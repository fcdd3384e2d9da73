// fmmhip_kernel.h -- the fused FMM kernel for gfx950 (device templates),
// included by the per-datatype translation units fmmhip_gemm_*.hip.
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
//                (the products-materialised mode passes one W entry per
//                product instead, the product's slot: a plain store, W
//                applied afterwards by fmmhip_wsum.hip)
//   R loop       frame/base/bli_strassen_ab.c:248-294 -- inside the kernel
//
// Work decomposition: one workgroup per TILE x TILE position of the leaf,
// looping over all R products.  It is the only writer of that tile position
// in every C part, so the read-modify-write of C is race-free without
// atomics and deterministic.  The first product that writes a part applies
// beta (beta == 0 never reads C).
#ifndef FMMHIP_KERNEL_H
#define FMMHIP_KERNEL_H

#include <hip/hip_runtime.h>

#include "fmm_internal.h"

namespace fmmk {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d2_t __attribute__((ext_vector_type(2)));
typedef float f2_t __attribute__((ext_vector_type(2)));

constexpr int BM = 128;          // tile rows and columns of the default shapes (0, 1)
constexpr int BN = 128;
constexpr int BK = 16;           // k per LDS stage
// LDS tiles are kept [BK][row] (mn innermost): rows of 16 lanes read 128
// contiguous bytes.  The transposing store of a k-contiguous operand writes
// 8 even k rows of one mn column per 8 lanes; the row stride spreads them
// over the banks: T + 1 doubles (8 distinct bank pairs per 8 lanes) and
// T + 2 floats (32 distinct banks per 32 lanes), T the tile's rows.  A
// stride of T + 16 put all 8 rows in one bank (8-way conflicts, 6% of the
// FP64 kernel's time).
template <typename TL, int T>
__device__ __host__ constexpr int lds_row() { return sizeof(TL) == 8 ? T + 1 : T + 2; }
template <typename TL, int T>
__device__ __host__ constexpr int lds_tile() { return BK * lds_row<TL, T>(); }   // elements per tile buffer

// Element (mn, k) of an LDS tile.
template <typename TL, int T>
__device__ __forceinline__ constexpr int lds_at(int mn, int k) { return k * lds_row<TL, T>() + mn; }

// Launch shape: NT threads as a WAVES_M x WAVES_N grid of waves over a
// TILE x TILE tile of C, OCC waves per SIMD.
// MULTI: the parts of a product after the first are interleaved with the
// MFMA sub-blocks (one extra register set in flight); otherwise they are
// loaded after the MFMA block.
//   Shape 0: 128 x 128 tile, 512 threads, 64 x 32 per wave, two workgroups per CU, single
//   Shape 1: 128 x 128 tile, 256 threads, 64 x 64 per wave, two workgroups per CU, multi
//   Shape 2:  64 x  64 tile, 256 threads, 32 x 32 per wave, four workgroups per CU, single
//            (0.92x shape 0's rate on full tiles, measured; no padding waste
//            on dimensions that are odd multiples of 64)
template <int SHAPE> struct Cfg;
template <> struct Cfg<0> { static constexpr int NT = 512, WAVES_M = 2, WAVES_N = 4, OCC = 4, TILE = 128; static constexpr bool MULTI = false; };
template <> struct Cfg<1> { static constexpr int NT = 256, WAVES_M = 2, WAVES_N = 2, OCC = 2, TILE = 128; static constexpr bool MULTI = true; };
template <> struct Cfg<2> { static constexpr int NT = 256, WAVES_M = 2, WAVES_N = 2, OCC = 4, TILE = 64; static constexpr bool MULTI = false; };

// Tile rows (= columns) of a launch shape, on the host.
inline int tile_size(int shape) { return shape == 2 ? Cfg<2>::TILE : Cfg<0>::TILE; }

template <int SHAPE> struct Geo {
    static constexpr int NT = Cfg<SHAPE>::NT;
    static constexpr int T = Cfg<SHAPE>::TILE;
    static constexpr int WTM = T / Cfg<SHAPE>::WAVES_M;
    static constexpr int WTN = T / Cfg<SHAPE>::WAVES_N;
    static constexpr int FI = WTM / 16;            // 16x16 MFMA blocks per wave along M
    static constexpr int FJ = WTN / 16;            // ... along N
    static constexpr int PAIRS = T * BK / 2 / NT;  // staged pairs per thread per operand
    static constexpr int MN_PAIRS = T / 2;         // pairs along mn in one k row
    static constexpr int K_ROWS = NT / MN_PAIRS;   // k rows per pass (mn-contiguous)
    static constexpr int K_PAIRS = BK / 2;         // pairs along k in one mn column
    static constexpr int MN_COLS = NT / K_PAIRS;   // mn columns per pass (k-contiguous)
    static_assert(PAIRS * NT * 2 == T * BK, "staging covers the tile");
    static_assert(PAIRS * K_ROWS == BK && PAIRS * MN_COLS == T, "staging passes cover the tile");
};

struct KArgs {
    const void* A;
    const void* B;
    void* C;
    long lda;                    // the non-unit stride of A (A_MN: A(i,p) at A[i + p*lda])
    long ldb;                    // the non-unit stride of B (B_MN: B(p,j) at B[p*ldb + j])
    long rsC, csC;
    int lm, lk, ln;              // leaf dims
    int min_m, min_k, min_n;     // smallest A-rows / k / B-cols part extents: tiles inside need no checks
    int min_cm, min_cn;          // smallest C part extents (epilogue)
    int R;
    int tiles_m, tiles_n;
    int va, vb;                  // pairs of A / B are 16-byte aligned (vector loads allowed)
    // Product split: `groups` (1, 2, 4, 8) workgroups per tile, group g
    // taking products [g*R/groups, (g+1)*R/groups) and scattering into its
    // own temporary tmp + g*tmp_elems (C's strides; zeroed beforehand, reduced
    // into C afterwards), first writers from bit gbit of the entries' flags.
    // ksplit: the groups split the k-steps instead (R == 1, classical).
    // xcd_groups: the launch deals whole groups to the XCDs (group_tile_of)
    int groups, gbit, ksplit, xcd_groups;
    void* tmp;
    long tmp_elems;
    const int* ptr;              // u, v, w CSR pointers, 3*(R+1)
    const fmm_entry_t* ent;
    double alpha, beta;
};

// Accumulator layout of the 16x16 output block, lane l, register q:
//   FP64 (v_mfma_f64_16x16x4_f64):  row l/16 + 4q,    col l%16
//   FP32 (v_mfma_f32_16x16x4_f32):  row 4(l/16) + q,  col l%16
// (both checked on the device with known operands).  Operands: lane l holds
// A[l%16][l/16] and B[l/16][l%16] of the 16x4 / 4x16 blocks.
template <typename TL> struct Leaf;
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

// Global-memory views of pointers.  The batched kernels read a problem's
// arguments from memory, where the compiler cannot tell the address space
// of the pointers in them and emits flat loads and stores; flat ones also
// count against lgkmcnt, which ties the waits for the LDS fragment reads to
// the global loads of the next tile.  Loads and stores through these types
// are global ones (the pointers do point to global memory).
template <typename T> struct GPtr {
    typedef const __attribute__((address_space(1))) T c;
    typedef __attribute__((address_space(1))) T m;
};

// An entry read field by field through a global-memory pointer (a struct
// copy from another address space has no constructor in C++).
typedef typename GPtr<fmm_entry_t>::c gent_t;
__device__ __forceinline__ fmm_entry_t load_entry(gent_t* p)
{
    fmm_entry_t e;
    e.off = p->off;
    e.rows = p->rows;
    e.cols = p->cols;
    e.coef = p->coef;
    e.first = p->first;
    e.part = p->part;
    return e;
}

template <typename TI> struct Pair;
template <> struct Pair<double> { typedef d2_t t; };
template <> struct Pair<float> { typedef f2_t t; };

// Staging of one operand tile (TILE along "mn", BK along k) for one
// k-step: the weighted sum over the product's nonzero parts.  The raw values
// of up to NE parts are loaded before the MFMA block of the current k-step
// (their loads are in flight while the matrix cores run) and combined after
// it; parts beyond NE are loaded and accumulated after the block.
//   MN_CONTIG: the operand is contiguous along mn (pairs along mn),
//              otherwise along k (pairs along k).
//   IS_B:      the entry's (rows, cols) are (k, mn) rather than (mn, k).
// Addresses are a uniform tile pointer (scalar registers) plus a per-thread
// 32-bit offset fixed for the whole kernel.
template <typename TI, bool MN_CONTIG, bool IS_B, int SHAPE>
struct Stage {
    static constexpr int PAIRS = Geo<SHAPE>::PAIRS, MN_PAIRS = Geo<SHAPE>::MN_PAIRS, K_ROWS = Geo<SHAPE>::K_ROWS,
                         K_PAIRS = Geo<SHAPE>::K_PAIRS, MN_COLS = Geo<SHAPE>::MN_COLS;
    double sum[PAIRS][2];        // the weighted sum being formed
    double tmp[PAIRS][2];        // raw values of the part in flight
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
    template <bool VEC>
    __device__ __forceinline__ void load(const TI* __restrict__ base, const fmm_entry_t& e, long ld,
                                         int mn0, int k0, bool interior, double (&x)[PAIRS][2]) const
    {
        typedef typename Pair<TI>::t pair_t;
        typedef typename GPtr<TI>::c gTI;
        typedef typename GPtr<pair_t>::c gpair_t;
        gTI* t = (gTI*)base + e.off + (MN_CONTIG ? (long)mn0 + (long)k0 * ld : (long)mn0 * ld + k0) + off0;
        const int ld32 = (int)ld;
        if (interior) {
#pragma unroll
            for (int s = 0; s < PAIRS; s++) {
                gTI* p = t + (MN_CONTIG ? dk(s) * ld32 : dmn(s) * ld32);
                if (VEC) {
                    pair_t q = *(gpair_t*)p;
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
            gTI* p = t + (MN_CONTIG ? dk(s) * ld32 : dmn(s) * ld32);
            const int mn_0 = lmn0 + dmn(s), k_0 = lk0 + dk(s);
            // element 1 is one step along the contiguous dimension
            const bool ok0 = mn_0 < emn && k_0 < ek;
            const bool ok1 = MN_CONTIG ? (mn_0 + 1 < emn && k_0 < ek) : (mn_0 < emn && k_0 + 1 < ek);
            double x0 = 0.0, x1 = 0.0;
            if (VEC && ok1) {
                pair_t q = *(gpair_t*)p;
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

    // sum := raw part e (scaled later by scale_first)
    template <bool VEC>
    __device__ __forceinline__ void load_first(const TI* base, const fmm_entry_t& e, long ld, int mn0, int k0,
                                               bool interior)
    {
        load<VEC>(base, e, ld, mn0, k0, interior, sum);
    }

    __device__ __forceinline__ void scale_first(double c)
    {
        if (c != 1.0) {
#pragma unroll
            for (int s = 0; s < PAIRS; s++) { sum[s][0] *= c; sum[s][1] *= c; }
        }
    }

    // tmp := raw part e (in flight until add_tmp)
    template <bool VEC>
    __device__ __forceinline__ void load_tmp(const TI* base, const fmm_entry_t& e, long ld, int mn0, int k0,
                                             bool interior)
    {
        load<VEC>(base, e, ld, mn0, k0, interior, tmp);
    }

    __device__ __forceinline__ void add_tmp(double c)
    {
#pragma unroll
        for (int s = 0; s < PAIRS; s++) {
            sum[s][0] = fma(c, tmp[s][0], sum[s][0]);
            sum[s][1] = fma(c, tmp[s][1], sum[s][1]);
        }
    }

    // Writes the staged sums into an LDS tile (layout lds_at).
    template <typename TL>
    __device__ __forceinline__ void store(TL* lds) const
    {
        constexpr int T = Geo<SHAPE>::T;
        TL* q = lds + lds_at<TL, T>(lmn0, lk0);
#pragma unroll
        for (int s = 0; s < PAIRS; s++) {
            TL* qs = q + lds_at<TL, T>(dmn(s), dk(s));
            qs[0] = (TL)sum[s][0];
            qs[lds_at<TL, T>(MN_CONTIG ? 1 : 0, MN_CONTIG ? 0 : 1)] = (TL)sum[s][1];   // next along the contiguous dim
        }
    }
};

// XCD-aware tile order: consecutive workgroup ids are dealt round-robin to
// the 8 XCDs, so ids are first regrouped so that each XCD receives a
// contiguous run of tiles, then the run is walked in groups of 8 tile rows
// (column-major inside a group) so that concurrently running workgroups on
// one XCD share A row panels and B column panels in that XCD's L2.
__device__ __forceinline__ void tile_walk(int lin, int tiles_m, int tiles_n, int& tm, int& tn)
{
    const int GROUP = 8;
    const int per_group = GROUP * tiles_n;
    const int g = lin / per_group;
    const int first_m = g * GROUP;
    const int gsz = min(tiles_m - first_m, GROUP);
    const int in = lin - g * per_group;
    tm = first_m + in % gsz;
    tn = in / gsz;
}

__device__ __forceinline__ void tile_of(int bid, int tiles_m, int tiles_n, int& tm, int& tn)
{
    const int nwg = tiles_m * tiles_n;
    int lin = bid;
    if ((nwg & 7) == 0) lin = (bid & 7) * (nwg >> 3) + (bid >> 3);
    tile_walk(lin, tiles_m, tiles_n, tm, tn);
}

// The order over a grid of `groups` product (or k-step) groups of tiles.
// By default block g * tiles + t is tile t (tile_of's order) of group g, so
// the XCDs share each product's tiles.  With xcd, the ids are regrouped over
// the whole grid so that each XCD receives a contiguous run of (group,
// tile) pairs, group-major: whole groups per XCD, whose workgroups share
// their products' leaves in that XCD's L2.  The host asks for it where a
// product has no more tiles than an XCD holds workgroups at once (444 L2 at
// 13824^3, 49 tiles per product: spread over the XCDs, every product was
// read into every XCD's L2, L2 hit 0.17); a product of many tiles walked by
// one XCD re-reads its panels more often than eight XCDs sharing it.
__device__ __forceinline__ void group_tile_of(int bid, int groups, bool xcd, int tiles_m, int tiles_n, int& grp,
                                              int& tm, int& tn)
{
    const int tiles = tiles_m * tiles_n;
    if (groups == 1 || !xcd) {
        grp = bid / tiles;
        tile_of(bid - grp * tiles, tiles_m, tiles_n, tm, tn);
        return;
    }
    const int nwg = groups * tiles;
    const int x = bid & 7, q = nwg >> 3, r = nwg & 7;
    const int lin = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + (bid >> 3);
    grp = lin / tiles;
    tile_walk(lin - grp * tiles, tiles_m, tiles_n, tm, tn);
}

// One 16-column block (column block j of the wave's accumulators) of the W
// epilogue for one C part: C := (first ? beta*C : C) + lam*acc.  Every read
// of C is issued before any write, so the reads are in flight together, and
// each write goes out from its own register.  CHECK clips to the part
// (erows/ecols: the part's extent from the block's first row/column).
template <bool CHECK, int FI, int ROW_REG, bool C_COL, typename TO, typename ACC, int FJ>
__device__ __forceinline__ void epi_block(TO* __restrict__ cb, int rs32, int erows, int ecols,
                                          const ACC (&acc)[FI][FJ], int j, double lam, bool first, double beta,
                                          bool rd)
{
    typename GPtr<TO>::m* g = (typename GPtr<TO>::m*)cb;
    double v[FI][4];
#pragma unroll
    for (int i = 0; i < FI; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int dr = 16 * i + ROW_REG * q;
            typename GPtr<TO>::c* p = g + (C_COL ? dr : dr * rs32);
            v[i][q] = (rd && (!CHECK || (dr < erows && 0 < ecols))) ? (double)*p : 0.0;
        }
#pragma unroll
    for (int i = 0; i < FI; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) v[i][q] = fma(lam, (double)acc[i][j][q], first ? beta * v[i][q] : v[i][q]);
#pragma unroll
    for (int i = 0; i < FI; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int dr = 16 * i + ROW_REG * q;
            if (!CHECK || (dr < erows && 0 < ecols)) g[C_COL ? dr : dr * rs32] = (TO)v[i][q];
        }
}

// One TILE x TILE position (tm, tn) of the leaf, all R products: the
// body of both the one-problem kernel and the batched persistent kernel.
// sA/sB are the workgroup's double-buffered LDS tiles.
// KSPL: the groups split the k-steps (a.ksplit; instantiated apart, so that
// the other kernels keep their registers).
template <typename TI, typename TL, typename TO, int SHAPE, bool VEC, bool A_MN, bool B_MN, bool C_COL,
          bool KSPL = false>
__device__ __forceinline__ void gemm_tile(const KArgs& a, int tm, int tn, int grp,
                                          TL (*sA)[lds_tile<TL, Cfg<SHAPE>::TILE>()],
                                          TL (*sB)[lds_tile<TL, Cfg<SHAPE>::TILE>()])
{
    constexpr int BM = Cfg<SHAPE>::TILE, BN = Cfg<SHAPE>::TILE;   // this shape's tile
    constexpr int WAVES_M = Cfg<SHAPE>::WAVES_M;
    constexpr bool MULTI = Cfg<SHAPE>::MULTI;
    constexpr int WTM = Geo<SHAPE>::WTM, WTN = Geo<SHAPE>::WTN, FI = Geo<SHAPE>::FI, FJ = Geo<SHAPE>::FJ;
    typedef typename Leaf<TL>::acc_t acc_t;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = (wave % WAVES_M) * WTM;
    const int wn = (wave / WAVES_M) * WTN;
    const int m0 = tm * BM, n0 = tn * BN;

    const TI* __restrict__ Ab = (const TI*)a.A;
    const TI* __restrict__ Bb = (const TI*)a.B;
    const int R = a.R;
    const bool split = a.groups > 1;
    TO* __restrict__ Cb = split ? (TO*)a.tmp + (long)grp * a.tmp_elems : (TO*)a.C;
    constexpr bool ks = KSPL;
    const int r_begin = ks ? 0 : (int)((long)grp * R / a.groups);
    const int r_end = ks ? R : (int)((long)(grp + 1) * R / a.groups);
    // lists as global-memory pointers (the batched form reads the arguments
    // from memory; see GPtr)
    typedef typename GPtr<int>::c gint;
    gint* __restrict__ uptr = (gint*)a.ptr;
    gint* __restrict__ vptr = uptr + (R + 1);
    gint* __restrict__ wptr = uptr + 2 * (R + 1);
    gent_t* __restrict__ ent = (gent_t*)a.ent;
    const int nk = (a.lk + BK - 1) / BK;
    const int kt_begin = ks ? (int)((long)grp * nk / a.groups) : 0;
    const int kt_end = ks ? (int)((long)(grp + 1) * nk / a.groups) : nk;
    const long lda = a.lda, ldb = a.ldb;

    // tiles inside every part need no bounds checks
    const bool a_rows_in = m0 + BM <= a.min_m;
    const bool b_cols_in = n0 + BN <= a.min_n;
    const bool c_in = m0 + BM <= a.min_cm && n0 + BN <= a.min_cn;

    Stage<TI, A_MN, false, SHAPE> stA;
    Stage<TI, B_MN, true, SHAPE> stB;
    stA.init(tid, lda);
    stB.init(tid, ldb);


    for (int r = r_begin; r < r_end; r++) {
        const int u0 = uptr[r], nu = uptr[r + 1] - u0;
        const int v0 = vptr[r], nv = vptr[r + 1] - v0;
        const int ng = nu > nv ? nu : nv;   // load groups: part g of A and part g of B
        // the product's first parts, held for all k-steps
        const fmm_entry_t ea0 = load_entry(ent + u0), eb0 = load_entry(ent + v0);

        acc_t acc[FI][FJ];
#pragma unroll
        for (int i = 0; i < FI; i++)
#pragma unroll
            for (int j = 0; j < FJ; j++) acc[i][j] = (acc_t){0, 0, 0, 0};

        // prologue: stage the first k-step (all groups, one after another)
        {
            const int kp = kt_begin * BK;
            const bool kpin = kp + BK <= a.min_k;
            const bool ia = a_rows_in && kpin, ib = b_cols_in && kpin;
            if (ia && ib) {
                stA.template load_first<VEC>(Ab, ea0, lda, m0, kp, true);
                stB.template load_first<VEC>(Bb, eb0, ldb, n0, kp, true);
            } else {
                stA.template load_first<VEC>(Ab, ea0, lda, m0, kp, ia);
                stB.template load_first<VEC>(Bb, eb0, ldb, n0, kp, ib);
            }
            stA.scale_first(ea0.coef);
            stB.scale_first(eb0.coef);
            for (int g = 1; g < nu; g++) {
                stA.template load_tmp<VEC>(Ab, load_entry(ent + u0 + g), lda, m0, kp, ia);
                stA.add_tmp(ent[u0 + g].coef);
            }
            for (int g = 1; g < nv; g++) {
                stB.template load_tmp<VEC>(Bb, load_entry(ent + v0 + g), ldb, n0, kp, ib);
                stB.add_tmp(ent[v0 + g].coef);
            }
            stA.template store<TL>(sA[0]);
            stB.template store<TL>(sB[0]);
        }
        __syncthreads();

        // Main loop.  While the MFMAs of k-step kt run out of LDS buffer
        // kt&1, the next k-step's parts stream in: group 0 and 1 are issued
        // before MFMA sub-block 0; after sub-block g the group g+1 that
        // arrived is folded into the sum and group g+2 is issued, so each
        // part's load is covered by a quarter of the MFMA block.
        constexpr int KS = BK / 4;
        for (int kt = kt_begin; kt < kt_end; kt++) {
            const int cur = (kt - kt_begin) & 1;
            const bool more = kt + 1 < kt_end;
            const int k0 = (kt + 1) * BK;
            const bool kin = k0 + BK <= a.min_k;
            const bool ia = a_rows_in && kin, ib = b_cols_in && kin;
            if (more) {
                // both operands' first parts in one block when the tiles are
                // interior (no per-operand branches between the loads)
                if (ia && ib) {
                    stA.template load_first<VEC>(Ab, ea0, lda, m0, k0, true);
                    stB.template load_first<VEC>(Bb, eb0, ldb, n0, k0, true);
                } else {
                    stA.template load_first<VEC>(Ab, ea0, lda, m0, k0, ia);
                    stB.template load_first<VEC>(Bb, eb0, ldb, n0, k0, ib);
                }
                if (MULTI) {
                    if (1 < nu) stA.template load_tmp<VEC>(Ab, load_entry(ent + u0 + 1), lda, m0, k0, ia);
                    if (1 < nv) stB.template load_tmp<VEC>(Bb, load_entry(ent + v0 + 1), ldb, n0, k0, ib);
                }
            }
            const TL* la = sA[cur];
            const TL* lb = sB[cur];
            // LLVM's single-wave small-GEMM interleaving of the LDS reads
            // with the MFMAs, for the multi-part shape; the single shapes
            // run 1.9% faster (FP64, m = n = k = 16384) without it and with
            // the next stage stored at the end of the MFMA loop
            if (MULTI) __builtin_amdgcn_iglp_opt(1);
            // single-part operands (classical, pre-summed): nothing to add
            const bool single = !MULTI && nu == 1 && nv == 1;
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                const int kk = 4 * ks + (lane >> 4);
                TL fa[FI], fb[FJ];
#pragma unroll
                for (int i = 0; i < FI; i++) fa[i] = la[lds_at<TL, BM>(wm + 16 * i + (lane & 15), kk)];
#pragma unroll
                for (int j = 0; j < FJ; j++) fb[j] = lb[lds_at<TL, BN>(wn + 16 * j + (lane & 15), kk)];
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) acc[i][j] = Leaf<TL>::mfma(fa[i], fb[j], acc[i][j]);
                if (single && more && ks == KS - 1) {
                    stA.scale_first(ea0.coef);
                    stB.scale_first(eb0.coef);
                    stA.template store<TL>(sA[cur ^ 1]);
                    stB.template store<TL>(sB[cur ^ 1]);
                }
                if (MULTI && more) {
                    if (ks == 0) {
                        stA.scale_first(ea0.coef);
                        stB.scale_first(eb0.coef);
                    }
                    const int g = ks + 1;            // group that was in flight
                    if (g < ng) {
                        if (g < nu) stA.add_tmp(ent[u0 + g].coef);
                        if (g < nv) stB.add_tmp(ent[v0 + g].coef);
                        if (g + 1 < nu) stA.template load_tmp<VEC>(Ab, load_entry(ent + u0 + g + 1), lda, m0, k0, ia);
                        if (g + 1 < nv) stB.template load_tmp<VEC>(Bb, load_entry(ent + v0 + g + 1), ldb, n0, k0, ib);
                    }
                }
            }
            if (more && !single) {
                if (MULTI) {
                    // groups past the MFMA block: the one in flight, then serially
                    for (int g = KS + 1; g < ng; g++) {
                        if (g < nu) stA.add_tmp(ent[u0 + g].coef);
                        if (g < nv) stB.add_tmp(ent[v0 + g].coef);
                        if (g + 1 < nu) stA.template load_tmp<VEC>(Ab, load_entry(ent + u0 + g + 1), lda, m0, k0, ia);
                        if (g + 1 < nv) stB.template load_tmp<VEC>(Bb, load_entry(ent + v0 + g + 1), ldb, n0, k0, ib);
                    }
                } else {
                    // single: every further part after the MFMA block, one
                    // at a time (fewest registers)
                    stA.scale_first(ea0.coef);
                    stB.scale_first(eb0.coef);
                    for (int g = 1; g < nu; g++) {
                        stA.template load_tmp<VEC>(Ab, load_entry(ent + u0 + g), lda, m0, k0, ia);
                        stA.add_tmp(ent[u0 + g].coef);
                    }
                    for (int g = 1; g < nv; g++) {
                        stB.template load_tmp<VEC>(Bb, load_entry(ent + v0 + g), ldb, n0, k0, ib);
                        stB.add_tmp(ent[v0 + g].coef);
                    }
                }
                stA.template store<TL>(sA[cur ^ 1]);
                stB.template store<TL>(sB[cur ^ 1]);
            }
            __syncthreads();
        }

        // W epilogue: C_part += alpha*w*acc for every nonzero w of product r;
        // the part's first writer applies beta (beta == 0: C is not read).
        const int w0 = wptr[r], w1 = wptr[r + 1];
        // epilogue coordinates of this lane's accumulator element (i, j, q):
        // row = m0 + wm + 16 i + ROW_LANE (lane/16) + ROW_REG q, col = n0 + wn + 16 j + lane%16
        const int er0 = wm + Leaf<TL>::ROW_LANE * (lane >> 4);
        const int ec0 = wn + (lane & 15);
        for (int e = w0; e < w1; e++) {
            const fmm_entry_t we = load_entry(ent + e);
            const double lam = a.alpha * we.coef;
            const bool first = ((we.first >> a.gbit) & 1) != 0;
            const double beta = split ? 0.0 : a.beta;      // temporaries start at zero
            const long rsC = C_COL ? 1 : a.rsC;
            const int rs32 = (int)rsC, cs32 = (int)a.csC;
            // lane's element (i, j, q) is at tile row er0 + 16 i + ROW_REG q, col ec0 + 16 j
            TO* ct = Cb + we.off + (long)(m0 + er0) * rsC + (long)(n0 + ec0) * a.csC;
            const int erows = we.rows - m0 - er0, ecols = we.cols - n0 - ec0;
            const bool rd = !first || beta != 0.0;          // uniform: C is read
#pragma unroll
            for (int j = 0; j < FJ; j++) {
                TO* cb = ct + 16 * j * cs32;
                if (c_in)
                    epi_block<false, FI, Leaf<TL>::ROW_REG, C_COL>(cb, rs32, 0, 0, acc, j, lam, first, beta, rd);
                else
                    epi_block<true, FI, Leaf<TL>::ROW_REG, C_COL>(cb, rs32, erows, ecols - 16 * j, acc, j, lam, first,
                                                                 beta, rd);
            }
        }
    }
}

template <typename TI, typename TL, typename TO, int SHAPE, bool VEC, bool A_MN, bool B_MN, bool C_COL,
          bool KSPL = false>
__global__ __attribute__((amdgpu_flat_work_group_size(Cfg<SHAPE>::NT, Cfg<SHAPE>::NT),
                          amdgpu_waves_per_eu(Cfg<SHAPE>::OCC, Cfg<SHAPE>::OCC)))
void fmm_gemm_kernel(KArgs a)
{
    __shared__ TL sA[2][lds_tile<TL, Cfg<SHAPE>::TILE>()];
    __shared__ TL sB[2][lds_tile<TL, Cfg<SHAPE>::TILE>()];
    int grp, tm, tn;
    group_tile_of((int)blockIdx.x, a.groups, a.xcd_groups != 0, a.tiles_m, a.tiles_n, grp, tm, tn);
    gemm_tile<TI, TL, TO, SHAPE, VEC, A_MN, B_MN, C_COL, KSPL>(a, tm, tn, grp, sA, sB);
}

// Fast kernel for single-part operands (classical, pre-summed, products
// materialised) whose tiles all lie inside their parts: column-major A,
// k-contiguous B, column-major C, leaf extents multiples of the 128 x 128
// tile and of the k-step, 16-byte aligned pairs, no k-split.  256 threads as
// 2 x 2 waves of 64 x 64, two workgroups per CU; k-steps of 16 (FP64 leaf)
// or 32 (FP32 leaf: its MFMAs take half the time, so twice the k per
// barrier).  Per k-step the next tile is loaded through pointers walked per
// step, the fragments of sub-step ks+1 are read into the other register set
// while the MFMAs of sub-step ks run (the reads spread evenly among the
// MFMAs), and the next tile is stored at the end.  The MFMAs of each
// accumulator run in the same k order as in gemm_tile, operands are scaled
// and rounded to the leaf type the same way, and the W epilogue is its
// unclipped form, so results equal the other shapes' bit for bit.
template <bool B> struct BoolC { static constexpr bool value = B; };
constexpr int kFastNT = 256;
struct FastCfg { static constexpr int T = 128, WAVES = 2, WT = 64, FI = 4, FJ = 4; };
template <typename TL>
__device__ __host__ constexpr int fast_bk() { return sizeof(TL) == 8 ? 16 : 32; }
template <typename TL>
__device__ __host__ constexpr int fast_lds() { return fast_bk<TL>() * lds_row<TL, FastCfg::T>(); }

// One tile position (tm, tn) of product group grp: the body of the fast
// kernel and of its batched persistent form.
template <typename TI, typename TL, typename TO>
__device__ __forceinline__ void fast_tile(const KArgs& a, int tm, int tn, int grp, TL (*sA)[fast_lds<TL>()],
                                          TL (*sB)[fast_lds<TL>()])
{
    constexpr int T = FastCfg::T, FI = FastCfg::FI, FJ = FastCfg::FJ, FBK = fast_bk<TL>(), KS = FBK / 4;
    constexpr int RS = lds_row<TL, T>();
    // staging geometry: pairs along m for A (k rows), along k for B (n columns)
    constexpr int AP = T * FBK / 2 / kFastNT, A_MNP = T / 2, A_KR = kFastNT / A_MNP;
    constexpr int B_KP = FBK / 2, B_NC = kFastNT / B_KP;
    static_assert(AP * A_KR == FBK && AP * B_NC == T, "staging covers the tile");
    static_assert(KS % 2 == 0, "the fragment register sets alternate per sub-step");
    typedef typename Leaf<TL>::acc_t acc_t;
    typedef typename Pair<TI>::t pair_t;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave % FastCfg::WAVES) * FastCfg::WT, wn = (wave / FastCfg::WAVES) * FastCfg::WT;
    const int m0 = tm * T, n0 = tn * T;
    const int a_m = 2 * (tid % A_MNP), a_k = tid / A_MNP;
    const int b_k = 2 * (tid % B_KP), b_n = tid / B_KP;
    const long lda = a.lda, ldb = a.ldb;
    const bool split = a.groups > 1;
    // operands and lists as global-memory pointers (the batched form reads
    // the arguments from memory; see GPtr)
    typedef typename GPtr<TI>::c gTI;
    typedef typename GPtr<pair_t>::c gpair_t;
    TO* __restrict__ Cb = split ? (TO*)a.tmp + (long)grp * a.tmp_elems : (TO*)a.C;
    const int R = a.R;
    const int r_begin = (int)((long)grp * R / a.groups), r_end = (int)((long)(grp + 1) * R / a.groups);
    typedef typename GPtr<int>::c gint;
    gint* __restrict__ uptr = (gint*)a.ptr;
    gint* __restrict__ vptr = uptr + (R + 1);
    gint* __restrict__ wptr = uptr + 2 * (R + 1);
    gent_t* __restrict__ ent = (gent_t*)a.ent;
    const int nk = a.lk / FBK;
    const int l16 = lane & 15, kq = lane >> 4;

    pair_t ra[AP], rb[AP];
    TL fa[2][FI], fb[2][FJ];
    auto frag = [&](int buf, int ks, int f) {
        const TL* la = sA[buf];
        const TL* lb = sB[buf];
        const int kk = 4 * ks + kq;
#pragma unroll
        for (int i = 0; i < FI; i++) fa[f][i] = la[kk * RS + wm + 16 * i + l16];
#pragma unroll
        for (int j = 0; j < FJ; j++) fb[f][j] = lb[kk * RS + wn + 16 * j + l16];
    };

    for (int r = r_begin; r < r_end; r++) {
        const fmm_entry_t ea = load_entry(ent + uptr[r]), eb = load_entry(ent + vptr[r]);
        const double ca = ea.coef, cb = eb.coef;
        gTI* pa = (gTI*)(const TI*)a.A + ea.off + m0 + a_m + (long)a_k * lda;
        gTI* pb = (gTI*)(const TI*)a.B + eb.off + (long)(n0 + b_n) * ldb + b_k;
        const long da = (long)FBK * lda;
        auto gload = [&]() {
#pragma unroll
            for (int s = 0; s < AP; s++) ra[s] = *(gpair_t*)(pa + (long)(A_KR * s) * lda);
#pragma unroll
            for (int s = 0; s < AP; s++) rb[s] = *(gpair_t*)(pb + (long)(B_NC * s) * ldb);
            pa += da;
            pb += FBK;
        };
        // scaled as gemm_tile's staging does: in FP64 by the coefficients
        // (only for products where one is not 1: multiplying the other by 1
        // is exact), then rounded to the leaf type
        auto lstore = [&](int buf, bool scale) {
            if (scale) {
#pragma unroll
                for (int s = 0; s < AP; s++) {
                    ra[s][0] = (TI)((double)ra[s][0] * ca);
                    ra[s][1] = (TI)((double)ra[s][1] * ca);
                    rb[s][0] = (TI)((double)rb[s][0] * cb);
                    rb[s][1] = (TI)((double)rb[s][1] * cb);
                }
            }
            TL* la = sA[buf];
            TL* lb = sB[buf];
#pragma unroll
            for (int s = 0; s < AP; s++) {
                TL* q = la + (a_k + A_KR * s) * RS + a_m;
                q[0] = (TL)ra[s][0];
                q[1] = (TL)ra[s][1];
            }
#pragma unroll
            for (int s = 0; s < AP; s++) {
                TL* q = lb + b_k * RS + b_n + B_NC * s;
                q[0] = (TL)rb[s][0];
                q[RS] = (TL)rb[s][1];
            }
        };
        acc_t acc[FI][FJ];
#pragma unroll
        for (int i = 0; i < FI; i++)
#pragma unroll
            for (int j = 0; j < FJ; j++) acc[i][j] = (acc_t){0, 0, 0, 0};
        constexpr int NR = (FI + FJ) / 2, NM = FI * FJ;   // per sub-step: NR groups of NM / NR MFMAs, one LDS read each
        auto mfmas = [&](int f) {
#pragma unroll
            for (int i = 0; i < FI; i++)
#pragma unroll
                for (int j = 0; j < FJ; j++) acc[i][j] = Leaf<TL>::mfma(fa[f][i], fb[f][j], acc[i][j]);
        };
        auto interleave_reads = [&]() {
#pragma unroll
            for (int q = 0; q < NR; q++) {
                __builtin_amdgcn_sched_group_barrier(0x008, NM / NR, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);         // LDS read
            }
        };
        // FP32 leaf (eight sub-steps of 32-cycle MFMAs per k-step): the
        // staged next tile goes into the other buffer among the MFMAs of
        // sub-step KS-2 and the barrier follows that sub-step, so the first
        // fragments of the next step are read under the MFMAs of the last
        // sub-step; the tile after the next is loaded right after the
        // barrier (the staging registers are free then) and has the rest of
        // this step and most of the next to arrive.  Every LDS read of
        // buffer cur comes before the barrier, and every store into the
        // other buffer before it too, so one barrier per step orders both.
        // STORE: a next step exists; LOAD: a step after the next exists.
        // (FP64 leaf: the same order ran 2% slower than the one below.)
        auto kstep_mid = [&](int cur, auto sc, auto st, auto ld) {
            constexpr bool STORE = decltype(st)::value, LOAD = decltype(ld)::value;
#pragma unroll
            for (int ks = 0; ks + 1 < KS; ks++) {
                const int f = ks & 1;
                frag(cur, ks + 1, f ^ 1);
                mfmas(f);
                if (STORE && ks == KS - 2) {
                    lstore(cur ^ 1, decltype(sc)::value);
#pragma unroll
                    for (int q = 0; q < NR; q++) {
                        __builtin_amdgcn_sched_group_barrier(0x008, NM / NR, 0);       // MFMA
                        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);             // LDS read
                        __builtin_amdgcn_sched_group_barrier(0x200, 2 * AP / NR, 0);   // LDS write
                    }
                } else {
                    interleave_reads();
                }
            }
            __syncthreads();
            if (LOAD) gload();
            if (STORE) frag(cur ^ 1, 0, 0);
            mfmas((KS - 1) & 1);
            if (STORE) {
#pragma unroll
                for (int q = 0; q < NR; q++) {
                    __builtin_amdgcn_sched_group_barrier(0x008, NM / NR, 0);   // MFMA
                    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);         // LDS read
                    if (LOAD) __builtin_amdgcn_sched_group_barrier(0x020, 2 * AP / NR, 0);   // global load
                }
            }
        };
        // FP64 leaf (four sub-steps of 64-cycle MFMAs): the next tile is
        // loaded at the start of the step and stored after its MFMAs, then
        // the barrier.  (Measured slower: the order above, -2%; loads two
        // steps ahead in two staging sets, -4.5%; direct-to-LDS loads, -2%.)
        auto kstep_end = [&](int cur, auto sc, bool more) {
            if (more) gload();
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                const int f = ks & 1;
                if (ks + 1 < KS) frag(cur, ks + 1, f ^ 1);
                mfmas(f);
                if (ks + 1 < KS) interleave_reads();
                if (ks == KS - 1 && more) lstore(cur ^ 1, decltype(sc)::value);
            }
            __syncthreads();
            if (more) frag(cur ^ 1, 0, 0);
        };
        // the stores' scaling decided per product at compile time (one
        // instantiation each; a runtime branch in the FP64 order measured 5%
        // slower)
        auto kloop = [&](auto sc) {
            typedef BoolC<true> T_;
            typedef BoolC<false> F_;
            gload();
            lstore(0, decltype(sc)::value);
            if constexpr (sizeof(TL) == 4) {
                if (nk > 1) gload();
                __syncthreads();
                frag(0, 0, 0);
                int kt = 0;
                for (; kt + 2 < nk; kt++) kstep_mid(kt & 1, sc, T_(), T_());
                if (kt + 1 < nk) kstep_mid((kt++) & 1, sc, T_(), F_());
                kstep_mid(kt & 1, sc, F_(), F_());
            } else {
                __syncthreads();
                frag(0, 0, 0);
                for (int kt = 0; kt < nk; kt++) kstep_end(kt & 1, sc, kt + 1 < nk);
            }
        };
        if (ca != 1.0 || cb != 1.0)
            kloop(BoolC<true>());
        else
            kloop(BoolC<false>());
        // W epilogue, every tile inside every C part (unclipped form)
        const int er0 = wm + Leaf<TL>::ROW_LANE * kq;
        const int ec0 = wn + l16;
        for (int e = wptr[r]; e < wptr[r + 1]; e++) {
            const fmm_entry_t we = load_entry(ent + e);
            const double lam = a.alpha * we.coef;
            const bool first = ((we.first >> a.gbit) & 1) != 0;
            const double beta = split ? 0.0 : a.beta;
            const int cs32 = (int)a.csC;
            TO* ct = Cb + we.off + (long)(m0 + er0) + (long)(n0 + ec0) * a.csC;
            const bool rd = !first || beta != 0.0;
#pragma unroll
            for (int j = 0; j < FJ; j++)
                epi_block<false, FI, Leaf<TL>::ROW_REG, true>(ct + 16 * j * cs32, 1, 0, 0, acc, j, lam, first, beta,
                                                              rd);
        }
    }
}

template <typename TI, typename TL, typename TO>
__global__ __attribute__((amdgpu_flat_work_group_size(kFastNT, kFastNT), amdgpu_waves_per_eu(2, 2)))
void fmm_fast_kernel(KArgs a)
{
    __shared__ TL sA[2][fast_lds<TL>()];
    __shared__ TL sB[2][fast_lds<TL>()];
    int grp, tm, tn;
    group_tile_of((int)blockIdx.x, a.groups, a.xcd_groups != 0, a.tiles_m, a.tiles_n, grp, tm, tn);
    fast_tile<TI, TL, TO>(a, tm, tn, grp, sA, sB);
}

// Whether a prepared launch with leaf type TL can take fmm_fast_kernel
// (single-part operands are the caller's to guarantee).
template <typename TL>
inline bool fast_eligible(const KArgs& ka, bool a_mn, bool b_mn, bool c_col)
{
    constexpr int T = FastCfg::T;
    return a_mn && !b_mn && c_col && ka.va && ka.vb && !ka.ksplit && ka.lm % T == 0 && ka.ln % T == 0 &&
           ka.lk % fast_bk<TL>() == 0 && ka.lk >= fast_bk<TL>() && ka.min_m == ka.lm && ka.min_n == ka.ln &&
           ka.min_k == ka.lk && ka.min_cm >= ka.lm && ka.min_cn >= ka.ln && ka.tiles_m * T == ka.lm &&
           ka.tiles_n * T == ka.ln;
}

// Batched persistent kernel: a fixed grid of workgroups takes work items
// (problem, tile, product group) from device queues until they are empty.  There is one
// queue per XCD (workgroup b of a grid that is resident at once runs on XCD
// b mod 8): the host deals each problem's tiles to the queues in contiguous
// blocks of its tile grid, so the workgroups of one XCD share A and B panels
// in that XCD's L2, and lists each queue most expensive first.  A workgroup
// takes from its own queue, then from the next ones once it is empty.  The
// queue heads are counters, reset by batch_reset_kernel before each launch.
constexpr int kBatchQueues = 8;
struct BatchItem {
    int problem;
    int tile;                    // linear tile index in the problem's leaf grid, + product group x tiles
};

// The next work item of the workgroup (uniform), or -1 when every queue
// from `q` on is empty; q and tries carry the workgroup's position.
__device__ __forceinline__ int batch_next(const int* __restrict__ qoff, int* __restrict__ heads, int& q, int& tries,
                                          int* slot)
{
    for (;;) {
        if (threadIdx.x == 0) *slot = atomicAdd(&heads[q], 1);
        __syncthreads();
        // uniform across the workgroup: readfirstlane keeps the item and its
        // problem descriptor in scalar registers
        const int w = __builtin_amdgcn_readfirstlane(*slot);
        __syncthreads();                       // everyone has read slot before it is rewritten
        const int q0 = qoff[q], qn = qoff[q + 1] - q0;
        if (w < qn) return q0 + w;
        if (++tries == kBatchQueues) return -1;
        q = (q + 1) % kBatchQueues;
    }
}

template <typename TI, typename TL, typename TO, int SHAPE, bool VEC>
__global__ __attribute__((amdgpu_flat_work_group_size(Cfg<SHAPE>::NT, Cfg<SHAPE>::NT),
                          amdgpu_waves_per_eu(Cfg<SHAPE>::OCC, Cfg<SHAPE>::OCC)))
void fmm_batch_kernel(const KArgs* __restrict__ probs, const BatchItem* __restrict__ items,
                      const int* __restrict__ qoff, int* __restrict__ heads)
{
    __shared__ TL sA[2][lds_tile<TL, Cfg<SHAPE>::TILE>()];
    __shared__ TL sB[2][lds_tile<TL, Cfg<SHAPE>::TILE>()];
    __shared__ int slot;
    int q = (int)blockIdx.x % kBatchQueues, tries = 0;
    for (;;) {
        const int w = batch_next(qoff, heads, q, tries, &slot);
        if (w < 0) break;
        const BatchItem it = items[w];
        const KArgs a = probs[it.problem];
        const int nt = a.tiles_m * a.tiles_n, grp = it.tile / nt, t = it.tile - grp * nt;
        const int tm = t % a.tiles_m, tn = t / a.tiles_m;
        gemm_tile<TI, TL, TO, SHAPE, VEC, true, false, true>(a, tm, tn, grp, sA, sB);
        __syncthreads();                       // LDS free before the next item's prologue
    }
}

// The batched form of the fast kernel (problems the host found eligible,
// and the interior tiles of the others; an item is one tile of one product
// group).
template <typename TI, typename TL, typename TO>
__global__ __attribute__((amdgpu_flat_work_group_size(kFastNT, kFastNT), amdgpu_waves_per_eu(2, 2)))
void fmm_batch_fast_kernel(const KArgs* __restrict__ probs, const BatchItem* __restrict__ items,
                           const int* __restrict__ qoff, int* __restrict__ heads)
{
    __shared__ TL sA[2][fast_lds<TL>()];
    __shared__ TL sB[2][fast_lds<TL>()];
    __shared__ int slot;
    int q = (int)blockIdx.x % kBatchQueues, tries = 0;
    for (;;) {
        const int w = batch_next(qoff, heads, q, tries, &slot);
        if (w < 0) break;
        const BatchItem it = items[w];
        const KArgs a = probs[it.problem];
        const int nt = a.tiles_m * a.tiles_n, grp = it.tile / nt, t = it.tile - grp * nt;
        const int tm = t % a.tiles_m, tn = t / a.tiles_m;
        fast_tile<TI, TL, TO>(a, tm, tn, grp, sA, sB);
        __syncthreads();                       // LDS free before the next item's prologue
    }
}

template <int UNUSED>   // a template, so each translation unit may instantiate it
__global__ void batch_reset_kernel(int* heads)
{
    if (threadIdx.x < kBatchQueues) heads[threadIdx.x] = 0;
}

// Turns runtime layout flags into template arguments.
template <typename TI, typename TL, typename TO, int SHAPE, bool VEC>
int launch_layouts(const KArgs& ka, bool a_mn, bool b_mn, bool c_col, int nwg, hipStream_t st)
{
#define FMMK_GO(AM, BMN, CC)                                                                           \
    hipLaunchKernelGGL((fmm_gemm_kernel<TI, TL, TO, SHAPE, VEC, AM, BMN, CC>), dim3(nwg),                \
                       dim3(Cfg<SHAPE>::NT), 0, st, ka)
    const int sel = (a_mn ? 4 : 0) | (b_mn ? 2 : 0) | (c_col ? 1 : 0);
    if (ka.ksplit) {
        // the k-split is prepared for column-major A, B and C only
        if (sel != 5) return FMMHIP_ERR_UNSUPPORTED;
        hipLaunchKernelGGL((fmm_gemm_kernel<TI, TL, TO, SHAPE, VEC, true, false, true, true>), dim3(nwg),
                           dim3(Cfg<SHAPE>::NT), 0, st, ka);
        return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
    }
    switch (sel) {
    case 0: FMMK_GO(false, false, false); break;
    case 1: FMMK_GO(false, false, true); break;
    case 2: FMMK_GO(false, true, false); break;
    case 3: FMMK_GO(false, true, true); break;
    case 4: FMMK_GO(true, false, false); break;
    case 5: FMMK_GO(true, false, true); break;
    case 6: FMMK_GO(true, true, false); break;
    default: FMMK_GO(true, true, true); break;
    }
#undef FMMK_GO
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}

// Vector loads need both operands' pairs 16-byte aligned (ka.va && ka.vb).
// Shape 2 (64 x 64 tiles) is instantiated for column-major A, B and C
// without the k-split only.
template <typename TI, typename TL, typename TO>
int launch_typed(const KArgs& ka, int shape, bool a_mn, bool b_mn, bool c_col, int nwg, hipStream_t st)
{
    const bool vec = ka.va && ka.vb;
    if (shape == 2) {
        if (ka.ksplit || !(a_mn && !b_mn && c_col)) return FMMHIP_ERR_UNSUPPORTED;
        if (vec)
            hipLaunchKernelGGL((fmm_gemm_kernel<TI, TL, TO, 2, true, true, false, true>), dim3(nwg), dim3(Cfg<2>::NT),
                               0, st, ka);
        else
            hipLaunchKernelGGL((fmm_gemm_kernel<TI, TL, TO, 2, false, true, false, true>), dim3(nwg), dim3(Cfg<2>::NT),
                               0, st, ka);
        return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
    }
    if (shape == 1) {
        if (vec) return launch_layouts<TI, TL, TO, 1, true>(ka, a_mn, b_mn, c_col, nwg, st);
        return launch_layouts<TI, TL, TO, 1, false>(ka, a_mn, b_mn, c_col, nwg, st);
    }
    if (shape == 3) {
        // the fast kernel (the host selects it for eligible single-part calls)
        if (!fast_eligible<TL>(ka, a_mn, b_mn, c_col)) return FMMHIP_ERR_UNSUPPORTED;
        hipLaunchKernelGGL((fmm_fast_kernel<TI, TL, TO>), dim3(nwg), dim3(kFastNT), 0, st, ka);
        return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
    }
    if (vec) return launch_layouts<TI, TL, TO, 0, true>(ka, a_mn, b_mn, c_col, nwg, st);
    return launch_layouts<TI, TL, TO, 0, false>(ka, a_mn, b_mn, c_col, nwg, st);
}

// Batched launch (column-major A, B, C): reset the queue head, then the
// persistent kernel of launch shape `shape` (0, 3: 128 x 128 tiles, 2: 64 x 64)
// on `nwg` workgroups.
template <typename TI, typename TL, typename TO>
int launch_batch_typed(const KArgs* probs, const BatchItem* items, const int* qoff, int* heads, bool vec, int shape,
                       int nwg, hipStream_t st)
{
    hipLaunchKernelGGL(batch_reset_kernel<0>, dim3(1), dim3(64), 0, st, heads);
#define FMMK_BATCH(SH, V)                                                                                       \
    hipLaunchKernelGGL((fmm_batch_kernel<TI, TL, TO, SH, V>), dim3(nwg), dim3(Cfg<SH>::NT), 0, st, probs, items, \
                       qoff, heads)
    if (shape == 2) {
        if (vec) FMMK_BATCH(2, true);
        else FMMK_BATCH(2, false);
    } else if (shape == 3) {
        hipLaunchKernelGGL((fmm_batch_fast_kernel<TI, TL, TO>), dim3(nwg), dim3(kFastNT), 0, st, probs, items, qoff,
                           heads);
    } else {
        if (vec) FMMK_BATCH(0, true);
        else FMMK_BATCH(0, false);
    }
#undef FMMK_BATCH
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}

}  // namespace fmmk

// Per-datatype launchers (one translation unit each).
int fmmk_launch_d64(const fmmk::KArgs& ka, int shape, bool a_mn, bool b_mn, bool c_col, int nwg, hipStream_t st);
int fmmk_launch_d_f32leaf(const fmmk::KArgs& ka, int shape, bool a_mn, bool b_mn, bool c_col, int nwg,
                          hipStream_t st);
// FP32 pre-summed leaves of FP64 operands, FP32 MFMA leaf, FP64 C or slots
// (column-major A, B, C; shapes 0, 2 and 3)
int fmmk_launch_f32leaves_d(const fmmk::KArgs& ka, int shape, int nwg, hipStream_t st);
int fmmk_launch_s32(const fmmk::KArgs& ka, int shape, bool a_mn, bool b_mn, bool c_col, int nwg, hipStream_t st);
// items: kBatchQueues queues, queue q at items[qoff[q] .. qoff[q+1]); heads: kBatchQueues counters
int fmmk_launch_batch_d64(const fmmk::KArgs* probs, const fmmk::BatchItem* items, const int* qoff, int* heads,
                          bool vec, int shape, int nwg, hipStream_t st);
int fmmk_launch_batch_s32(const fmmk::KArgs* probs, const fmmk::BatchItem* items, const int* qoff, int* heads,
                          bool vec, int shape, int nwg, hipStream_t st);

#endif
