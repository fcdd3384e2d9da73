// fmmhip_int8x4.hip -- the INT8x4 leaf of the FP64 FMM (and classical) GEMM.
//
// Leaf products run on the int8 MFMA (v_mfma_i32_16x16x64_i8, twice the
// bf16 rate) with exact int32 accumulation; operand summation, the
// recombination of the slice products and the W write-back into C stay
// FP64.  The method is the slicing of Ozaki, Ogita, Oishi and Rump
// ("Error-free transformations of matrix multiplication by using fast
// routines of matrix multiplication", Numer. Algorithms 2012) with integer
// slices, as Ootomo, Ozaki and Yokota run it on integer matrix units ("DGEMM
// on integer matrix multiplication unit", IJHPCA 2024): every value becomes
// a few small integers with one scale per matrix row, the integer products
// are exact, and only the slicing and the dropped slice products err.
// Three stages per call:
//
//  1. row maxima (HBM-bound): for every product r, the U- (V-) weighted sum
//     x of the operand's parts is formed in FP64 (as the pre-summed pack
//     mode does) per block (64 rows x 64 k where rows are contiguous in
//     memory, 16 x 256 where k is), and the block's row maxima
//     of |x| go into MX[r][row] by an atomic maximum over the double's bits
//     (a non-finite row stores a quiet NaN's bits, above every finite one).
//  2. slices (HBM-bound): the same sums again; with mx = f 2^E the row
//     maximum (f in [1/2, 1)), y = x 2^-E (127 / f) has |y| <= 127 and
//       d0 = rint(y), d1 = rint(254 (y - d0)), d2 = rint(254 (.. - d1)),
//       d3 = rint(254 (.. - d2)),  y = d0 + d1/254 + d2/254^2 + d3/254^3,
//     every digit in [-127, 127] (a remainder is at most 1/2, times 254),
//     written as four int8 planes; the row's inverse scale mx / 127 goes to
//     SC[r][row] (NaN for a non-finite row, 0 for a zero row).
//     Representation error: half a unit of d3, 0.5 / (127 * 254^3) = 2.4e-10
//     (about 2^-32) of the row maximum.
//  3. GEMM: one workgroup per 128 x 128 leaf tile of C, all R products; per
//     64-wide k step the ten slice products of weight at least 254^-3 into
//     four int32 accumulators, one per weight:
//       G0 = sum d0 d0',  G1 = sum d0 d1' + d1 d0',
//       G2 = sum d0 d2' + d1 d1' + d2 d0',
//       G3 = sum d0 d3' + d1 d2' + d2 d1' + d3 d0'.
//     Each term is at most 127^2 = 16129 in magnitude, so G3 stays exact for
//     k <= 32768 (4 * 16129 * 32768 < 2^31); the host splits longer k.  At
//     the product's end, in FP64, c = (G0 + (G1 + (G2 + G3/254)/254)/254)
//     scaled by the row's and the column's inverse scales, and the W
//     epilogue scatters alpha*w*c into the C parts (race-free: the workgroup
//     owns the tile in every part).  The dropped products (weights 254^-4
//     and below) are each below 254^-4 (about 2^-32) of the row maximum times the column
//     maximum.
//
// Slices in HBM: P[r][g][kb][row][64 bytes] (kb = k/64): one k step of a
// 128-row tile is 8 KB contiguous, the same bytes per step as the BF16x3
// leaf's planes, so the moves into LDS are the same.  Rows are padded to a
// multiple of 128 and k to a multiple of 256 with zeros.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "fmmhip_host.h"

namespace i8s {

typedef int intx4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));   // 16 bytes, a native vector
typedef __attribute__((address_space(3))) void lds_void_t;
template <int I> struct IntC { static constexpr int value = I; };

constexpr int KSTEP = 64;        // k per MFMA stage (bytes per row and slice)
constexpr int CHUNK = 256;       // k padding of the slices
constexpr int ROW_PAD = 128;     // slice rows are padded to this
// Split blocks: rows x k per block.  Rows contiguous in memory (A column-
// major): 64 x 64, each k column 512 contiguous bytes and each slice's 64
// rows x 64 bytes one contiguous write (16 x 256 blocks touched 256 lines
// a leading dimension apart and re-read them from HBM for every product);
// k contiguous: 16 x 256, each row 2 KB contiguous.
template <bool ROWC> struct Blk {
    static constexpr int SR = ROWC ? 64 : 16, CK = ROWC ? 64 : 256;
    static constexpr int TPR = 256 / SR;       // threads per row in the row passes
};
constexpr int SLICES = 4;
constexpr long KMAX = 32768;     // k per product for exact int32 sums
constexpr unsigned long long kBadBits = 0x7FF8000000000000ULL;   // a quiet NaN: above +inf and every finite

// ---- row maxima and slices ----------------------------------------------------

struct SplitArgs {
    const double* X;
    long sr, sk;                 // element (row i, k) of a part at X[off + i*sr + k*sk]
    const int* ptr;              // CSR of this operand's entries over products
    const fmm_entry_t* ent;
    int swap;                    // entries' rows are the k extent (operand B)
    int Lp, ldk;                 // padded rows, padded k
    int8_t* P;                   // slices
    unsigned long long* MX;      // row maxima (double bits) [r][row]
    double* SC;                  // inverse scales [r][row]
    int R;
    int nrc;                     // (row block, k block) pairs: (Lp / SR) * (ldk / CK)
};

// Block ids are dealt so that the R blocks of one (row block, chunk) run
// back to back on one XCD (ids go round-robin to the 8 XCDs): the parts
// they all read come from that XCD's L2 after the first product.
template <bool ROWC>
__device__ __forceinline__ bool block_of(const SplitArgs& a, int& r, int& row0, int& k0, int& ch)
{
    typedef Blk<ROWC> K;
    const int xcd = (int)blockIdx.x & 7, t8 = (int)blockIdx.x >> 3;
    r = t8 % a.R;
    const int q = (t8 / a.R) * 8 + xcd;             // (row block, chunk) index
    if (q >= a.nrc) return false;
    const int nrb = a.Lp / K::SR;
    row0 = (q % nrb) * K::SR;
    ch = q / nrb;
    k0 = ch * K::CK;
    return true;
}

// The block's weighted sum of the parts into T (zero outside their
// extents): one part at a time, the thread's elements of it loaded together
// (clamped into the part, so every load is in bounds and unconditional).
template <bool ROWC>
__device__ __forceinline__ void sum_to_lds(const SplitArgs& a, int r, int row0, int k0,
                                           double (*T)[Blk<ROWC>::CK + 1])
{
    constexpr int SR = Blk<ROWC>::SR, CK = Blk<ROWC>::CK;
    constexpr int PER = SR * CK / 256;
    const int t = threadIdx.x;
    double v[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) v[i] = 0.0;
    const int e0 = a.ptr[r], e1 = a.ptr[r + 1];
    for (int e = e0; e < e1; e++) {
        const fmm_entry_t en = a.ent[e];
        const int er = a.swap ? en.cols : en.rows;
        const int ek = a.swap ? en.rows : en.cols;
        if (er <= 0 || ek <= 0) continue;             // an empty part (dimension below the split)
        const double* __restrict__ Xe = a.X + en.off;
        double x[PER];
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int idx = t + 256 * i;
            const int gi = row0 + (ROWC ? (idx % SR) : (idx / CK));
            const int gk = k0 + (ROWC ? (idx / SR) : (idx % CK));
            x[i] = Xe[(long)min(gi, er - 1) * a.sr + (long)min(gk, ek - 1) * a.sk];
        }
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int idx = t + 256 * i;
            const int gi = row0 + (ROWC ? (idx % SR) : (idx / CK));
            const int gk = k0 + (ROWC ? (idx / SR) : (idx % CK));
            if (gi < er && gk < ek) v[i] = fma(en.coef, x[i], v[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int idx = t + 256 * i;
        T[ROWC ? (idx % SR) : (idx / CK)][ROWC ? (idx / SR) : (idx % CK)] = v[i];
    }
}

template <bool ROWC>
__global__ __launch_bounds__(256) void rowmax_kernel(SplitArgs a)
{
    typedef Blk<ROWC> K;
    __shared__ double T[K::SR][K::CK + 1];
    int r, row0, k0, ch;
    if (!block_of<ROWC>(a, r, row0, k0, ch)) return;
    sum_to_lds<ROWC>(a, r, row0, k0, T);
    __syncthreads();
    const int row = threadIdx.x / K::TPR, s = threadIdx.x % K::TPR;
    double mx = 0.0;
    int bad = 0;
#pragma unroll
    for (int j = 0; j < K::CK / K::TPR; j++) {
        const double x = T[row][s + K::TPR * j];
        mx = fmax(mx, fabs(x));
        bad |= !isfinite(x);
    }
#pragma unroll
    for (int o = 1; o < K::TPR; o <<= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        bad |= __shfl_xor(bad, o);
    }
    if (s == 0 && (bad || mx > 0.0))
        atomicMax(a.MX + (long)r * a.Lp + row0 + row, bad ? kBadBits : (unsigned long long)__double_as_longlong(mx));
}

template <bool ROWC>
__global__ __launch_bounds__(256) void slice_kernel(SplitArgs a)
{
    typedef Blk<ROWC> K;
    constexpr int SR = K::SR, CK = K::CK, TPR = K::TPR;
    static_assert(SLICES * SR * CK <= SR * (CK + 1) * 8, "the slices fit in the sum's LDS");
    static_assert(CK % KSTEP == 0 && (SLICES * SR * CK / 16) % 256 == 0, "whole 16-byte pieces per thread");
    __shared__ __attribute__((aligned(16))) double T[SR][CK + 1];
    int r, row0, k0, ch;
    if (!block_of<ROWC>(a, r, row0, k0, ch)) return;
    sum_to_lds<ROWC>(a, r, row0, k0, T);
    __syncthreads();
    const int t = threadIdx.x, row = t / TPR, s = t % TPR;
    const double mx = __longlong_as_double((long long)a.MX[(long)r * a.Lp + row0 + row]);
    const bool bad = !isfinite(mx);
    int E = 0;
    double f = 0.5;
    if (!bad && mx > 0.0) f = frexp(mx, &E);        // mx = f 2^E, f in [1/2, 1)
    const double up = 127.0 / f;                     // |x| 2^-E up <= 127
    double x[CK / TPR];
#pragma unroll
    for (int j = 0; j < CK / TPR; j++) x[j] = T[row][s + TPR * j];
    __syncthreads();                                 // T is reused for the slices
    int8_t* O = (int8_t*)&T[0][0];                   // O[g][row][CK]
#pragma unroll
    for (int j = 0; j < CK / TPR; j++) {
        const int kk = s + TPR * j;
        double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
        if (!bad) {
            // y - rint(y) is exact; a remainder of at most 1/2 times 254 is
            // at most 127, so every digit fits int8
            const double y = ldexp(x[j], -E) * up;
            d0 = rint(y);
            const double r1 = (y - d0) * 254.0;
            d1 = rint(r1);
            const double r2 = (r1 - d1) * 254.0;
            d2 = rint(r2);
            d3 = rint((r2 - d2) * 254.0);
        }
        O[(0 * SR + row) * CK + kk] = (int8_t)(int)d0;
        O[(1 * SR + row) * CK + kk] = (int8_t)(int)d1;
        O[(2 * SR + row) * CK + kk] = (int8_t)(int)d2;
        O[(3 * SR + row) * CK + kk] = (int8_t)(int)d3;
    }
    if (s == 0 && ch == 0) a.SC[(long)r * a.Lp + row0 + row] = bad ? __longlong_as_double((long long)kBadBits) : mx / 127.0;
    __syncthreads();
    // write out: per (slice, k step) SR rows of 64 B, contiguous
    const long plane = (long)a.ldk * a.Lp;
#pragma unroll
    for (int i = 0; i < SLICES * SR * CK / 16 / 256; i++) {
        const int idx = t + 256 * i;
        const int g = idx / (SR * CK / 16);
        const int rem = idx % (SR * CK / 16);
        const int kbl = rem / (SR * 4);
        const int rem2 = rem % (SR * 4);
        const int rr = rem2 >> 2, c = rem2 & 3;
        const uint4 v = *(const uint4*)&O[(g * SR + rr) * CK + kbl * KSTEP + c * 16];
        int8_t* dst = a.P + ((long)r * SLICES + g) * plane + ((long)(k0 / KSTEP + kbl) * a.Lp + row0 + rr) * KSTEP + c * 16;
        *(uint4*)dst = v;
    }
}

// Clears n 8-byte words (the row maxima before their atomic maxima): a fill
// kernel rather than a memset node, so the call can be captured in a graph.
__global__ __launch_bounds__(256) void zero_kernel(unsigned long long* p, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = 0ULL;
}

// ---- GEMM -------------------------------------------------------------------

struct GemmArgs {
    const int8_t* A;
    const double* SA;            // inverse row scales [r][LpA]
    const int8_t* B;
    const double* SB;            // inverse column scales [r][LpB]
    long a_plane, b_plane;       // bytes per slice
    int LpA, LpB;
    int nkb;                     // k steps per product (a multiple of 4)
    int R;
    const int* wptr;             // W CSR over products (absolute entry indices)
    const fmm_entry_t* ent;
    double* C;
    long rsC, csC;
    double alpha, beta;
    int tiles_m, tiles_n;
    int min_cm, min_cn;          // tiles inside [0,min_cm) x [0,min_cn) are inside every C part
    // product split: `groups` workgroups per tile, group g taking products
    // [g*R/groups, (g+1)*R/groups) into its zeroed temporary tmp + g*tmp_elems
    // (C's strides), first writers from bit gbit of the entries' flags
    int groups, gbit;
    double* tmp;
    long tmp_elems;
};

template <int WM, int WN, int FI, int FJ>
struct GCfg {
    static constexpr int NT = WM * WN * 64;
    static constexpr int BM = WM * FI * 16, BN = WN * FJ * 16;
    static constexpr int A_CH = SLICES * BM * 4, B_CH = SLICES * BN * 4;   // 16-byte chunks per stage
};

template <int WM, int WN, int FI, int FJ>
__global__ __attribute__((amdgpu_flat_work_group_size(WM * WN * 64, WM * WN * 64),
                          amdgpu_waves_per_eu(WM * WN / 4, WM * WN / 4))) void int8x4_gemm_kernel(GemmArgs a)
{
    typedef GCfg<WM, WN, FI, FJ> G;
    constexpr int BM = G::BM, BN = G::BN, NT = G::NT;
    static_assert(BM + BN <= NT, "one scale per thread");
    // two LDS stages as separate objects, so that the compiler tells the
    // direct-to-LDS writes into one from the fragment reads of the other
    __shared__ u32x4 sA0[G::A_CH], sA1[G::A_CH];
    __shared__ u32x4 sB0[G::B_CH], sB1[G::B_CH];
    __shared__ double sS[2][BM + BN];

    // blocks [g*tiles, (g+1)*tiles) are product group g
    const int tiles = a.tiles_m * a.tiles_n;
    const int grp = (int)blockIdx.x / tiles;
    const int r_begin = (int)((long)grp * a.R / a.groups), r_end = (int)((long)(grp + 1) * a.R / a.groups);
    double* const Cb = a.groups > 1 ? a.tmp + (long)grp * a.tmp_elems : a.C;
    const double beta_c = a.groups > 1 ? 0.0 : a.beta;      // temporaries start at zero
    int tm, tn;
    fmmk::tile_of((int)blockIdx.x - grp * tiles, a.tiles_m, a.tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave % WM) * (FI * 16), wn = (wave / WM) * (FJ * 16);

    // One k step of the A and B slices goes from HBM straight into an LDS
    // stage: 2 x 8 SLICES moves of 1 KB (16 rows x 64 bytes of one slice),
    // each one global_load_lds_dwordx4 in which lane i writes bytes
    // [16 i, 16 i + 16) of the move's LDS block, i.e. row i / 4 of the block,
    // stored chunk i % 4, which holds the row's 16-byte k chunk
    // (i % 4) ^ ((i >> 3) & 3) (a row swizzle: the fragment reads of eight
    // consecutive rows hit distinct banks).  Move q covers slice q / 8, rows
    // 16 (q % 8) .. + 15; wave w makes moves MPW w .. MPW w + MPW - 1 of each
    // operand.  The scales of a product's rows and columns go through a
    // register into LDS with its first step.
    static_assert(BM == 128 && BN == 128, "eight 16-row moves per slice and operand");
    constexpr int MPW = 8 * SLICES / (NT / 64);   // moves per wave and operand
    static_assert(MPW * (NT / 64) == 8 * SLICES, "the moves divide among the waves");
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int lane_off = (lane >> 2) * KSTEP + 16 * ((lane & 3) ^ ((lane >> 3) & 3));
    double re = 0.0;
    auto stage_in = [&](u32x4* dA, u32x4* dB, int r_, int kb_) {
        const int8_t* Ab_ = a.A + (long)r_ * SLICES * a.a_plane + ((long)kb_ * a.LpA + m0) * KSTEP + lane_off;
        const int8_t* Bb_ = a.B + (long)r_ * SLICES * a.b_plane + ((long)kb_ * a.LpB + n0) * KSTEP + lane_off;
#pragma unroll
        for (int t = 0; t < MPW; t++) {
            const int q = MPW * wv + t, g = q >> 3, rb = q & 7;
            __builtin_amdgcn_global_load_lds((const void*)(Ab_ + g * a.a_plane + rb * 16 * KSTEP),
                                             (lds_void_t*)(dA + 64 * q), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(Bb_ + g * a.b_plane + rb * 16 * KSTEP),
                                             (lds_void_t*)(dB + 64 * q), 16, 0, 0);
        }
        if (kb_ == 0) {
            if (tid < BM) re = a.SA[(long)r_ * a.LpA + m0 + tid];
            else if (tid < BM + BN) re = a.SB[(long)r_ * a.LpB + n0 + tid - BM];
        }
    };
    // the moves of this wave (and its scale load) have landed, the scales
    // are stored, then every wave's moves have landed
    auto stage_done = [&](bool prod_start, int ep) {
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        if (prod_start && tid < BM + BN) sS[ep][tid] = re;
        __syncthreads();
    };
    intx4 G0[FI][FJ], G1[FI][FJ], G2[FI][FJ], G3[FI][FJ];
#pragma unroll
    for (int i = 0; i < FI; i++)
#pragma unroll
        for (int j = 0; j < FJ; j++) G0[i][j] = G1[i][j] = G2[i][j] = G3[i][j] = (intx4){0, 0, 0, 0};

    stage_in(sA0, sB0, r_begin, 0);
    stage_done(true, r_begin & 1);

    // fragment reads: row wm + 16 i + (lane & 15) (resp. wn + 16 j + ...),
    // 16-byte k chunk kc = lane >> 4; the swizzle depends on lane & 15 only.
    // A and B fragments take the same k chunk and byte order, so every k
    // pairs an A value with the B value of the same k.
    const int kc = lane >> 4;
    const int rd_a = (wm + (lane & 15)) * 4 + (kc ^ (((lane & 15) >> 1) & 3));
    const int rd_b = (wn + (lane & 15)) * 4 + (kc ^ (((lane & 15) >> 1) & 3));
    // One k step of product r on LDS stage S_ (kb & 1: nkb is even, so every
    // product starts on stage 0) while the next step's slices move into the
    // other stage (read in the previous step, whose barrier every wave has
    // passed).
    auto step = [&](int r, int kb, auto stage) {
        constexpr int S_ = decltype(stage)::value;
        int rn = r, kbn = kb + 1;
        if (kbn == a.nkb) { rn = r + 1; kbn = 0; }
        const bool more = rn < r_end;
        if (more) stage_in(S_ ? sA0 : sA1, S_ ? sB0 : sB1, rn, kbn);

        __builtin_amdgcn_iglp_opt(0);
        const u32x4* la = S_ ? sA1 : sA0;
        const u32x4* lb = S_ ? sB1 : sB0;
        auto fa = [&](int g, int i) { return __builtin_bit_cast(intx4, la[rd_a + (g * BM + 16 * i) * 4]); };
        auto fb = [&](int g, int j) { return __builtin_bit_cast(intx4, lb[rd_b + (g * BN + 16 * j) * 4]); };
#define I8_MFMA(ACC, X, Y) ACC = __builtin_amdgcn_mfma_i32_16x16x64_i8(X, Y, ACC, 0, 0, 0)
        // ten slice products, ordered so that few fragments are live at
        // once: A slice 0 with every B slice, then slices 1, 2, 3 with theirs
        intx4 a0[FI], b0[FJ], b1[FJ], b2[FJ];
#pragma unroll
        for (int i = 0; i < FI; i++) a0[i] = fa(0, i);
#pragma unroll
        for (int j = 0; j < FJ; j++) b0[j] = fb(0, j);
#pragma unroll
        for (int i = 0; i < FI; i++)
#pragma unroll
            for (int j = 0; j < FJ; j++) I8_MFMA(G0[i][j], a0[i], b0[j]);
#pragma unroll
        for (int j = 0; j < FJ; j++) b1[j] = fb(1, j);
#pragma unroll
        for (int i = 0; i < FI; i++)
#pragma unroll
            for (int j = 0; j < FJ; j++) I8_MFMA(G1[i][j], a0[i], b1[j]);
#pragma unroll
        for (int j = 0; j < FJ; j++) b2[j] = fb(2, j);
#pragma unroll
        for (int i = 0; i < FI; i++)
#pragma unroll
            for (int j = 0; j < FJ; j++) I8_MFMA(G2[i][j], a0[i], b2[j]);
        {
            intx4 b3[FJ];
#pragma unroll
            for (int j = 0; j < FJ; j++) b3[j] = fb(3, j);
#pragma unroll
            for (int i = 0; i < FI; i++)
#pragma unroll
                for (int j = 0; j < FJ; j++) I8_MFMA(G3[i][j], a0[i], b3[j]);
        }
        {
            intx4 a1[FI];
#pragma unroll
            for (int i = 0; i < FI; i++) a1[i] = fa(1, i);
#pragma unroll
            for (int i = 0; i < FI; i++)
#pragma unroll
                for (int j = 0; j < FJ; j++) {
                    I8_MFMA(G1[i][j], a1[i], b0[j]);
                    I8_MFMA(G2[i][j], a1[i], b1[j]);
                    I8_MFMA(G3[i][j], a1[i], b2[j]);
                }
        }
        {
            intx4 a2[FI];
#pragma unroll
            for (int i = 0; i < FI; i++) a2[i] = fa(2, i);
#pragma unroll
            for (int i = 0; i < FI; i++)
#pragma unroll
                for (int j = 0; j < FJ; j++) {
                    I8_MFMA(G2[i][j], a2[i], b0[j]);
                    I8_MFMA(G3[i][j], a2[i], b1[j]);
                }
        }
        {
            intx4 a3[FI];
#pragma unroll
            for (int i = 0; i < FI; i++) a3[i] = fa(3, i);
#pragma unroll
            for (int i = 0; i < FI; i++)
#pragma unroll
                for (int j = 0; j < FJ; j++) I8_MFMA(G3[i][j], a3[i], b0[j]);
        }
#undef I8_MFMA

        if (kb == a.nkb - 1) {
            // product end: FP64 recombination with the row and column
            // inverse scales, then the W epilogue, C_part += alpha*w*c; the
            // part's first writer applies beta (beta == 0: C is not read)
            const double* sa = sS[r & 1];
            const double* sb = sa + BM;
            constexpr double inv = 1.0 / 254.0;
            const bool c_in = m0 + BM <= a.min_cm && n0 + BN <= a.min_cn;
            const int er0 = wm + 4 * (lane >> 4), ec0 = wn + (lane & 15);
            for (int e = a.wptr[r]; e < a.wptr[r + 1]; e++) {
                const fmm_entry_t we = a.ent[e];
                const double lam = a.alpha * we.coef;
                const bool first = ((we.first >> a.gbit) & 1) != 0;   // first writer within the group
                double* ct = Cb + we.off + (long)(m0 + er0) * a.rsC + (long)(n0 + ec0) * a.csC;
                const int erows = we.rows - m0 - er0, ecols = we.cols - n0 - ec0;
#pragma unroll
                for (int j = 0; j < FJ; j++) {
                    const double cs = sb[ec0 + 16 * j];
#pragma unroll
                    for (int i = 0; i < FI; i++)
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int dr = 16 * i + q, dc = 16 * j;
                            if (c_in || (dr < erows && dc < ecols)) {
                                const double g23 = fma((double)G3[i][j][q], inv, (double)G2[i][j][q]);
                                const double g123 = fma(g23, inv, (double)G1[i][j][q]);
                                const double c = fma(g123, inv, (double)G0[i][j][q]) * sa[er0 + dr] * cs;
                                double* p = ct + (long)dr * a.rsC + (long)dc * a.csC;
                                double old = 0.0;
                                if (!first) old = *p;
                                else if (beta_c != 0.0) old = beta_c * *p;
                                *p = fma(lam, c, old);
                            }
                        }
                }
            }
#pragma unroll
            for (int i = 0; i < FI; i++)
#pragma unroll
                for (int j = 0; j < FJ; j++) G0[i][j] = G1[i][j] = G2[i][j] = G3[i][j] = (intx4){0, 0, 0, 0};
        }
        stage_done(more && kbn == 0, rn & 1);
    };
    for (int r = r_begin; r < r_end; r++) {
        for (int kb = 0; kb < a.nkb; kb += 2) {
            step(r, kb, IntC<0>());
            step(r, kb + 1, IntC<1>());
        }
    }
}

}  // namespace i8s

// ---- host ---------------------------------------------------------------------

namespace {

constexpr int kWM = 2, kWN = 4, kFI = 4, kFJ = 2;
typedef i8s::GCfg<kWM, kWN, kFI, kFJ> CfgI8;

long up_to(long x, long m) { return (x + m - 1) / m * m; }

}  // namespace

int int8x4_gemm(fmmhip_alg_t alg, long m, long n, long k, double alpha, const double* A, long rsA, long csA,
                const double* B, long rsB, long csB, double beta, double* C, long rsC, long csC, hipStream_t st)
{
    if (!alg || m < 0 || n < 0 || k < 0) return FMMHIP_ERR_INVALID;
    if (m == 0 || n == 0) return FMMHIP_SUCCESS;
    if (!C) return FMMHIP_ERR_INVALID;
    if (k == 0 || alpha == 0.0) return fmmhip_scale_d(C, m, n, rsC, csC, beta, st);
    if (!A || !B) return FMMHIP_ERR_INVALID;
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;
    // the int32 sums are exact for leaf k up to KMAX: longer k runs as
    // consecutive calls over k ranges of KMAX leaf columns, C += A_p B_p
    const long kpiece = i8s::KMAX * alg->kt;
    if (k > kpiece) {
        for (long k0 = 0; k0 < k; k0 += kpiece) {
            const int rc = int8x4_gemm(alg, m, n, std::min(kpiece, k - k0), alpha, A + k0 * csA, rsA, csA,
                                       B + k0 * rsB, rsB, csB, k0 ? 1.0 : beta, C, rsC, csC, st);
            if (rc) return rc;
        }
        return FMMHIP_SUCCESS;
    }

    fmm_geom_t g;
    memset(&g, 0, sizeof(g));
    g.m = m; g.n = n; g.k = k;
    g.lm = (m + alg->mt - 1) / alg->mt;
    g.lk = (k + alg->kt - 1) / alg->kt;
    g.ln = (n + alg->nt - 1) / alg->nt;
    g.rsA = rsA; g.csA = csA; g.rsB = rsB; g.csB = csB; g.rsC = rsC; g.csC = csC;
    const int* dptr = nullptr;
    const fmm_entry_t* dent = nullptr;
    // held from the descriptor lookups to the last launch: another call
    // cannot evict the cached lists or regrow the workspace in between
    std::unique_lock<std::mutex> lk = enqueue_lock();
    int rc = fused_desc(alg, g, st, &dptr, &dent);
    if (rc) return rc;

    const int R = alg->R;
    const long LpA = up_to(g.lm, i8s::ROW_PAD), LpB = up_to(g.ln, i8s::ROW_PAD);
    const long ldk = up_to(g.lk, i8s::CHUNK);
    if (LpA > INT32_MAX || LpB > INT32_MAX || ldk > i8s::KMAX) return FMMHIP_ERR_UNSUPPORTED;
    const size_t slA = (size_t)R * i8s::SLICES * LpA * ldk, slB = (size_t)R * i8s::SLICES * LpB * ldk;
    const size_t rowA = (size_t)R * LpA * 8, rowB = (size_t)R * LpB * 8;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // slices A | slices B | maxima A, B (cleared together) | scales A | scales B
    const size_t oPB = al(slA), oMA = oPB + al(slB), oMB = oMA + rowA, oSA = oMB + al(rowB), oSB = oSA + al(rowA);
    const long tiles = (LpA / CfgI8::BM) * (LpB / CfgI8::BN);
    // products materialised (pack 2, or automatic when the slots fit in half
    // of the device's memory): the kernel stores each product once and the
    // W stage applies W -- the kernel's W epilogue then only stores
    bool wmat = R > 1 && (alg->pack == 2 || alg->pack < 0);
    if (wmat && alg->pack < 0) wmat = 8.0 * R * (double)g.lm * (double)g.ln < fmmhip_ws_bound();
    MatView mv;
    memset(&mv, 0, sizeof(mv));
    if (wmat && (rc = mat_view(alg, g, st, &mv))) return rc;
    int cus = 256;
    fmmhip_device_cu_count(&cus);
    // product split (int8 MFMA rate, ten slice products per leaf product;
    // one workgroup per CU, LDS); materialised, any count up to R, no copies
    const double c_span = (double)(m - 1) * rsC + (double)(n - 1) * csC + 1.0;   // the temporaries' extent
    const int groups = wmat ? (alg->split > 0 ? std::min(alg->split, R) : mat_groups(R, tiles, cus))
                            : split_groups(alg->split, R, tiles,
                                           2.0 * 10 * R * (double)g.lm * (double)g.lk * (double)g.ln, 8.0 * c_span,
                                           false, 5.0e15, 1);
    const long tmp_elems = (!wmat && groups > 1) ? (((m - 1) * rsC + (n - 1) * csC + 1 + 31) & ~31L) : 0;
    const size_t oT = oSB + al(rowB);        // split temporaries, or the product slots
    const size_t bytes = oT + (wmat ? (size_t)R * (size_t)mv.slot * 8 : (size_t)groups * tmp_elems * 8);
    char* ws = (char*)workspace_for(st, bytes);
    if (!ws) return FMMHIP_ERR_NOMEM;

    i8s::SplitArgs sa;
    sa.X = A; sa.sr = rsA; sa.sk = csA; sa.ptr = dptr; sa.ent = dent; sa.swap = 0;
    sa.Lp = (int)LpA; sa.ldk = (int)ldk; sa.P = (int8_t*)ws;
    sa.MX = (unsigned long long*)(ws + oMA); sa.SC = (double*)(ws + oSA);
    i8s::SplitArgs sb;
    sb.X = B; sb.sr = csB; sb.sk = rsB; sb.ptr = dptr + (R + 1); sb.ent = dent; sb.swap = 1;
    sb.Lp = (int)LpB; sb.ldk = (int)ldk; sb.P = (int8_t*)(ws + oPB);
    sb.MX = (unsigned long long*)(ws + oMB); sb.SC = (double*)(ws + oSB);

    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    const bool timed = timing_active();
    if (timed) timing_record(&e0, st);
    const dim3 blk(256);
    // (row block, k block) pairs of each operand for its block shape
    auto nrc = [&](long Lp, bool rowc) {
        return rowc ? Lp / i8s::Blk<true>::SR * (ldk / i8s::Blk<true>::CK)
                    : Lp / i8s::Blk<false>::SR * (ldk / i8s::Blk<false>::CK);
    };
    const long nrcA = nrc(LpA, rsA == 1), nrcB = nrc(LpB, csB == 1);
    if ((double)up_to(nrcA, 8) * R > INT32_MAX || (double)up_to(nrcB, 8) * R > INT32_MAX)
        return FMMHIP_ERR_UNSUPPORTED;
    sa.R = sb.R = R;
    sa.nrc = (int)nrcA;
    sb.nrc = (int)nrcB;
    // grid: (row block, chunk) pairs rounded up to 8 (one run per XCD) x products
    const dim3 gA((unsigned)(up_to(sa.nrc, 8) * R));
    const dim3 gB((unsigned)(up_to(sb.nrc, 8) * R));
    const long nmax = (long)((oSA - oMA) / 8);   // both maxima arrays and the padding between
    hipLaunchKernelGGL(i8s::zero_kernel, dim3((unsigned)std::min((nmax + 255) / 256, 4096L)), blk, 0, st,
                       (unsigned long long*)(ws + oMA), nmax);
    if (rsA == 1) hipLaunchKernelGGL(i8s::rowmax_kernel<true>, gA, blk, 0, st, sa);
    else hipLaunchKernelGGL(i8s::rowmax_kernel<false>, gA, blk, 0, st, sa);
    if (csB == 1) hipLaunchKernelGGL(i8s::rowmax_kernel<true>, gB, blk, 0, st, sb);
    else hipLaunchKernelGGL(i8s::rowmax_kernel<false>, gB, blk, 0, st, sb);
    if (rsA == 1) hipLaunchKernelGGL(i8s::slice_kernel<true>, gA, blk, 0, st, sa);
    else hipLaunchKernelGGL(i8s::slice_kernel<false>, gA, blk, 0, st, sa);
    if (csB == 1) hipLaunchKernelGGL(i8s::slice_kernel<true>, gB, blk, 0, st, sb);
    else hipLaunchKernelGGL(i8s::slice_kernel<false>, gB, blk, 0, st, sb);
    if (hipGetLastError() != hipSuccess) return FMMHIP_ERR_HIP;
    if (!wmat && groups > 1 && (rc = fmmhip_split_zero(1, ws + oT, tmp_elems, groups, st))) return rc;
    if (timed) timing_record(&e1, st);

    i8s::GemmArgs ga;
    ga.A = (const int8_t*)ws; ga.SA = (const double*)(ws + oSA);
    ga.B = (const int8_t*)(ws + oPB); ga.SB = (const double*)(ws + oSB);
    ga.a_plane = LpA * ldk; ga.b_plane = LpB * ldk;
    ga.LpA = (int)LpA; ga.LpB = (int)LpB;
    ga.nkb = (int)(ldk / i8s::KSTEP);
    ga.R = R;
    ga.wptr = dptr + 2 * (R + 1);
    ga.ent = dent;
    ga.C = C; ga.rsC = rsC; ga.csC = csC;
    ga.alpha = alpha; ga.beta = beta;
    ga.tiles_m = (int)(LpA / CfgI8::BM);
    ga.tiles_n = (int)(LpB / CfgI8::BN);
    ga.min_cm = (int)(m / alg->mt);
    ga.min_cn = (int)(n / alg->nt);
    ga.groups = groups;
    ga.gbit = groups == 8 ? 3 : groups == 4 ? 2 : groups == 2 ? 1 : 0;
    ga.tmp = groups > 1 ? (double*)(ws + oT) : nullptr;
    ga.tmp_elems = tmp_elems;
    if (wmat) {
        // slot r of lm x ln per product: one entry each, coefficient 1, first writer
        ga.wptr = mv.wptr;
        ga.ent = mv.ent;
        ga.C = (double*)(ws + oT);
        ga.rsC = mv.swap ? mv.ldm : 1;
        ga.csC = mv.swap ? 1 : mv.ldm;
        ga.alpha = 1.0;
        ga.beta = 0.0;
        ga.min_cm = (int)g.lm;
        ga.min_cn = (int)g.ln;
        ga.gbit = 0;
        ga.tmp = ga.C;           // every group's Cb: the slots
        ga.tmp_elems = 0;
    }
    const long nwg = (long)ga.tiles_m * ga.tiles_n * groups;
    if (nwg > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;
    hipLaunchKernelGGL((i8s::int8x4_gemm_kernel<kWM, kWN, kFI, kFJ>), dim3((unsigned)nwg), dim3(CfgI8::NT), 0, st,
                       ga);
    if (hipGetLastError() != hipSuccess) return FMMHIP_ERR_HIP;
    if (timed) {
        timing_record(&e2, st);
        timing_push(e0, e1, e2, true);
    }
    if (wmat) {
        fmmhip_wsum_t w;
        memset(&w, 0, sizeof(w));
        w.M = ws + oT;
        w.ldm = mv.ldm;
        w.swap = mv.swap;
        w.Lc = mv.Lc;
        w.Ls = mv.Ls;
        w.C = C;
        w.rsC = rsC;
        w.csC = csC;
        w.alpha = alpha;
        w.beta = beta;
        w.np = mv.np;
        w.parts = mv.parts;
        w.terms = mv.terms;
        w.nslots = R;
        w.ngroups = 1;
        w.gparts = mv.np;
        return fmmhip_wsum(1, &w, st);
    }
    if (groups > 1) return fmmhip_split_reduce(1, C, m, n, rsC, csC, beta, ws + oT, tmp_elems, groups, st);
    return FMMHIP_SUCCESS;
}
