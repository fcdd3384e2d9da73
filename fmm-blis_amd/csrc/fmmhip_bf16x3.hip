// fmmhip_bf16x3.hip -- the BF16x3 leaf of the FP64 FMM (and classical) GEMM.
//
// Leaf products run on the bf16 MFMA with FP32 accumulation; operand
// summation, the recombination of split terms and the W write-back into C
// stay FP64.  Two stages per call:
//
//  1. split (HBM-bound): for every product r, the U- (V-) weighted sum of
//     the operand's parts is formed in FP64 (as the pre-summed pack mode
//     does) and written as bf16 planes with one power-of-two scale per
//     (row, 256-wide k chunk).  With y = x 2^-e, e chosen so that |y| < 256
//     over the chunk row, the digits are d0 = rint(y), d1 = rint(512 (y -
//     d0)), then either
//       three planes:  P0 = d0, P1 = d1, P2 = bf16(s2) with s2 = 512 (y - d0)
//                      - d1 (|s2| <= 1/2): y = P0 + 2^-9 (P1 + P2),
//                      representation error <= 2^-27 of the chunk row max;
//       four planes:   d2 = rint(512 s2), s3 = 512 s2 - d2, P2 = d2 2^-9
//                      (exact), P3 = bf16(s3 2^-9): y = P0 + 2^-9 (P1 + P2 +
//                      P3), representation error <= 2^-36 of the row max.
//     d0, d1, d2 are integers in [-256, 256], exact in bf16.
//  2. GEMM: one workgroup per 128 x 128 leaf tile of C, all R products; per
//     32-wide k step, the plane products on v_mfma_f32_16x16x32_bf16 into
//     three FP32 accumulators of weights 1, 2^-9, 2^-18:
//       three planes, eight terms:
//         G0 = sum P0 P0',  G1 = sum P0 P1' + P1 P0' + P0 P2' + P2 P0',
//         G2 = sum P1 P1' + P1 P2' + P2 P1'   (six terms: P1 P1' only);
//       four planes, ten terms:
//         G0 = sum P0 P0',
//         G1 = sum P0 P1' + P1 P0' + P0 P2' + P2 P0' + P0 P3' + P3 P0',
//         G2 = sum P1 P1' + P1 P2' + P2 P1'.
//     G0 is a sum of at most 256 integer products of magnitude <= 2^16, so
//     it is exact in FP32; at every chunk end the three are recombined in
//     FP64, acc += 2^(eA_i + eB_j) * (G0 + 2^-9 (G1 + 2^-9 G2)), and after
//     the last chunk the W epilogue scatters alpha*w*acc into the C parts
//     (race-free: the workgroup owns the tile in every part).
//     Error (normalised, uniform[-1,1]): three planes are limited by the
//     representation; four planes by the FP32 rounding of G1, which the bf16
//     MFMA applies after every 8 k-terms of a product (measured): about 13x
//     below three planes.  FMM algorithms need that margin, as their operand
//     sums and W recombination amplify the leaf error (two-level Strassen at
//     16384^3: 3.4e-11 with four planes, 4.4e-10 with three).  Six terms drop
//     P1 P2', P2 P1' (about 1.6x the eight-term error).
//
// Operand planes in HBM: P[r][g][kb][row][32] (kb = k/32): one k step of a
// 128-row tile is 8 KB contiguous.  Rows are padded to a multiple of 128
// and k to a multiple of 256 with zeros, so the GEMM loads need no bounds.
// Exponents E[r][chunk][row].  Non-finite inputs make their chunk row NaN.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "fmmhip_host.h"

namespace bf3 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));   // 16 bytes, a native vector
typedef __attribute__((address_space(3))) void lds_void_t;
template <int I> struct IntC { static constexpr int value = I; };

constexpr int KSTEP = 32;        // k per MFMA stage
constexpr int CHUNK = 256;       // k per exact FP32 chunk
constexpr int STEPS_PER_CHUNK = CHUNK / KSTEP;
constexpr int ROW_PAD = 128;     // plane rows are padded to this
constexpr int SROWS = 16;        // rows per split block

// ---- split ------------------------------------------------------------------

struct SplitArgs {
    const double* X;
    long sr, sk;                 // element (row i, k) of a part at X[off + i*sr + k*sk]
    const int* ptr;              // CSR of this operand's entries over products
    const fmm_entry_t* ent;
    int swap;                    // entries' rows are the k extent (operand B)
    int Lp, ldk;                 // padded rows, padded k
    u16* P;                      // planes
    int* E;                      // exponents
    int R;                       // products
    int nrc;                     // (row block, chunk) pairs: (Lp / SROWS) * (ldk / CHUNK)
};

__device__ __forceinline__ u16 bf16_exact(double d)   // |d| integer <= 256
{
    return (u16)(__float_as_uint((float)d) >> 16);
}

__device__ __forceinline__ u16 bf16_rne(double d)
{
    uint32_t u = __float_as_uint((float)d);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}

// Block: SROWS rows x one CHUNK of k of one product.  Block ids are mapped
// so that the R blocks of one (row block, chunk) run back to back on one
// XCD (ids are dealt round-robin to the 8 XCDs): the parts they all read
// come from that XCD's L2 after the first product, not from HBM R times.
template <bool ROWC, int PLANES>
__global__ __launch_bounds__(256) void split_kernel(SplitArgs a)
{
    static_assert(PLANES == 3 || PLANES == 4, "three or four planes");
    static_assert(PLANES * SROWS * CHUNK * 2 <= SROWS * (CHUNK + 1) * 8, "the planes fit in the sum's LDS");
    __shared__ __attribute__((aligned(16))) double T[SROWS][CHUNK + 1];
    const int xcd = (int)blockIdx.x & 7, t8 = (int)blockIdx.x >> 3;
    const int r = t8 % a.R;
    const int q = (t8 / a.R) * 8 + xcd;             // (row block, chunk) index
    if (q >= a.nrc) return;
    const int nrb = a.Lp / SROWS;
    const int rb = q % nrb, ch = q / nrb;
    const int row0 = rb * SROWS;
    const int k0 = ch * CHUNK;
    const int t = threadIdx.x;
    const int e0 = a.ptr[r], e1 = a.ptr[r + 1];
    // 1. weighted sum of the parts, zero outside their extents: one part
    //    at a time, the thread's elements of it loaded together
    constexpr int PER = SROWS * CHUNK / 256;
    double v[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) v[i] = 0.0;
    for (int e = e0; e < e1; e++) {
        const fmm_entry_t en = a.ent[e];
        const int er = a.swap ? en.cols : en.rows;
        const int ek = a.swap ? en.rows : en.cols;
        if (er <= 0 || ek <= 0) continue;             // an empty part (dimension below the split)
        const double* __restrict__ Xe = a.X + en.off;
        // every load in bounds (clamped into the part) and unconditional, so
        // they are in flight together; elements outside the part are not added
        double x[PER];
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int idx = t + 256 * i;
            const int gi = row0 + (ROWC ? (idx % SROWS) : (idx / CHUNK));
            const int gk = k0 + (ROWC ? (idx / SROWS) : (idx % CHUNK));
            x[i] = Xe[(long)min(gi, er - 1) * a.sr + (long)min(gk, ek - 1) * a.sk];
        }
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int idx = t + 256 * i;
            const int gi = row0 + (ROWC ? (idx % SROWS) : (idx / CHUNK));
            const int gk = k0 + (ROWC ? (idx / SROWS) : (idx % CHUNK));
            if (gi < er && gk < ek) v[i] = fma(en.coef, x[i], v[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int idx = t + 256 * i;
        T[ROWC ? (idx % SROWS) : (idx / CHUNK)][ROWC ? (idx / SROWS) : (idx % CHUNK)] = v[i];
    }
    __syncthreads();
    // 2. per row: scale from the chunk row max, then the three digits
    const int row = t >> 4, s = t & 15;
    double x[CHUNK / 16];
    double mx = 0.0;
    int bad = 0;
#pragma unroll
    for (int j = 0; j < CHUNK / 16; j++) {
        x[j] = T[row][s + 16 * j];
        mx = fmax(mx, fabs(x[j]));
        bad |= !isfinite(x[j]);
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        bad |= __shfl_xor(bad, o);
    }
    int E = 0;
    if (!bad && mx > 0.0) (void)frexp(mx, &E);      // mx < 2^E
    const int esh = 8 - E;                           // |x| 2^esh < 256
    __syncthreads();                                 // T is reused for the planes
    u16* O = (u16*)&T[0][0];                         // O[g][row][CHUNK]
#pragma unroll
    for (int j = 0; j < CHUNK / 16; j++) {
        const int kk = s + 16 * j;
        u16 p0, p1, p2, p3 = 0;
        if (bad) {
            p0 = 0x7FC0; p1 = 0; p2 = 0;
        } else {
            // every step is exact in FP64 (y - rint(y) and the scalings by
            // 512); only the last plane's bf16 rounding loses bits
            const double y = ldexp(x[j], esh);
            const double d0 = rint(y);
            const double r1 = (y - d0) * 512.0;
            const double d1 = rint(r1);
            const double s2 = r1 - d1;              // |s2| <= 1/2, in units of d1
            p0 = bf16_exact(d0);
            p1 = bf16_exact(d1);
            if (PLANES == 3) {
                p2 = bf16_rne(s2);
            } else {
                const double r2 = s2 * 512.0;
                const double d2 = rint(r2);
                p2 = bf16_exact(d2 * 0x1p-9);       // an integer times 2^-9: exact
                p3 = bf16_rne((r2 - d2) * 0x1p-9);
            }
        }
        O[(0 * SROWS + row) * CHUNK + kk] = p0;
        O[(1 * SROWS + row) * CHUNK + kk] = p1;
        O[(2 * SROWS + row) * CHUNK + kk] = p2;
        if (PLANES == 4) O[(3 * SROWS + row) * CHUNK + kk] = p3;
    }
    if (s == 0) a.E[((long)r * (a.ldk / CHUNK) + ch) * a.Lp + row0 + row] = bad ? 0 : -esh;
    __syncthreads();
    // 3. write out: per (plane, k step) SROWS rows of 64 B, contiguous
    const long plane = (long)a.ldk * a.Lp;
#pragma unroll
    for (int i = 0; i < PLANES * SROWS * CHUNK / 8 / 256; i++) {
        const int idx = t + 256 * i;
        const int g = idx / (SROWS * CHUNK / 8);
        const int rem = idx % (SROWS * CHUNK / 8);
        const int kbl = rem / (SROWS * 4);
        const int rem2 = rem % (SROWS * 4);
        const int rr = rem2 >> 2, c = rem2 & 3;
        const uint4 v = *(const uint4*)&O[(g * SROWS + rr) * CHUNK + kbl * KSTEP + c * 8];
        u16* dst =
            a.P + ((long)r * PLANES + g) * plane + ((long)(k0 / KSTEP + kbl) * a.Lp + row0 + rr) * KSTEP + c * 8;
        *(uint4*)dst = v;
    }
}

// ---- GEMM -------------------------------------------------------------------

struct GemmArgs {
    const u16* A;
    const int* EA;
    const u16* B;
    const int* EB;
    long a_plane, b_plane;       // elements per plane
    long ea_prod, eb_prod;       // exponents per product
    int LpA, LpB;
    int nkb;                     // k steps per product (multiple of STEPS_PER_CHUNK)
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

template <int WM, int WN, int FI, int FJ, int PLANES>
struct GCfg {
    static constexpr int NT = WM * WN * 64;
    static constexpr int BM = WM * FI * 16, BN = WN * FJ * 16;
    static constexpr int A_CH = PLANES * BM * 4, B_CH = PLANES * BN * 4;   // 16-byte chunks per stage
    static constexpr int A_PER = A_CH / NT, B_PER = B_CH / NT;
    static_assert((BM * 4) % NT == 0 && (BN * 4) % NT == 0, "a plane's chunks divide among threads");
    static_assert(BM + BN <= NT, "one exponent per thread");
};

// chunk c (16 bytes = 8 k) of row `row` in an LDS stage: swizzled so that
// the 16-byte reads of eight consecutive rows hit distinct banks
__device__ __forceinline__ int lds_chunk(int row, int c) { return row * 4 + (c ^ ((row >> 1) & 3)); }

template <int WM, int WN, int FI, int FJ, int PLANES, int TERMS>
__global__ __attribute__((amdgpu_flat_work_group_size(WM * WN * 64, WM * WN * 64),
                          amdgpu_waves_per_eu(WM * WN / 4, WM * WN / 4))) void bf16x3_gemm_kernel(GemmArgs a)
{
    static_assert((PLANES == 3 && (TERMS == 6 || TERMS == 8)) || (PLANES == 4 && TERMS == 10), "plane products");
    typedef GCfg<WM, WN, FI, FJ, PLANES> G;
    constexpr int BM = G::BM, BN = G::BN, NT = G::NT;
    // two LDS stages as separate objects, so that the compiler tells the
    // direct-to-LDS writes into one from the fragment reads of the other
    __shared__ u32x4 sA0[G::A_CH], sA1[G::A_CH];
    __shared__ u32x4 sB0[G::B_CH], sB1[G::B_CH];
    __shared__ int sE[2][BM + BN];

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

    // One k step of the A and B planes goes from HBM straight into an LDS
    // stage: 2 x 8 PLANES moves of 1 KB (16 rows x 64 bytes of one plane),
    // each one global_load_lds_dwordx4 in which lane i writes bytes
    // [16 i, 16 i + 16) of the move's LDS block, i.e. row i / 4 of the block,
    // stored chunk i % 4, which holds the row's k chunk (i % 4) ^ ((i >> 3) & 3)
    // (the swizzle of lds_chunk).  Move q covers plane q / 8, rows 16 (q % 8)
    // .. + 15; wave w makes moves MPW w .. MPW w + MPW - 1 of each operand
    // (MPW = 8 PLANES / waves).  No registers and no LDS store instructions are involved; the
    // moves count against vmcnt.  The chunk exponents go through a register.
    static_assert(BM == 128 && BN == 128, "eight 16-row moves per plane and operand");
    constexpr int MPW = 8 * PLANES / (NT / 64);   // moves per wave and operand
    static_assert(MPW * (NT / 64) == 8 * PLANES, "the moves divide among the waves");
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int lane_off = (lane >> 2) * KSTEP + 8 * ((lane & 3) ^ ((lane >> 3) & 3));
    int re = 0;
    auto stage_in = [&](u32x4* dA, u32x4* dB, int r_, int kb_) {
        const u16* Ab_ = a.A + (long)r_ * PLANES * a.a_plane + ((long)kb_ * a.LpA + m0) * KSTEP + lane_off;
        const u16* Bb_ = a.B + (long)r_ * PLANES * a.b_plane + ((long)kb_ * a.LpB + n0) * KSTEP + lane_off;
#pragma unroll
        for (int t = 0; t < MPW; t++) {
            const int q = MPW * wv + t, g = q >> 3, rb = q & 7;
            __builtin_amdgcn_global_load_lds((const void*)(Ab_ + g * a.a_plane + rb * 16 * KSTEP),
                                             (lds_void_t*)(dA + 64 * q), 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void*)(Bb_ + g * a.b_plane + rb * 16 * KSTEP),
                                             (lds_void_t*)(dB + 64 * q), 16, 0, 0);
        }
        if (kb_ % STEPS_PER_CHUNK == 0) {
            const int ch_ = kb_ / STEPS_PER_CHUNK;
            if (tid < BM) re = a.EA[(long)r_ * a.ea_prod + (long)ch_ * a.LpA + m0 + tid];
            else if (tid < BM + BN) re = a.EB[(long)r_ * a.eb_prod + (long)ch_ * a.LpB + n0 + tid - BM];
        }
    };
    // the moves of this wave (and its exponent load) have landed, the
    // exponents are stored, then every wave's moves have landed
    auto stage_done = [&](bool chunk_start, int ep) {
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        if (chunk_start && tid < BM + BN) sE[ep][tid] = re;
        __syncthreads();
    };
    floatx4 G0[FI][FJ], G1[FI][FJ], G2[FI][FJ];
    double acc[FI][FJ][4];
#pragma unroll
    for (int i = 0; i < FI; i++)
#pragma unroll
        for (int j = 0; j < FJ; j++) {
            G0[i][j] = G1[i][j] = G2[i][j] = (floatx4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 4; q++) acc[i][j][q] = 0.0;
        }

    stage_in(sA0, sB0, r_begin, 0);
    stage_done(true, 0);

    // fragment reads: row wm + 16 i + (lane & 15) (resp. wn + 16 j + ...),
    // 8-wide k block kc = lane >> 4; the swizzle depends on lane & 15 only
    const int kc = lane >> 4;
    const int rd_a = (wm + (lane & 15)) * 4 + (kc ^ (((lane & 15) >> 1) & 3));
    const int rd_b = (wn + (lane & 15)) * 4 + (kc ^ (((lane & 15) >> 1) & 3));
    int chunk = 0;        // chunks completed (exponent buffer parity)
    // One k step of product r on LDS stage S_ (kb & 1: nkb is a multiple of
    // STEPS_PER_CHUNK, so every product starts on stage 0) while the next
    // step's planes move into the other stage (read in the previous step,
    // whose barrier every wave has passed).
    auto step = [&](int r, int kb, auto stage) {
        constexpr int S_ = decltype(stage)::value;
        int rn = r, kbn = kb + 1;
        if (kbn == a.nkb) { rn = r + 1; kbn = 0; }
        const bool more = rn < r_end;
        if (more) stage_in(S_ ? sA0 : sA1, S_ ? sB0 : sB1, rn, kbn);

            // the plane products of the step; with four planes LLVM's
            // MFMA/LDS-read interleaving strategy 0 (3% faster than none at
            // 16384^3, classical and 222 L2; strategies 1 and 2 measured 0-4%
            // apart from it); the three-plane forms run 2.6x slower with it
            if (PLANES == 4) __builtin_amdgcn_iglp_opt(0);
            const u32x4* la = S_ ? sA1 : sA0;
            const u32x4* lb = S_ ? sB1 : sB0;
            auto fa = [&](int g, int i) { return __builtin_bit_cast(bf16x8, la[rd_a + (g * BM + 16 * i) * 4]); };
            auto fb = [&](int g, int j) { return __builtin_bit_cast(bf16x8, lb[rd_b + (g * BN + 16 * j) * 4]); };
#define BF3_MFMA(ACC, X, Y) ACC = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X, Y, ACC, 0, 0, 0)
            if (PLANES == 4) {
                // ten terms, ordered so that few fragments are live at once:
                // P0 with every B plane, then P1, P2, P3 with theirs
                bf16x8 a0[FI], b0[FJ], b1[FJ], b2[FJ];
#pragma unroll
                for (int i = 0; i < FI; i++) a0[i] = fa(0, i);
#pragma unroll
                for (int j = 0; j < FJ; j++) b0[j] = fb(0, j);
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) BF3_MFMA(G0[i][j], a0[i], b0[j]);
#pragma unroll
                for (int j = 0; j < FJ; j++) b1[j] = fb(1, j);
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) BF3_MFMA(G1[i][j], a0[i], b1[j]);
#pragma unroll
                for (int j = 0; j < FJ; j++) b2[j] = fb(2, j);
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) BF3_MFMA(G1[i][j], a0[i], b2[j]);
                {
                    bf16x8 b3[FJ];
#pragma unroll
                    for (int j = 0; j < FJ; j++) b3[j] = fb(3, j);
#pragma unroll
                    for (int i = 0; i < FI; i++)
#pragma unroll
                        for (int j = 0; j < FJ; j++) BF3_MFMA(G1[i][j], a0[i], b3[j]);
                }
                {
                    bf16x8 a1[FI];
#pragma unroll
                    for (int i = 0; i < FI; i++) a1[i] = fa(1, i);
#pragma unroll
                    for (int i = 0; i < FI; i++)
#pragma unroll
                        for (int j = 0; j < FJ; j++) {
                            BF3_MFMA(G1[i][j], a1[i], b0[j]);
                            BF3_MFMA(G2[i][j], a1[i], b1[j]);
                            BF3_MFMA(G2[i][j], a1[i], b2[j]);
                        }
                }
                {
                    bf16x8 a2[FI];
#pragma unroll
                    for (int i = 0; i < FI; i++) a2[i] = fa(2, i);
#pragma unroll
                    for (int i = 0; i < FI; i++)
#pragma unroll
                        for (int j = 0; j < FJ; j++) {
                            BF3_MFMA(G1[i][j], a2[i], b0[j]);
                            BF3_MFMA(G2[i][j], a2[i], b1[j]);
                        }
                }
                {
                    bf16x8 a3[FI];
#pragma unroll
                    for (int i = 0; i < FI; i++) a3[i] = fa(3, i);
#pragma unroll
                    for (int i = 0; i < FI; i++)
#pragma unroll
                        for (int j = 0; j < FJ; j++) BF3_MFMA(G1[i][j], a3[i], b0[j]);
                }
            } else {
                bf16x8 a0[FI], b0[FJ], a1[FI], b1[FJ];
#pragma unroll
                for (int i = 0; i < FI; i++) a0[i] = fa(0, i);
#pragma unroll
                for (int j = 0; j < FJ; j++) b0[j] = fb(0, j);
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) BF3_MFMA(G0[i][j], a0[i], b0[j]);
#pragma unroll
                for (int j = 0; j < FJ; j++) b1[j] = fb(1, j);
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) BF3_MFMA(G1[i][j], a0[i], b1[j]);
#pragma unroll
                for (int i = 0; i < FI; i++) a1[i] = fa(1, i);
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) {
                        BF3_MFMA(G1[i][j], a1[i], b0[j]);
                        BF3_MFMA(G2[i][j], a1[i], b1[j]);
                    }
                bf16x8 b2[FJ];
#pragma unroll
                for (int j = 0; j < FJ; j++) b2[j] = fb(2, j);
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) {
                        BF3_MFMA(G1[i][j], a0[i], b2[j]);
                        if (TERMS == 8) BF3_MFMA(G2[i][j], a1[i], b2[j]);
                    }
                bf16x8 a2[FI];
#pragma unroll
                for (int i = 0; i < FI; i++) a2[i] = fa(2, i);
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) {
                        BF3_MFMA(G1[i][j], a2[i], b0[j]);
                        if (TERMS == 8) BF3_MFMA(G2[i][j], a2[i], b1[j]);
                    }
            }
#undef BF3_MFMA

            if (kb % STEPS_PER_CHUNK == STEPS_PER_CHUNK - 1) {
                // chunk end: FP64 recombination with the chunk's row/column scales
                const int* eA = sE[chunk & 1];
                const int* eB = eA + BM;
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++) {
                        const int eb = eB[wn + 16 * j + (lane & 15)];
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int ea = eA[wm + 16 * i + 4 * (lane >> 4) + q];
                            const float g12 = fmaf(G2[i][j][q], 0x1p-9f, G1[i][j][q]);
                            const double t = fma((double)g12, 0x1p-9, (double)G0[i][j][q]);
                            acc[i][j][q] += ldexp(t, ea + eb);
                        }
                        G0[i][j] = G1[i][j] = G2[i][j] = (floatx4){0.f, 0.f, 0.f, 0.f};
                    }
                chunk++;
            }
            if (kb == a.nkb - 1) {
                // product end: W epilogue, C_part += alpha*w*acc; the part's
                // first writer applies beta (beta == 0: C is not read)
                const bool c_in = m0 + BM <= a.min_cm && n0 + BN <= a.min_cn;
                const int er0 = wm + 4 * (lane >> 4), ec0 = wn + (lane & 15);
                for (int e = a.wptr[r]; e < a.wptr[r + 1]; e++) {
                    const fmm_entry_t we = a.ent[e];
                    const double lam = a.alpha * we.coef;
                    const bool first = ((we.first >> a.gbit) & 1) != 0;   // first writer within the group
                    double* ct = Cb + we.off + (long)(m0 + er0) * a.rsC + (long)(n0 + ec0) * a.csC;
                    const int erows = we.rows - m0 - er0, ecols = we.cols - n0 - ec0;
#pragma unroll
                    for (int j = 0; j < FJ; j++)
#pragma unroll
                        for (int i = 0; i < FI; i++)
#pragma unroll
                            for (int q = 0; q < 4; q++) {
                                const int dr = 16 * i + q, dc = 16 * j;
                                if (c_in || (dr < erows && dc < ecols)) {
                                    double* p = ct + (long)dr * a.rsC + (long)dc * a.csC;
                                    double old = 0.0;
                                    if (!first) old = *p;
                                    else if (beta_c != 0.0) old = beta_c * *p;
                                    *p = fma(lam, acc[i][j][q], old);
                                }
                            }
                }
#pragma unroll
                for (int i = 0; i < FI; i++)
#pragma unroll
                    for (int j = 0; j < FJ; j++)
#pragma unroll
                        for (int q = 0; q < 4; q++) acc[i][j][q] = 0.0;
            }
        stage_done(more && kbn % STEPS_PER_CHUNK == 0, chunk & 1);
    };
    for (int r = r_begin; r < r_end; r++) {
        for (int kb = 0; kb < a.nkb; kb += 2) {
            step(r, kb, IntC<0>());
            step(r, kb + 1, IntC<1>());
        }
    }
}

}  // namespace bf3

// ---- host ---------------------------------------------------------------------

namespace {

constexpr int kWM = 2, kWN = 4, kFI = 4, kFJ = 2;
typedef bf3::GCfg<kWM, kWN, kFI, kFJ, 3> Cfg3;   // tile geometry (independent of the plane count)

long round_up(long x, long m) { return (x + m - 1) / m * m; }

}  // namespace

int bf16x3_gemm(fmmhip_alg_t alg, int terms, long m, long n, long k, double alpha, const double* A, long rsA, long csA,
                const double* B, long rsB, long csB, double beta, double* C, long rsC, long csC, hipStream_t st)
{
    if (!alg || m < 0 || n < 0 || k < 0 || (terms != 6 && terms != 8 && terms != 10)) return FMMHIP_ERR_INVALID;
    const int planes = terms == 10 ? 4 : 3;
    if (m == 0 || n == 0) return FMMHIP_SUCCESS;
    if (!C) return FMMHIP_ERR_INVALID;
    if (k == 0 || alpha == 0.0) return fmmhip_scale_d(C, m, n, rsC, csC, beta, st);
    if (!A || !B) return FMMHIP_ERR_INVALID;
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;

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
    const long LpA = round_up(g.lm, bf3::ROW_PAD), LpB = round_up(g.ln, bf3::ROW_PAD);
    const long ldk = round_up(g.lk, bf3::CHUNK);
    const long nch = ldk / bf3::CHUNK;
    if (LpA > INT32_MAX || LpB > INT32_MAX || ldk > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;
    const size_t planesA = (size_t)R * planes * LpA * ldk * 2, planesB = (size_t)R * planes * LpB * ldk * 2;
    const size_t expA = (size_t)R * nch * LpA * 4, expB = (size_t)R * nch * LpB * 4;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t oPB = al(planesA), oEA = oPB + al(planesB), oEB = oEA + al(expA);
    const long tiles = (LpA / Cfg3::BM) * (LpB / Cfg3::BN);
    // products materialised (pack 2, or automatic when the slots fit in half
    // of the device's memory): the kernel stores each product once and the
    // W stage applies W -- the kernel's W epilogue then only stores (same C
    // bit for bit either way)
    bool wmat = R > 1 && (alg->pack == 2 || alg->pack < 0);
    if (wmat && alg->pack < 0) wmat = 8.0 * R * (double)g.lm * (double)g.ln < fmmhip_ws_bound();
    MatView mv;
    memset(&mv, 0, sizeof(mv));
    if (wmat && (rc = mat_view(alg, g, st, &mv))) return rc;
    int cus = 256;
    fmmhip_device_cu_count(&cus);
    // product split (bf16 MFMA rate: 8 or 6 plane products per leaf product;
    // one workgroup per CU, LDS); materialised, any count up to R, no copies
    const double c_span = (double)(m - 1) * rsC + (double)(n - 1) * csC + 1.0;   // the temporaries' extent
    const int groups = wmat ? (alg->split > 0 ? std::min(alg->split, R) : mat_groups(R, tiles, cus))
                            : split_groups(alg->split, R, tiles,
                                           2.0 * terms * R * (double)g.lm * (double)g.lk * (double)g.ln, 8.0 * c_span,
                                           false, 2.5e15, 1);
    const long tmp_elems = (!wmat && groups > 1) ? (((m - 1) * rsC + (n - 1) * csC + 1 + 31) & ~31L) : 0;
    const size_t oT = oEB + al(expB);        // split temporaries, or the product slots
    const size_t bytes = oT + (wmat ? (size_t)R * (size_t)mv.slot * 8 : (size_t)groups * tmp_elems * 8);
    char* ws = (char*)workspace_for(st, bytes);
    if (!ws) return FMMHIP_ERR_NOMEM;

    bf3::SplitArgs sa;
    sa.X = A; sa.sr = rsA; sa.sk = csA; sa.ptr = dptr; sa.ent = dent; sa.swap = 0;
    sa.Lp = (int)LpA; sa.ldk = (int)ldk; sa.P = (bf3::u16*)ws; sa.E = (int*)(ws + oEA);
    bf3::SplitArgs sb;
    sb.X = B; sb.sr = csB; sb.sk = rsB; sb.ptr = dptr + (R + 1); sb.ent = dent; sb.swap = 1;
    sb.Lp = (int)LpB; sb.ldk = (int)ldk; sb.P = (bf3::u16*)(ws + oPB); sb.E = (int*)(ws + oEB);

    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    const bool timed = timing_active();
    if (timed) timing_record(&e0, st);
    const dim3 blk(256);
    if ((double)round_up(LpA / bf3::SROWS * nch, 8) * R > INT32_MAX ||
        (double)round_up(LpB / bf3::SROWS * nch, 8) * R > INT32_MAX)
        return FMMHIP_ERR_UNSUPPORTED;
    sa.R = sb.R = R;
    sa.nrc = (int)(LpA / bf3::SROWS * nch);
    sb.nrc = (int)(LpB / bf3::SROWS * nch);
    // grid: (row block, chunk) pairs rounded up to 8 (one run per XCD) x products
    const dim3 gA((unsigned)(round_up(sa.nrc, 8) * R));
    const dim3 gB((unsigned)(round_up(sb.nrc, 8) * R));
    auto split = [&](const bf3::SplitArgs& sg, bool rowc, dim3 grid) {
        if (planes == 4) {
            if (rowc) hipLaunchKernelGGL((bf3::split_kernel<true, 4>), grid, blk, 0, st, sg);
            else hipLaunchKernelGGL((bf3::split_kernel<false, 4>), grid, blk, 0, st, sg);
        } else {
            if (rowc) hipLaunchKernelGGL((bf3::split_kernel<true, 3>), grid, blk, 0, st, sg);
            else hipLaunchKernelGGL((bf3::split_kernel<false, 3>), grid, blk, 0, st, sg);
        }
    };
    split(sa, rsA == 1, gA);
    split(sb, csB == 1, gB);
    if (hipGetLastError() != hipSuccess) return FMMHIP_ERR_HIP;
    if (!wmat && groups > 1 && (rc = fmmhip_split_zero(1, ws + oT, tmp_elems, groups, st))) return rc;
    if (timed) timing_record(&e1, st);

    bf3::GemmArgs ga;
    ga.A = (const bf3::u16*)ws; ga.EA = (const int*)(ws + oEA);
    ga.B = (const bf3::u16*)(ws + oPB); ga.EB = (const int*)(ws + oEB);
    ga.a_plane = LpA * ldk; ga.b_plane = LpB * ldk;
    ga.ea_prod = nch * LpA; ga.eb_prod = nch * LpB;
    ga.LpA = (int)LpA; ga.LpB = (int)LpB;
    ga.nkb = (int)(ldk / bf3::KSTEP);
    ga.R = R;
    ga.wptr = dptr + 2 * (R + 1);
    ga.ent = dent;
    ga.C = C; ga.rsC = rsC; ga.csC = csC;
    ga.alpha = alpha; ga.beta = beta;
    ga.tiles_m = (int)(LpA / Cfg3::BM);
    ga.tiles_n = (int)(LpB / Cfg3::BN);
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
    if (terms == 10)
        hipLaunchKernelGGL((bf3::bf16x3_gemm_kernel<kWM, kWN, kFI, kFJ, 4, 10>), dim3((unsigned)nwg), dim3(Cfg3::NT),
                           0, st, ga);
    else if (terms == 8)
        hipLaunchKernelGGL((bf3::bf16x3_gemm_kernel<kWM, kWN, kFI, kFJ, 3, 8>), dim3((unsigned)nwg), dim3(Cfg3::NT),
                           0, st, ga);
    else
        hipLaunchKernelGGL((bf3::bf16x3_gemm_kernel<kWM, kWN, kFI, kFJ, 3, 6>), dim3((unsigned)nwg), dim3(Cfg3::NT),
                           0, st, ga);
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
