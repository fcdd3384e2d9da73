// fmmhip_presum.hip -- the summation stage of the pre-summed pack mode: for
// every product r, the U- (or V-) weighted sum of the operand's parts,
//   S_r = sum_{(i,p)} U[(i,p)][r] A_ip      (zero padded to the leaf),
// written once to a workspace as R stacked leaves.  The fused kernel then
// reads S_r / T_r with one part per product.  This is the work of
// plugin/ref_kernels/bli_packm_fmm_ref.c:89-147 done for whole leaves at
// HBM bandwidth instead of per micro-panel.
//
// Layout: the operand is viewed along its contiguous dimension c (extent Lc)
// and its strided dimension s (extent Ls): element (c, s) of a part lies at
// X[off + c + s*ld].  S_r is stored [Ls][ldo] (c innermost, ldo >= Lc even)
// at S + r*ldo*Ls, so the GEMM reads it with the operand's original
// orientation and 16-byte pairs.  Consecutive threads walk c: reads and
// writes coalesce.
//
// Two-level tables with many parts per operand (333, 444) form the sums in
// two stages (presum_stage_kernel): the inner level's sums of every outer
// block's parts, then the outer level's sums of those, each stage over at
// most 16 parts.
#include <hip/hip_runtime.h>

#include "fmm_internal.h"

namespace {

// Global-memory views of pointers: the batched kernels read each segment's
// pointers from memory, where the compiler cannot tell their address space
// and would emit flat loads and stores (which also count against lgkmcnt);
// loads and stores through these types are global ones.
template <typename T> struct G1 {
    typedef const __attribute__((address_space(1))) T c;
    typedef __attribute__((address_space(1))) T m;
};
__device__ __forceinline__ fmm_entry_t entry_at(const fmm_entry_t* base, int i)
{
    typename G1<fmm_entry_t>::c* p = (typename G1<fmm_entry_t>::c*)base + i;
    fmm_entry_t e;
    e.off = p->off;
    e.rows = p->rows;
    e.cols = p->cols;
    e.coef = p->coef;
    e.first = p->first;
    e.part = p->part;
    return e;
}

// Thread (c0..c0+1, s) of the leaf; loops over all products.
template <typename T>
__global__ __launch_bounds__(256) void presum_kernel(const T* __restrict__ X, long ld, const int* __restrict__ ptr,
                                                     const fmm_entry_t* __restrict__ ent, int R, int swap,
                                                     int Lc, int Ls, int ldo, T* __restrict__ S)
{
    const long pairs_c = (Lc + 1) / 2;
    const long total = pairs_c * Ls;
    const long plane = (long)ldo * Ls;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const int s = (int)(idx / pairs_c);
        const int c = 2 * (int)(idx - (long)s * pairs_c);
        for (int r = 0; r < R; r++) {
            double v0 = 0.0, v1 = 0.0;
            for (int e = ptr[r]; e < ptr[r + 1]; e++) {
                const fmm_entry_t en = ent[e];
                const int ec = swap ? en.cols : en.rows;   // extent along c
                const int es = swap ? en.rows : en.cols;   // extent along s
                if (s >= es) continue;
                const T* p = X + en.off + c + (long)s * ld;
                if (c < ec) v0 = fma(en.coef, (double)p[0], v0);
                if (c + 1 < ec) v1 = fma(en.coef, (double)p[1], v1);
            }
            T* q = S + r * plane + c + (long)s * ldo;
            q[0] = (T)v0;
            if (c + 1 < Lc) q[1] = (T)v1;
        }
    }
}

// Batched form: segment s = one operand of one problem; blocks
// [boff[s], boff[s+1]) process it, each block a contiguous range of pairs.
template <typename T>
__global__ __launch_bounds__(256) void presum_batch_kernel(const fmmhip_presum_seg_t* __restrict__ segs, int nseg,
                                                           const int* __restrict__ boff)
{
    // the block's segment: binary search over the block offsets (uniform)
    int lo = 0, hi = nseg - 1;
    const int b = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boff[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const fmmhip_presum_seg_t sg = segs[lo];
    typename G1<T>::c* __restrict__ X = (typename G1<T>::c*)(const T*)sg.X;
    typename G1<T>::m* __restrict__ S = (typename G1<T>::m*)(T*)sg.S;
    typename G1<int>::c* __restrict__ sptr = (typename G1<int>::c*)sg.ptr;
    const long pairs_c = (sg.Lc + 1) / 2;
    const long total = pairs_c * sg.Ls;
    const long plane = (long)sg.ldo * sg.Ls;
    const long per_block = (total + (boff[lo + 1] - boff[lo]) - 1) / (boff[lo + 1] - boff[lo]);
    const long i0 = (long)(b - boff[lo]) * per_block;
    const long i1 = i0 + per_block < total ? i0 + per_block : total;
    for (long idx = i0 + threadIdx.x; idx < i1; idx += blockDim.x) {
        const int s = (int)(idx / pairs_c);
        const int c = 2 * (int)(idx - (long)s * pairs_c);
        for (int r = 0; r < sg.R; r++) {
            double v0 = 0.0, v1 = 0.0;
            for (int e = sptr[r]; e < sptr[r + 1]; e++) {
                const fmm_entry_t en = entry_at(sg.ent, e);
                const int ec = sg.swap ? en.cols : en.rows;
                const int es = sg.swap ? en.rows : en.cols;
                if (s >= es) continue;
                typename G1<T>::c* p = X + en.off + c + (long)s * sg.ld;
                if (c < ec) v0 = fma(en.coef, (double)p[0], v0);
                if (c + 1 < ec) v1 = fma(en.coef, (double)p[1], v1);
            }
            typename G1<T>::m* q = S + r * plane + c + (long)s * sg.ldo;
            q[0] = (T)v0;
            if (c + 1 < sg.Lc) q[1] = (T)v1;
        }
    }
}

// Part-staged form: block = TB consecutive elements along c of one s;
// thread t loads element (c, s) of every part once into its own LDS slot
// xs[p][t] (read back only by itself: no barrier), then forms every S_r with
// the entries' coefficients in CSR order -- the same FMA sequence as
// presum_kernel, so the sums are bit-identical; HBM reads drop from one per
// (entry, element) to one per (part, element).
// The parts' offsets and extents of one operand, into LDS (once per block).
template <int NPMAX, int TB>
__device__ __forceinline__ void parts_meta(const fmm_entry_t* __restrict__ parts, int np, int swap, long* p_off,
                                           int* p_c, int* p_s)
{
    for (int p = threadIdx.x; p < np; p += TB) {
        const fmm_entry_t pe = entry_at(parts, p);
        p_off[p] = pe.off;
        p_c[p] = swap ? pe.cols : pe.rows;
        p_s[p] = swap ? pe.rows : pe.cols;
    }
}

// One unit: element (c, s) of every part into the thread's LDS slots, then
// every S_r at (c, s).  16 parts at a time, their loads in flight together:
// every address is clamped into its part (or to X[0] for an empty part), and
// elements outside the part are replaced by zero afterwards.
template <typename T, typename TO, int NPMAX, int TB>
__device__ __forceinline__ void parts_unit(const T* __restrict__ X, long ld, const int* __restrict__ ptr,
                                           const fmm_entry_t* __restrict__ ent, int np, int R, int Lc, int ldo,
                                           long plane, TO* __restrict__ S, int s, int c, double (*xs)[TB],
                                           const long* p_off, const int* p_c, const int* p_s)
{
    const int t = threadIdx.x;
    typename G1<T>::c* __restrict__ Xg = (typename G1<T>::c*)X;
    typename G1<TO>::m* __restrict__ Sg = (typename G1<TO>::m*)S;
    typename G1<int>::c* __restrict__ ptrg = (typename G1<int>::c*)ptr;
    for (int p0 = 0; p0 < np; p0 += 16) {
        double xv[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int p = min(p0 + j, np - 1);
            const int ec = p_c[p], es = p_s[p];
            const long at = ((ec > 0) & (es > 0)) ? p_off[p] + min(c, ec - 1) + (long)min(s, es - 1) * ld : 0;
            const double x = (double)Xg[at];
            xv[j] = ((c < ec) & (s < es)) ? x : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (p0 + j < np) xs[p0 + j][t] = xv[j];
    }
    if (c < Lc) {
        for (int r = 0; r < R; r++) {
            double v = 0.0;
            for (int e = ptrg[r]; e < ptrg[r + 1]; e++) {
                const fmm_entry_t en = entry_at(ent, e);
                v = fma(en.coef, xs[en.part][t], v);
            }
            Sg[r * plane + c + (long)s * ldo] = (TO)v;
        }
    }
}

// TO: the leaves' type (FP32 leaves of FP64 operands for the FP32-leaf kernel)
template <typename T, typename TO, int NPMAX, int TB>
__global__ __launch_bounds__(TB) void presum_parts_kernel(const T* __restrict__ X, long ld,
                                                          const int* __restrict__ ptr,
                                                          const fmm_entry_t* __restrict__ ent,
                                                          const fmm_entry_t* __restrict__ parts, int np, int R,
                                                          int swap, int Lc, int Ls, int ldo, TO* __restrict__ S)
{
    __shared__ double xs[NPMAX][TB];
    __shared__ long p_off[NPMAX];
    __shared__ int p_c[NPMAX], p_s[NPMAX];
    parts_meta<NPMAX, TB>(parts, np, swap, p_off, p_c, p_s);
    __syncthreads();
    const long nbc = (Lc + TB - 1) / TB;
    const long plane = (long)ldo * Ls;
    for (long blk = blockIdx.x; blk < nbc * Ls; blk += gridDim.x) {
        const int s = (int)(blk / nbc);
        const int c = (int)(blk - (long)s * nbc) * TB + threadIdx.x;
        parts_unit<T, TO, NPMAX, TB>(X, ld, ptr, ent, np, R, Lc, ldo, plane, S, s, c, xs, p_off, p_c, p_s);
    }
}

// Batched part-staged form (segments with at most NPMAX parts): blocks
// [boff[g], boff[g+1]) take the units (TB elements along c of one s) of
// segment g in a grid stride.
template <typename T, int NPMAX, int TB>
__global__ __launch_bounds__(TB) void presum_parts_batch_kernel(const fmmhip_presum_seg_t* __restrict__ segs,
                                                                int nseg, const int* __restrict__ boff)
{
    int lo = 0, hi = nseg - 1;
    const int b = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boff[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const fmmhip_presum_seg_t sg = segs[lo];
    __shared__ double xs[NPMAX][TB];
    __shared__ long p_off[NPMAX];
    __shared__ int p_c[NPMAX], p_s[NPMAX];
    parts_meta<NPMAX, TB>(sg.parts, sg.np, sg.swap, p_off, p_c, p_s);
    __syncthreads();
    const long nbc = (sg.Lc + TB - 1) / TB;
    const long plane = (long)sg.ldo * sg.Ls;
    const int nbl = boff[lo + 1] - boff[lo];
    for (long u = b - boff[lo]; u < nbc * sg.Ls; u += nbl) {
        const int s = (int)(u / nbc);
        const int c = (int)(u - (long)s * nbc) * TB + threadIdx.x;
        parts_unit<T, T, NPMAX, TB>((const T*)sg.X, sg.ld, sg.ptr, sg.ent, sg.np, sg.R, sg.Lc, sg.ldo, plane, (T*)sg.S, s,
                                 c, xs, p_off, p_c, p_s);
    }
}

// One stage of the two-stage summation of a two-level algorithm: segment
// si = blockIdx.x / bps forms its R sums of up to 16 parts over the whole
// leaf (units of 2*TB elements along c of one s, grid-strided over the
// segment's bps blocks).  Each thread holds two adjacent elements of every
// part in registers; the sums walk the segment's sparse term list (part,
// coefficient; ascending part per sum, uniform scalar loads), the part
// register chosen by a uniform switch, so only nonzero terms cost VALU
// work and no load depends on a term.
template <typename T, int TB>
__global__ __launch_bounds__(TB) void presum_stage_kernel(const T* __restrict__ X, long ld, T* __restrict__ S,
                                                          int swap, int Lc, int Ls, int ldo,
                                                          const fmmhip_stage_seg_t* __restrict__ segs, int bps)
{
    constexpr int NP = FMMHIP_STAGE_PARTS;
    const int si = (int)blockIdx.x / bps;
    const fmmhip_stage_seg_t sg = segs[si];
    // The segment's part descriptors and term list into LDS, read back
    // below as uniform (scalar) values: read from global memory in the
    // loops, each would be a vector load whose wait also waits for every
    // store issued before it.
    __shared__ long p_off_l[NP];
    __shared__ int p_c_l[NP], p_s_l[NP];
    __shared__ int tp_l[FMMHIP_STAGE_MAX_R + 1];
    __shared__ fmmhip_stage_term_t tm_l[FMMHIP_STAGE_PARTS * FMMHIP_STAGE_MAX_R];
    parts_meta<NP, TB>(sg.parts, sg.np, swap, p_off_l, p_c_l, p_s_l);
    const int nterm = sg.tptr[sg.R];
    for (int i = threadIdx.x; i <= sg.R; i += TB) tp_l[i] = sg.tptr[i];
    for (int i = threadIdx.x; i < nterm; i += TB) tm_l[i] = sg.terms[i];
    __syncthreads();
    const long nbc = (Lc + 2 * TB - 1) / (2 * TB);
    T* __restrict__ So = S + sg.s_off;
    for (long u = (long)blockIdx.x - (long)si * bps; u < nbc * Ls; u += bps) {
        const int s = (int)(u / nbc);
        const int c0 = (int)(u - (long)s * nbc) * 2 * TB + 2 * (int)threadIdx.x;
        // s is uniform too, so each part's row start is a scalar base
        // address and only the offset along c is per lane; every address is
        // clamped into its part (or to X[0] for an empty part), all 32 loads
        // are in flight together, and elements outside the part are replaced
        // by zero afterwards.
        double r0[NP], r1[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const int p = min(j, sg.np - 1);
            const int ec = __builtin_amdgcn_readfirstlane(p_c_l[p]), es = __builtin_amdgcn_readfirstlane(p_s_l[p]);
            const long off = (long)(((unsigned long)(unsigned)__builtin_amdgcn_readfirstlane((int)(p_off_l[p] >> 32)) << 32) |
                                    (unsigned)__builtin_amdgcn_readfirstlane((int)p_off_l[p]));
            const bool live = (ec > 0) & (es > 0);
            const T* __restrict__ base = X + (live ? off + (long)min(s, es - 1) * ld : 0);
            r0[j] = (double)base[live ? min(c0, ec - 1) : 0];
            r1[j] = (double)base[live ? min(c0 + 1, ec - 1) : 0];
        }
        double x0[NP], x1[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const int p = min(j, sg.np - 1);
            const int ec = __builtin_amdgcn_readfirstlane(p_c_l[p]), es = __builtin_amdgcn_readfirstlane(p_s_l[p]);
            const bool in = (s < es) & (j < sg.np);
            x0[j] = (in & (c0 < ec)) ? r0[j] : 0.0;
            x1[j] = (in & (c0 + 1 < ec)) ? r1[j] : 0.0;
        }
        if (c0 < Lc) {
            T* out = So + c0 + (long)s * ldo;
            const bool two = c0 + 1 < Lc;
            for (int r = 0; r < sg.R; r++) {
                double v0 = 0.0, v1 = 0.0;
                const int t1 = __builtin_amdgcn_readfirstlane(tp_l[r + 1]);
                for (int t = __builtin_amdgcn_readfirstlane(tp_l[r]); t < t1; t++) {
                    const fmmhip_stage_term_t tm = tm_l[t];
                    switch (__builtin_amdgcn_readfirstlane(tm.part)) {
#define STAGE_TERM(K)                                                                                         \
    case K:                                                                                                   \
        v0 = fma(tm.coef, x0[K], v0);                                                                         \
        v1 = fma(tm.coef, x1[K], v1);                                                                         \
        break;
                        STAGE_TERM(0) STAGE_TERM(1) STAGE_TERM(2) STAGE_TERM(3)
                        STAGE_TERM(4) STAGE_TERM(5) STAGE_TERM(6) STAGE_TERM(7)
                        STAGE_TERM(8) STAGE_TERM(9) STAGE_TERM(10) STAGE_TERM(11)
                        STAGE_TERM(12) STAGE_TERM(13) STAGE_TERM(14) STAGE_TERM(15)
#undef STAGE_TERM
                    default: break;
                    }
                }
                // one store of the pair: S, s_off, pstride and ldo are even
                // element counts from a 256-byte aligned base, and c0 is even
                if (two) {
                    typedef T pair2 __attribute__((ext_vector_type(2)));
                    *(pair2*)(out + (long)r * sg.pstride) = pair2{(T)v0, (T)v1};
                } else {
                    out[(long)r * sg.pstride] = (T)v0;
                }
            }
        }
    }
}

template <typename T, typename TO, int NPMAX, int TB>
int launch_parts(const fmmhip_presum_seg_t* g, hipStream_t st)
{
    const long nbc = (g->Lc + TB - 1) / TB;
    long blocks = nbc * g->Ls;
    if (blocks <= 0) return FMMHIP_SUCCESS;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((presum_parts_kernel<T, TO, NPMAX, TB>), dim3((unsigned)blocks), dim3(TB), 0, st,
                       (const T*)g->X, g->ld, g->ptr, g->ent, g->parts, g->np, g->R, g->swap, g->Lc, g->Ls, g->ldo,
                       (TO*)g->S);
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}

template <typename T, typename TO = T>
int launch_parts_typed(const fmmhip_presum_seg_t* g, hipStream_t st)
{
    // LDS per block = NPMAX * TB * 8 bytes: 32 KB, 64 KB, 40 KB, 64 KB,
    // 128 KB (36 parts, 2-level <2,3,2> and <3,2,2>: the 40-part form holds
    // four blocks per CU where the 64-part one held two)
    if (g->np <= 16) return launch_parts<T, TO, 16, 256>(g, st);
    if (g->np <= 32) return launch_parts<T, TO, 32, 256>(g, st);
    if (g->np <= 40) return launch_parts<T, TO, 40, 128>(g, st);
    if (g->np <= 64) return launch_parts<T, TO, 64, 128>(g, st);
    if (g->np <= 256) return launch_parts<T, TO, 256, 64>(g, st);
    return FMMHIP_ERR_UNSUPPORTED;
}

}  // namespace

extern "C" int fmmhip_presum_seg(int is_double, const fmmhip_presum_seg_t* seg, void* stream)
{
    if (!seg || !seg->parts || seg->np <= 0) return FMMHIP_ERR_INVALID;
    return is_double ? launch_parts_typed<double>(seg, (hipStream_t)stream)
                     : launch_parts_typed<float>(seg, (hipStream_t)stream);
}

// FP64 operand, FP32 leaves: each FP64 sum rounded once to FP32 (the value
// the FP32-leaf kernel would stage from an FP64 leaf)
extern "C" int fmmhip_presum_seg_d2s(const fmmhip_presum_seg_t* seg, void* stream)
{
    if (!seg || !seg->parts || seg->np <= 0) return FMMHIP_ERR_INVALID;
    return launch_parts_typed<double, float>(seg, (hipStream_t)stream);
}

extern "C" int fmmhip_presum_stage(int is_double, const void* X, long ld, void* S, int swap, int Lc, int Ls, int ldo,
                                   const fmmhip_stage_seg_t* segs, int nseg, void* stream)
{
    if (nseg <= 0 || Lc <= 0 || Ls <= 0) return FMMHIP_SUCCESS;
    // the kernel stores element pairs as one vector (the segments' offsets
    // and strides are even element counts by construction)
    if ((ldo & 1) || ((uintptr_t)S & (2 * (is_double ? sizeof(double) : sizeof(float)) - 1))) return FMMHIP_ERR_INVALID;
    constexpr int TB = 256;
    // about 65536 blocks in all (grid-strided), at most one per unit of a segment
    const long units = (long)((Lc + 2 * TB - 1) / (2 * TB)) * Ls;
    const long want = (65536 + nseg - 1) / nseg;
    const int bps = (int)(units < want ? units : want);
    if ((long)bps * nseg > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;
    if (is_double)
        hipLaunchKernelGGL((presum_stage_kernel<double, TB>), dim3((unsigned)(bps * nseg)),
                           dim3(TB), 0, (hipStream_t)stream, (const double*)X, ld, (double*)S, swap, Lc, Ls, ldo,
                           segs, bps);
    else
        hipLaunchKernelGGL((presum_stage_kernel<float, TB>), dim3((unsigned)(bps * nseg)),
                           dim3(TB), 0, (hipStream_t)stream, (const float*)X, ld, (float*)S, swap, Lc, Ls, ldo, segs,
                           bps);
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}

extern "C" int fmmhip_presum_batch(int is_double, const fmmhip_presum_seg_t* segs, int nseg, const int* boff,
                                   int nblocks, int parts, void* stream)
{
    if (nseg <= 0 || nblocks <= 0) return FMMHIP_SUCCESS;
    if (parts) {
        if (is_double)
            hipLaunchKernelGGL((presum_parts_batch_kernel<double, FMMHIP_BATCH_PARTS_MAX, 256>), dim3(nblocks), dim3(256), 0,
                               (hipStream_t)stream, segs, nseg, boff);
        else
            hipLaunchKernelGGL((presum_parts_batch_kernel<float, FMMHIP_BATCH_PARTS_MAX, 256>), dim3(nblocks), dim3(256), 0,
                               (hipStream_t)stream, segs, nseg, boff);
        return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
    }
    if (is_double)
        hipLaunchKernelGGL(presum_batch_kernel<double>, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, segs, nseg,
                           boff);
    else
        hipLaunchKernelGGL(presum_batch_kernel<float>, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, segs, nseg,
                           boff);
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}

// ptr: CSR of the operand's entries over products (R+1 values, indexing ent).
// swap: the entries' (rows, cols) are (s, c) rather than (c, s).
extern "C" int fmmhip_presum_d(const double* X, long ld, const int* ptr, const fmm_entry_t* ent, int R, int swap,
                               int Lc, int Ls, int ldo, double* S, void* stream)
{
    const long total = (long)((Lc + 1) / 2) * Ls;
    if (total <= 0) return FMMHIP_SUCCESS;
    long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(presum_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, X, ld,
                       ptr, ent, R, swap, Lc, Ls, ldo, S);
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}

extern "C" int fmmhip_presum_s(const float* X, long ld, const int* ptr, const fmm_entry_t* ent, int R, int swap,
                               int Lc, int Ls, int ldo, float* S, void* stream)
{
    const long total = (long)((Lc + 1) / 2) * Ls;
    if (total <= 0) return FMMHIP_SUCCESS;
    long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(presum_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, X, ld,
                       ptr, ent, R, swap, Lc, Ls, ldo, S);
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}
