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
#include <hip/hip_runtime.h>

#include "fmm_internal.h"

namespace {

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
    const T* __restrict__ X = (const T*)sg.X;
    T* __restrict__ S = (T*)sg.S;
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
            for (int e = sg.ptr[r]; e < sg.ptr[r + 1]; e++) {
                const fmm_entry_t en = sg.ent[e];
                const int ec = sg.swap ? en.cols : en.rows;
                const int es = sg.swap ? en.rows : en.cols;
                if (s >= es) continue;
                const T* p = X + en.off + c + (long)s * sg.ld;
                if (c < ec) v0 = fma(en.coef, (double)p[0], v0);
                if (c + 1 < ec) v1 = fma(en.coef, (double)p[1], v1);
            }
            T* q = S + r * plane + c + (long)s * sg.ldo;
            q[0] = (T)v0;
            if (c + 1 < sg.Lc) q[1] = (T)v1;
        }
    }
}

}  // namespace

extern "C" int fmmhip_presum_batch(int is_double, const fmmhip_presum_seg_t* segs, int nseg, const int* boff,
                                   int nblocks, void* stream)
{
    if (nseg <= 0 || nblocks <= 0) return FMMHIP_SUCCESS;
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
