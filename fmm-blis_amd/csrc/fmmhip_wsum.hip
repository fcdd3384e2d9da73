// fmmhip_wsum.hip -- the W stage of the products-materialised mode: the
// MFMA kernel stores every product M_r = S_r T_r once, as leaf r of a slot
// array, and this stage forms every C part from them,
//   C_c := beta C_c + sum_r (alpha W[c][r]) M_r      (r ascending),
// over the part's extent.  It is the scatter of
// plugin/ref_kernels/bli_gemm_fmm_ref.c:101-120 turned around: one pass per
// C part that reads the products it needs, instead of one read-modify-write
// of C per (product, part).  The FMA chain per element is the one the fused
// epilogue runs (beta C, or 0 when beta == 0, then one fma per product in
// ascending order, rounded to the C type after each product), so the two
// modes give the same C.
//
// Layout: element (c, s) of slot r lies at M + r*slot + c + s*ldm, c along
// C's contiguous dimension (rows when swap == 0, columns when swap == 1);
// consecutive threads walk c, so the slot reads and the C accesses coalesce.
#include <hip/hip_runtime.h>

#include "fmm_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kEpt = 2;          // elements per thread (kThreads apart)
constexpr int kBatch = 8;        // products whose loads are in flight together

template <typename T>
__device__ __forceinline__ double round_to(double v)
{
    if constexpr (sizeof(T) == 4) return (double)(float)v;
    return v;
}

// Block bb: part p, strided index s, chunk cb of kThreads*kEpt elements
// along c.  Grid-stride over the blocks.
template <typename T>
__global__ __launch_bounds__(kThreads) void wsum_kernel(fmmhip_wsum_t w, int cblocks, long nblocks)
{
    const T* __restrict__ M = (const T*)w.M;
    T* __restrict__ C = (T*)w.C;
    const int tid = threadIdx.x;
    for (long bb = blockIdx.x; bb < nblocks; bb += gridDim.x) {
        const long per_part = (long)w.Ls * cblocks;
        const int p = (int)(bb / per_part);
        const long rem = bb - (long)p * per_part;
        const int s = (int)(rem / cblocks);
        const int cb = (int)(rem - (long)s * cblocks);
        const fmm_entry_t pe = w.parts[p];
        const int ec = w.swap ? pe.cols : pe.rows;      // part extent along c
        const int es = w.swap ? pe.rows : pe.cols;      // ... along s
        if (s >= es || ec <= 0) continue;               // uniform over the block
        const int t0 = pe.first, nt = pe.part;          // the part's terms
        const long cs_c = w.swap ? w.csC : w.rsC;       // C stride along c
        const long cs_s = w.swap ? w.rsC : w.csC;       // ... along s

        int cc[kEpt];
        bool ok[kEpt];
        double v[kEpt];
#pragma unroll
        for (int u = 0; u < kEpt; u++) {
            const int c = cb * kThreads * kEpt + u * kThreads + tid;
            ok[u] = c < ec;
            cc[u] = ok[u] ? c : ec - 1;                 // clamped: loads stay inside the part
            v[u] = 0.0;
        }
        T* cp = C + pe.off + (long)s * cs_s;
        if (w.beta != 0.0) {
#pragma unroll
            for (int u = 0; u < kEpt; u++) v[u] = w.beta * (double)cp[(long)cc[u] * cs_c];
        }
        const T* ms = M + (long)s * w.ldm;
        for (int t = 0; t < nt; t += kBatch) {
            double x[kBatch][kEpt];
#pragma unroll
            for (int j = 0; j < kBatch; j++) {
                const fmm_entry_t te = w.terms[t0 + min(t + j, nt - 1)];
#pragma unroll
                for (int u = 0; u < kEpt; u++) x[j][u] = (double)ms[te.off + cc[u]];
            }
#pragma unroll
            for (int j = 0; j < kBatch; j++) {
                if (t + j < nt) {
                    const double lam = w.alpha * w.terms[t0 + t + j].coef;
#pragma unroll
                    for (int u = 0; u < kEpt; u++) v[u] = round_to<T>(fma(lam, x[j][u], v[u]));
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kEpt; u++)
            if (ok[u]) cp[(long)cc[u] * cs_c] = (T)v[u];
    }
}

template <typename T>
int launch(const fmmhip_wsum_t* w, hipStream_t st)
{
    if (w->np <= 0 || w->Lc <= 0 || w->Ls <= 0) return FMMHIP_SUCCESS;
    const int cblocks = (w->Lc + kThreads * kEpt - 1) / (kThreads * kEpt);
    const long nblocks = (long)w->np * w->Ls * cblocks;
    const long grid = nblocks < (1L << 22) ? nblocks : (1L << 22);
    hipLaunchKernelGGL(wsum_kernel<T>, dim3((unsigned)grid), dim3(kThreads), 0, st, *w, cblocks, nblocks);
    return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}

}  // namespace

extern "C" int fmmhip_wsum(int is_double, const fmmhip_wsum_t* w, void* stream)
{
    if (!w) return FMMHIP_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    return is_double ? launch<double>(w, st) : launch<float>(w, st);
}
