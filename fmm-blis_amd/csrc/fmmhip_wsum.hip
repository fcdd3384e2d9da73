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
//
// Two passes: per part (wsum_kernel: each part reads its products, so a
// product is read once per W entry), or all parts of a group at once
// (wsum_cols_kernel, groups of at most 16 parts over at most 128 slots: a
// thread keeps every part's value at its element in registers and reads
// each slot once, applying it to the parts that take it; 222 L2, one
// group: 49 slot reads and 32 C accesses per element instead of 144 and
// 32; the two stages of 444 L2: 49 groups of 16 Y leaves over 49 products,
// then 16 groups of 16 C parts over 49 Y leaves).  Both run the same FMA
// chain per element, so they give the same C.
#include <hip/hip_runtime.h>

#include <algorithm>

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

constexpr int kColsMaxR = 128;   // products of the all-parts pass
constexpr int kColsMaxP = 16;    // parts of the all-parts pass

// All parts of one group at once.  blockIdx.y: the group; block item bb:
// strided index s, chunk cb of kThreads*kEpt elements along c; grid-stride
// over the items.  The block first lays out its group's W table by slot in
// LDS: lam[j][p] = alpha W (part p, slot j), bit p of mask[j] set where
// part p takes slot j, and the slot offsets.
template <typename T>
__global__ __launch_bounds__(kThreads) void wsum_cols_kernel(fmmhip_wsum_t w, int cblocks, long nblocks)
{
    __shared__ double lam_s[kColsMaxR][kColsMaxP];
    __shared__ long off_s[kColsMaxR];
    __shared__ unsigned char has_s[kColsMaxR][kColsMaxP];
    __shared__ unsigned mask_s[kColsMaxR];
    __shared__ int part_s[kColsMaxP];
    const T* __restrict__ M = (const T*)w.M;
    T* __restrict__ C = (T*)w.C;
    const int tid = threadIdx.x;
    const int R = w.nslots;
    const int g = (int)blockIdx.y;
    const int np = w.groups ? w.gparts : w.np;      // parts of this group (<= kColsMaxP, host-checked)
    if (tid < kColsMaxP) {
        int pi = -1;
        if (tid < np) pi = w.groups ? w.groups[(long)g * w.gparts + tid] : tid;
        part_s[tid] = (pi >= 0 && pi < w.np) ? pi : -1;
    }
    for (int i = tid; i < kColsMaxR * kColsMaxP; i += kThreads) (&has_s[0][0])[i] = 0;
    for (int i = tid; i < kColsMaxR; i += kThreads) off_s[i] = 0;
    __syncthreads();
    for (int p = 0; p < np; p++) {
        const int pi = part_s[p];
        if (pi < 0) continue;                // uniform
        const fmm_entry_t pe = w.parts[pi];
        for (int j = tid; j < pe.part; j += kThreads) {
            const fmm_entry_t te = w.terms[pe.first + j];
            const int r = te.cols;               // the slot's index in the group
            if (r >= 0 && r < R) {
                lam_s[r][p] = w.alpha * te.coef;
                has_s[r][p] = 1;
                off_s[r] = te.off;
            }
        }
        __syncthreads();                    // one part's writes per round: no two threads on one entry
    }
    if (tid < kColsMaxR) {
        unsigned m = 0;
        for (int p = 0; p < np; p++) m |= has_s[tid][p] ? 1u << p : 0u;
        mask_s[tid] = m;
    }
    __syncthreads();
    const long cs_c = w.swap ? w.csC : w.rsC;       // C stride along c
    const long cs_s = w.swap ? w.rsC : w.csC;       // ... along s
    for (long bb = blockIdx.x; bb < nblocks; bb += gridDim.x) {
        const int s = (int)(bb / cblocks);
        const int cb = (int)(bb - (long)s * cblocks);
        int c[kEpt], cc[kEpt];
#pragma unroll
        for (int u = 0; u < kEpt; u++) {
            c[u] = cb * kThreads * kEpt + u * kThreads + tid;
            cc[u] = c[u] < w.Lc ? c[u] : w.Lc - 1;  // clamped: slot loads stay inside the leaf
        }
        double acc[kColsMaxP][kEpt];
#pragma unroll
        for (int p = 0; p < kColsMaxP; p++) {
#pragma unroll
            for (int u = 0; u < kEpt; u++) acc[p][u] = 0.0;
            if (p < np && part_s[p] >= 0 && w.beta != 0.0) {
                const fmm_entry_t pe = w.parts[part_s[p]];
                const int ec = w.swap ? pe.cols : pe.rows, es = w.swap ? pe.rows : pe.cols;
                if (s < es) {
                    const T* cp = C + pe.off + (long)s * cs_s;
#pragma unroll
                    for (int u = 0; u < kEpt; u++)
                        if (c[u] < ec) acc[p][u] = w.beta * (double)cp[(long)c[u] * cs_c];
                }
            }
        }
        const T* ms = M + (long)s * w.ldm;
        for (int r0 = 0; r0 < R; r0 += kBatch) {
            double x[kBatch][kEpt];
#pragma unroll
            for (int j = 0; j < kBatch; j++) {
                const long o = off_s[min(r0 + j, R - 1)];
#pragma unroll
                for (int u = 0; u < kEpt; u++) x[j][u] = (double)ms[o + cc[u]];
            }
#pragma unroll
            for (int j = 0; j < kBatch; j++) {
                if (r0 + j < R) {
                    const unsigned m = __builtin_amdgcn_readfirstlane(mask_s[r0 + j]);
#pragma unroll
                    for (int p = 0; p < kColsMaxP; p++) {
                        if (m & (1u << p)) {
                            const double lam = lam_s[r0 + j][p];
#pragma unroll
                            for (int u = 0; u < kEpt; u++) acc[p][u] = round_to<T>(fma(lam, x[j][u], acc[p][u]));
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int p = 0; p < kColsMaxP; p++) {
            if (p < np && part_s[p] >= 0) {
                const fmm_entry_t pe = w.parts[part_s[p]];
                const int ec = w.swap ? pe.cols : pe.rows, es = w.swap ? pe.rows : pe.cols;
                if (s < es) {
                    T* cp = C + pe.off + (long)s * cs_s;
#pragma unroll
                    for (int u = 0; u < kEpt; u++)
                        if (c[u] < ec) cp[(long)c[u] * cs_c] = (T)acc[p][u];
                }
            }
        }
    }
}

template <typename T>
int launch(const fmmhip_wsum_t* w, hipStream_t st)
{
    if (w->np <= 0 || w->Lc <= 0 || w->Ls <= 0) return FMMHIP_SUCCESS;
    const int cblocks = (w->Lc + kThreads * kEpt - 1) / (kThreads * kEpt);
    const bool grouped = w->groups && w->ngroups > 0 && w->gparts > 0 && w->gparts <= kColsMaxP;
    const bool single = !w->groups && w->np <= kColsMaxP;
    if (w->nslots > 0 && w->nslots <= kColsMaxR && (grouped || single)) {
        const long nblocks = (long)w->Ls * cblocks;
        const int ng = grouped ? w->ngroups : 1;
        int cus = 256;
        fmmhip_device_cu_count(&cus);
        const long per = std::max(1L, std::min(nblocks, 8L * cus / ng));   // blocks per group
        hipLaunchKernelGGL(wsum_cols_kernel<T>, dim3((unsigned)per, (unsigned)ng), dim3(kThreads), 0, st, *w,
                           cblocks, nblocks);
        return hipGetLastError() == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
    }
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
