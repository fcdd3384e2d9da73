/*
 * fmm_internal.h -- declarations shared by the host C files (algorithm
 * tables, geometry, the obj_t driver) and the HIP layer (kernels and their
 * launchers).  Not installed; include/fmm_blis_amd.h is the public ABI.
 */
#ifndef FMM_INTERNAL_H
#define FMM_INTERNAL_H

#include <stdint.h>
#include "../../include/fmm_blis_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One nonzero coefficient of a product, resolved against a problem's
 * geometry: the part starts at element offset `off` of its operand, has
 * `rows` x `cols` valid elements (the rest of the leaf is zero padding),
 * and enters with coefficient `coef`.  For C entries, `first` marks the
 * first product that writes the part (where beta is applied). */
typedef struct fmm_entry_s {
    int64_t off;
    int32_t rows;
    int32_t cols;
    double  coef;
    int32_t first;
    int32_t part;
} fmm_entry_t;

/* Composed algorithm (levels already multiplied out) plus, per product r,
 * the compressed lists of its nonzero U, V, W coefficients. */
struct fmmhip_alg_s {
    int  mt, kt, nt, R, levels;
    int  classical;             /* R == 1, 1x1x1 */
    int *U, *V, *W;             /* dense [part][r] */
    int *u_ptr, *u_part, *u_coef;   /* CSR over r: entries u_ptr[r]..u_ptr[r+1]-1 */
    int *v_ptr, *v_part, *v_coef;
    int *w_ptr, *w_part, *w_coef, *w_first;
    long nnz_u, nnz_v, nnz_w;
    int  max_nnz_u, max_nnz_v, max_nnz_w;
    int  shape;                 /* kernel launch shape, -1 = chosen per call */
    int  pack;                  /* 0 fused, 1 pre-summed workspace, -1 = chosen per call */
    int  last_pack;             /* mode the last FP64/FP32 call took (-1: none yet) */
    int  last_shape;            /* kernel shape the last FP64/FP32 call launched (-1: none yet) */
    int  keep_all;              /* never evict cached descriptors (owners that capture graphs) */
    int  split;                 /* product groups per tile: -1 chosen per call, 1 (none), 2, 4, 8 */
    /* two-level algorithms: the two levels' own tables (level 0 outer),
     * for the two-stage operand summation; NULL otherwise */
    int  lv_m[2], lv_k[2], lv_n[2], lv_R[2];
    int *lv_U[2], *lv_V[2], *lv_W[2];   /* dense [part][r] */
    void *dev_cache;            /* owned by the HIP layer (fmmhip_alg_release_cache) */
};

/* Problem geometry of one call (host side). */
typedef struct fmm_geom_s {
    long m, n, k;               /* user dims of C (m x n) and the inner k */
    long lm, lk, ln;            /* leaf dims: ceil(m/mt), ceil(k/kt), ceil(n/nt) */
    long rsA, csA, rsB, csB, rsC, csC;
} fmm_geom_t;

/* Part split of `whole` into `tilde` contiguous parts, sizes
 * floor(whole/tilde) + (t < whole mod tilde), offsets the prefix sums
 * (frame/base/bli_strassen_ab.c:86-121). */
void fmm_part_split(long whole, int tilde, long* size, long* off);

/* Resolves the algorithm's coefficient lists against geometry g into
 * host arrays (caller frees with free()):  ptr[3*(R+1)] = u, v, w CSR
 * pointers; ent = u entries, then v, then w.  Returns 0 or an fmmhip
 * status. */
int fmm_build_entries(const struct fmmhip_alg_s* alg, const fmm_geom_t* g,
                      int32_t** ptr_out, fmm_entry_t** ent_out, long* nent_out);
/* The same lists into caller memory: ptr[3*(R+1)], ent[nnz_u+nnz_v+nnz_w]
 * (zeroed by the caller), and 2*(mt+kt+nt) longs of `sizes` to work in. */
void fmm_build_entries_into(const struct fmmhip_alg_s* alg, const fmm_geom_t* g, int32_t* ptr, fmm_entry_t* ent,
                            long* sizes);

/* HIP layer hooks used by the C files. */
void fmmhip_alg_release_cache(struct fmmhip_alg_s* alg);
/* Scales C by beta (beta == 0 writes zeros without reading C). */
int fmmhip_scale_d(double* C, long m, long n, long rsC, long csC, double beta, void* stream);
int fmmhip_scale_s(float* C, long m, long n, long rsC, long csC, float beta, void* stream);
/* Pointer kind: 1 device (or managed), 0 host, <0 error. */
int fmmhip_pointer_is_device(const void* p);
/* Set-up of the no-allocation entry point's scratch on the current device
 * (fmmhip_noalloc.hip); called by the library's init. */
int fmmhip_noalloc_setup(void);
/* The library's default stream (created on first use). */
void* fmmhip_default_stream(void);
/* Largest workspace (bytes) an automatic pack-mode choice may take: half of
 * the device's total memory, a property of the device rather than of what
 * is free at the time of a call. */
double fmmhip_ws_bound(void);
/* Y := X (m x n, element (i,j) at X[i*rs + j*cs]) into column-major Y
 * (ld = m), asynchronous on `stream`. */
int fmmhip_gather(int is_double, const void* X, long m, long n, long rs, long cs, void* Y, void* stream);
/* Split-product epilogue helpers (fmmhip_runtime.hip): zero G temporaries,
 * then C := beta*C + sum_g T_g over an m x n matrix with strides (rs, cs);
 * each temporary spans `elems` elements with C's strides. */
int fmmhip_split_zero(int is_double, void* T, long elems, int groups, void* stream);
int fmmhip_split_reduce(int is_double, void* C, long m, long n, long rs, long cs, double beta, const void* T,
                        long elems, int groups, void* stream);

/* One operand of one problem for the batched summation stage. */
typedef struct fmmhip_presum_seg_s {
    const void* X;              /* operand */
    void* S;                    /* its workspace leaves */
    long ld;                    /* operand leading dimension */
    const int* ptr;             /* device CSR over products (R+1) */
    const fmm_entry_t* ent;     /* device entries */
    int R, swap, Lc, Ls, ldo;
    int np;                     /* parts of the operand (entries' `part` < np) */
    const fmm_entry_t* parts;   /* device table of the np parts (offset and extents), or NULL */
} fmmhip_presum_seg_t;

/* Batched summation: segment s is processed by blocks [boff[s], boff[s+1]),
 * boff has nseg+1 entries, nblocks = boff[nseg] (all device memory).
 * parts: every segment has its part table with np <= FMMHIP_BATCH_PARTS_MAX
 * (part-staged form, same sums bit for bit); else the entry-driven form. */
#define FMMHIP_BATCH_PARTS_MAX 16
int fmmhip_presum_batch(int is_double, const fmmhip_presum_seg_t* segs, int nseg, const int* boff, int nblocks,
                        int parts, void* stream);

/* W stage of the products-materialised pack mode (fmmhip_wsum.hip): for
 * every C part c, C_c := beta C_c + sum_r (alpha W[c][r]) M_r over the
 * part's extent, r ascending, from the slots M_r the MFMA kernel stored. */
typedef struct fmmhip_wsum_s {
    const void* M;              /* R slots, slot r at M + off of its term */
    long ldm;                   /* slot leading dimension (along s) */
    int swap;                   /* 0: c runs along C's rows (column-major C), 1: along its columns */
    int Lc, Ls;                 /* leaf extents along c and s */
    void* C;
    long rsC, csC;
    double alpha, beta;
    int np;                     /* C parts */
    /* device part table: off, rows, cols of C part p; `first` = index of its
     * first term, `part` = number of its terms */
    const fmm_entry_t* parts;
    /* device terms, per part in ascending product order: off = element
     * offset of the product's slot, coef = W[c][r], part = r, cols = the
     * slot's index inside the part's group (below) */
    const fmm_entry_t* terms;
    /* All-parts pass (nslots > 0): the parts fall into ngroups groups of at
     * most 16 parts whose terms name at most nslots <= 64 slots each; a
     * group's parts are formed together, reading each of its slots once per
     * element.  groups: ngroups x gparts part indices (-1: none), or NULL
     * for one group of all np parts.  nslots = 0: one pass per part. */
    int nslots;
    int ngroups, gparts;
    const int32_t* groups;
} fmmhip_wsum_t;
int fmmhip_wsum(int is_double, const fmmhip_wsum_t* w, void* stream);

/* One segment of a stage of the two-stage operand summation
 * (fmmhip_presum.hip): R sums of np parts, every sum a full leaf.  Parts
 * lie at `parts[p].off` from the stage's input, sum r is written at
 * s_off + r*pstride from the stage's output. */
#define FMMHIP_STAGE_PARTS 16
/* Most products per level of the two-stage summation (its kernel holds a
 * level's term list, at most 16 terms per product, in LDS). */
#define FMMHIP_STAGE_MAX_R 64
typedef struct fmmhip_stage_term_s {
    double coef;
    int32_t part;               /* < np */
    int32_t pad;
} fmmhip_stage_term_t;
typedef struct fmmhip_stage_seg_s {
    long s_off;
    long pstride;
    int R, np;                  /* np <= FMMHIP_STAGE_PARTS */
    const int* tptr;            /* device CSR over the R sums into terms (ascending part per sum) */
    const fmmhip_stage_term_t* terms;
    const fmm_entry_t* parts;   /* off, rows, cols */
} fmmhip_stage_seg_t;
/* One stage: every segment over the leaf layout (Lc, Ls, ldo, swap as in
 * fmmhip_presum_seg_t), input X with leading dimension ld, output S. */
int fmmhip_presum_stage(int is_double, const void* X, long ld, void* S, int swap, int Lc, int Ls, int ldo,
                        const fmmhip_stage_seg_t* segs, int nseg, void* stream);

/* Summation stage of the pre-summed pack mode (fmmhip_presum.hip). */
int fmmhip_presum_d(const double* X, long ld, const int* ptr, const fmm_entry_t* ent, int R, int swap,
                    int Lc, int Ls, int ldo, double* S, void* stream);
int fmmhip_presum_s(const float* X, long ld, const int* ptr, const fmm_entry_t* ent, int R, int swap,
                    int Lc, int Ls, int ldo, float* S, void* stream);
/* The same sums from a part table: every part element is read once and all
 * R sums are formed from it (identical arithmetic, fewer HBM reads). */
int fmmhip_presum_seg(int is_double, const fmmhip_presum_seg_t* seg, void* stream);
/* FP64 operand, FP32 leaves (S of type float): each sum formed in FP64 and
 * rounded once to FP32. */
int fmmhip_presum_seg_d2s(const fmmhip_presum_seg_t* seg, void* stream);

#ifdef __cplusplus
}
#endif
#endif
