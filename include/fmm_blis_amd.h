/*
 * fmm_blis_amd.h -- C ABI of the MI355X fast-matrix-multiplication library
 * (libfmm_blis_amd.so).
 *
 * Two layers:
 *
 *  1. The fmm_blis drop-in.  Same names, argument meaning and types as the
 *     plugin and driver of dnparikh/fmm-blis, so a program written against
 *     libfmm.a + libblis_fmm_blis.a relinks against this library unchanged:
 *       bli_plugin_register_fmm_blis  plugin/bli_plugin_register.c:43-64
 *       FMM_BLIS_PACK_UKR/GEMM_UKR    plugin/bli_plugin_fmm_blis.h:45-46
 *       fmm_t                         plugin/bli_plugin_fmm_blis.h:58-68
 *       new_fmm / free_fmm            frame/base/bli_read_fmm.c:24-86
 *       STRASSEN_FMM                  frame/base/bli_strassen_ab.c:12-24
 *       bli_strassen_ab_ex            frame/base/bli_strassen_ab.c:123-295
 *       bli_strassen_ab               frame/base/bli_strassen_ab.c:297-300
 *     The operands are BLIS objects (obj_t, blis/frame/include/
 *     bli_type_defs.h:1215-1245).  When blis.h is included first its obj_t is
 *     used; otherwise a layout-identical declaration follows.
 *
 *  2. The native layer (fmmhip_*): plain pointers, sizes and strides, device
 *     resident operands, HIP streams passed as void*.  The drop-in calls it;
 *     tests, benchmarks, the batching and the serving forms call it directly.
 *
 * All entry points are C (extern "C"), no C++ or HIP types appear here.
 */
#ifndef FMM_BLIS_AMD_H
#define FMM_BLIS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* BLIS types (layout-identical mirror unless blis.h is already included)   */
/* ------------------------------------------------------------------------ */
#ifndef BLIS_H
typedef int64_t  dim_t;
typedef int64_t  inc_t;
typedef int64_t  doff_t;
typedef uint64_t siz_t;
typedef uint32_t objbits_t;
typedef int      err_t;      /* BLIS err_t is an int-sized enum */
typedef uint32_t num_t;

typedef struct { double real; double imag; } fmm_atom_t;   /* BLIS atom_t */

typedef struct obj_s {
    struct obj_s* root;
    dim_t         off[2];
    dim_t         dim[2];
    doff_t        diag_off;
    objbits_t     info;
    objbits_t     info2;
    siz_t         elem_size;
    void*         buffer;
    inc_t         rs;
    inc_t         cs;
    inc_t         is;
    fmm_atom_t    scalar;
    dim_t         m_padded;
    dim_t         n_padded;
    inc_t         ps;
    inc_t         pd;
    dim_t         m_panel;
    dim_t         n_panel;
} obj_t;

#define BLIS_SUCCESS  (-1)
#define BLIS_FAILURE  (-2)
#define BLIS_FLOAT    0x0u
#define BLIS_SCOMPLEX 0x1u
#define BLIS_DOUBLE   0x2u
#define BLIS_DCOMPLEX 0x3u
#define BLIS_INT      0x4u
#define BLIS_CONSTANT 0x5u
#endif /* BLIS_H */

/* Bits of obj_t.info this library reads (bli_type_defs.h:291-340). */
#define FMM_OBJ_DT_MASK    0x7u          /* datatype, bits 0-2 */
#define FMM_OBJ_TRANS_BIT  (0x1u << 3)   /* op(X) = X^T        */
#define FMM_OBJ_CONJ_BIT   (0x1u << 4)   /* no-op for real data */

/* ------------------------------------------------------------------------ */
/* fmm_blis drop-in                                                          */
/* ------------------------------------------------------------------------ */

/* Algorithm descriptor <m~, k~, n~; R> with integer coefficient tables
 * U[(i*k~+p)*R + r], V[(p*n~+j)*R + r], W[(i*n~+j)*R + r].
 * Replaces plugin/bli_plugin_fmm_blis.h:58-68 (same field order). */
typedef struct fmm_s {
    int  m_tilde;
    int  n_tilde;
    int  k_tilde;
    int  R;
    int* U;
    int* V;
    int* W;
} fmm_t;

/* Kernel slot ids.  Replaces plugin/bli_plugin_fmm_blis.h:45-46; set by
 * bli_plugin_register_fmm_blis, which registers the two slots in order
 * (ids 1 and 2) and fills them through the per-configuration init below;
 * fmm_blis_get_ukr_dt reads them back. */
extern siz_t FMM_BLIS_PACK_UKR;
extern siz_t FMM_BLIS_GEMM_UKR;

/* Initialises the HIP device, registers the two kernel slots and runs the
 * per-configuration init, once (thread-safe).  Returns BLIS_SUCCESS (-1),
 * or BLIS_FAILURE (-2) when no usable gfx950 device is present.
 * Replaces plugin/bli_plugin_register.c:43-64. */
err_t bli_plugin_register_fmm_blis(void);

/* Per-configuration init of the one configuration, the gfx950 device:
 * stores the float and double pack kernels (fmmhip_spack_fmm /
 * fmmhip_dpack_fmm) in slot FMM_BLIS_PACK_UKR and the gemm kernels
 * (fmmhip_sgemm_fmm_packed / fmmhip_dgemm_fmm_packed) in slot
 * FMM_BLIS_GEMM_UKR.  Called by bli_plugin_register_fmm_blis.
 * Replaces plugin/ref_kernels/bli_plugin_init_ref.c:61-71
 * (bli_plugin_init_fmm_blis_<cfg>_ref). */
void bli_plugin_init_fmm_blis_gfx950_ref(void);

/* The kernel registered in slot ukr_id for datatype dt (BLIS_FLOAT or
 * BLIS_DOUBLE), or NULL (unknown slot, complex type, no device); cast it to
 * fmm_ukr_s_ft / fmm_ukr_d_ft.  Registers first if needed.  The query a
 * caller of the reference makes with bli_cntx_get_ukr_dt(dt, ukr_id, cntx)
 * on the context the plugin's init filled (blis/frame/base/bli_cntx.h);
 * there is one context here, so it takes none. */
void* fmm_blis_get_ukr_dt(num_t dt, siz_t ukr_id);

/* Built-in <2,2,2;7> Strassen, identical to 222.txt.
 * Replaces frame/base/bli_strassen_ab.c:12-24. */
extern fmm_t STRASSEN_FMM;

/* Reads a coefficient file: header "m k n R", then m*k rows of U, k*n rows
 * of V, m*n rows of W (R integers each); lines starting with '#' are
 * comments, line length is unbounded.  U/V/W are malloc'd and owned by the
 * caller (free_fmm).  On a missing or malformed file a message goes to
 * stderr and the result has R == 0 and NULL tables.
 * Replaces frame/base/bli_read_fmm.c:24-80. */
fmm_t new_fmm(const char* file_name);

/* Frees the tables of a new_fmm result.  Replaces bli_read_fmm.c:82-86. */
void free_fmm(fmm_t fmm);

/* C := beta*C + alpha*op(A)*op(B) with one level of the algorithm fmm.
 * Real single or double objects of one datatype; any of A, B, C may be
 * row- or column-stored (or transposed through the obj_t trans bit); the
 * buffers may be host memory (staged through the device) or device memory
 * (computed in place).  Synchronous, like the reference.  Errors print a
 * message and abort(), as BLIS does with error checking on; see
 * fmm_blis_amd_set_abort_on_error.
 * Replaces frame/base/bli_strassen_ab.c:123-295. */
void bli_strassen_ab_ex(obj_t* alpha, obj_t* A, obj_t* B, obj_t* beta,
                        obj_t* C, fmm_t fmm);

/* bli_strassen_ab_ex(..., STRASSEN_FMM).
 * Replaces frame/base/bli_strassen_ab.c:297-300. */
void bli_strassen_ab(obj_t* alpha, obj_t* A, obj_t* B, obj_t* beta, obj_t* C);

/* ------------------------------------------------------------------------ */
/* Extensions of the drop-in (no counterpart in the reference)              */
/* ------------------------------------------------------------------------ */

/* Leaf-product arithmetic. */
typedef enum {
    FMMHIP_LEAF_NATIVE = 0,   /* the operands' type: FP64 for double, FP32 for float */
    FMMHIP_LEAF_F32    = 1,   /* FP64 operands, FP32 MFMA leaf, FP64 writeback       */
    FMMHIP_LEAF_BF16X3 = 2,   /* FP64 operands, split BF16 MFMA leaf (FP32 accumulate),
                                 FP64 recombination and writeback: four bf16 planes
                                 per operand, ten plane products (normalised error
                                 about 1e-11 for the classical algorithm, a few 1e-11
                                 for one- and two-level Strassen at 8192^3-16384^3)   */
    FMMHIP_LEAF_BF16X3_6 = 3, /* three planes, six plane products (fastest; about 1e-10
                                 classical, 1e-9 two-level Strassen at 16384^3)       */
    FMMHIP_LEAF_BF16X3_8 = 4, /* three planes, eight plane products (about 6e-11
                                 classical, 4e-10 two-level Strassen at 16384^3)      */
    FMMHIP_LEAF_INT8X4 = 5    /* FP64 operands, int8 MFMA leaf with exact int32 sums:
                                 four int8 slices per value with one scale per row
                                 of the summed operand (Ozaki-scheme slicing), ten
                                 slice products, FP64 recombination and writeback
                                 (about 4e-12 classical, 4e-11 two-level Strassen at
                                 16384^3; up to 1.2e-9 at k of a few); a
                                 non-finite value makes its whole row NaN          */
} fmmhip_leaf_t;

/* Like bli_strassen_ab_ex with `nlevels` levels; level l uses algs[l]
 * (mixing algorithms is allowed) and leaf arithmetic `leaf`. */
void bli_fmm_ab_ex(obj_t* alpha, obj_t* A, obj_t* B, obj_t* beta, obj_t* C,
                   const fmm_t* algs, int nlevels, fmmhip_leaf_t leaf);

/* Classical product C := beta*C + alpha*op(A)*op(B) through the same
 * object boundary (the role of BLIS bli_gemm on this path). */
void bli_gemm_hip(obj_t* alpha, obj_t* A, obj_t* B, obj_t* beta, obj_t* C);

/* When off (default on) the object entries record errors instead of
 * aborting; fmm_blis_amd_last_error returns the last one (0 = none). */
void fmm_blis_amd_set_abort_on_error(int on);
int  fmm_blis_amd_last_error(void);

/* The object entries cache up to 16 composed algorithms by table content;
 * a call pins the one it uses, and eviction takes unpinned ones only.
 * Reports the cached and the pinned count (diagnostics). */
void fmm_blis_amd_alg_cache_stats(int* cached, int* pinned);

/* Fills *obj as a view of an existing buffer, like BLIS
 * bli_obj_create_with_attached_buffer (blis/frame/base/bli_obj.c:59-75):
 * dt BLIS_FLOAT or BLIS_DOUBLE, m x n, element (i,j) at p + i*rs + j*cs. */
void fmm_obj_attach(num_t dt, dim_t m, dim_t n, void* p, inc_t rs, inc_t cs,
                    obj_t* obj);
/* Fills *obj as a 1x1 double scalar holding *storage (caller-owned). */
void fmm_obj_scalar(double* storage, obj_t* obj);

/* ------------------------------------------------------------------------ */
/* Native layer                                                              */
/* ------------------------------------------------------------------------ */

typedef enum {
    FMMHIP_SUCCESS         = 0,
    FMMHIP_ERR_INVALID     = 1,  /* bad argument or shape                  */
    FMMHIP_ERR_NOMEM       = 2,  /* host or device allocation failed       */
    FMMHIP_ERR_HIP         = 3,  /* a HIP runtime call or a launch failed  */
    FMMHIP_ERR_UNSUPPORTED = 4,  /* valid but not supported (e.g. complex) */
    FMMHIP_ERR_NODEVICE    = 5,  /* no usable gfx950 device                */
    FMMHIP_ERR_IO          = 6,  /* coefficient file unreadable/malformed  */
    FMMHIP_ERR_TOO_LARGE   = 7   /* beyond the no-allocation scratch        */
} fmmhip_status_t;

const char* fmmhip_status_string(int status);

/* Device set-up (idempotent, thread-safe); 0 on success. */
int fmmhip_init(void);
/* Name of the device's architecture (e.g. "gfx950") into buf. */
int fmmhip_device_arch(char* buf, size_t len);
int fmmhip_device_cu_count(int* cus);

/* Composed algorithm: nlevels levels, level l = algs[l].  The tables are
 * the Kronecker product of the levels (part (i1*m2+i2, p1*k2+p2), product
 * r1*R2+r2).  Validates every level against the Brent equations. */
typedef struct fmmhip_alg_s* fmmhip_alg_t;
int  fmmhip_alg_create(fmmhip_alg_t* out, const fmm_t* algs, int nlevels);
/* One level, R = 1, all coefficients 1: the classical product. */
int  fmmhip_alg_create_classical(fmmhip_alg_t* out);
void fmmhip_alg_destroy(fmmhip_alg_t alg);
/* Composed <m~,k~,n~;R> and nonzero counts of U, V, W. */
int  fmmhip_alg_info(fmmhip_alg_t alg, int* mt, int* kt, int* nt, int* R,
                     long* nnz_u, long* nnz_v, long* nnz_w);
/* Number of equations of the Brent system the tables violate (0 = exact). */
long fmmhip_alg_brent_violations(fmmhip_alg_t alg);
/* Kernel launch shape for this algorithm: 0 = 128 x 128 tiles, 512 threads
 * (one part per product loaded ahead), 1 = 128 x 128 tiles, 256 threads (two
 * parts ahead), 2 = 64 x 64 tiles, 256 threads (column-major A, B and C
 * without the classical k-split; other calls take shape 0), 3 = the fast
 * FP64 kernel (128 x 128 tiles, 256 threads, fragments read one sub-step
 * ahead; FP32 leaf and FP32 operands: k-steps of 32) for single-part
 * operands -- classical, pre-summed, products materialised -- with
 * column-major A, B and C whose leaf extents are multiples of 128 (k of the
 * k-step) and every tile inside its parts; with the products materialised
 * the library pads the leaves and slots in its workspace to whole tiles and
 * k-steps, so every such call qualifies (other calls take shape 0), -1 =
 * chosen per call (default: 3 where it applies, else 0, or 1 for the fused
 * pack mode).  The result does not depend on the shape (every element's
 * sums run in the same order).  The batch's tuner chooses between 128 x 128
 * tiles (shape 0, interior tiles on the fast kernel) and shape 2 per
 * problem.  fmmhip_alg_last_shape reports the shape the last FP64/FP32 call
 * launched (-1 before the first). */
int  fmmhip_alg_set_shape(fmmhip_alg_t alg, int shape);
int  fmmhip_alg_last_shape(fmmhip_alg_t alg, int* shape);
/* How the U/V-weighted operand sums are formed and the W-weighted products
 * applied: 0 = fused (operand sums streamed from the operands into LDS
 * inside the MFMA kernel, W scatter in its epilogue, no workspace); 1 =
 * pre-summed (an HBM-bandwidth pass writes every S_r and T_r once to a
 * per-stream device workspace of R*(leaf(A) + leaf(B)) elements; the MFMA
 * kernel reads one part per product and scatters with W in its epilogue);
 * 2 = pre-summed with the products materialised (the MFMA kernel stores
 * every M_r once into R more leaves of C's size, leaf(C) each, and a
 * bandwidth pass forms each C part from the products it needs: one pass
 * per part instead of one read-modify-write of C per (product, part), and
 * any number of workgroups per tile, up to R, without copies of C; the
 * same sums in the same order as 1 without split, so the same C, except
 * for two-level tables whose W stage runs in two stages, inner level then
 * outer (333, 444: fewer bytes, sums re-associated); the batch uses 1
 * instead); -1 = chosen per call (default: for multi-part
 * algorithms pre-summed when the call's whole workspace fits in half of
 * the device's total memory, materialised too when that fits as well and a
 * time model expects it to be faster; the choice depends on the algorithm,
 * the shape, the layouts and the device, never on what memory is free, so
 * repeated calls round alike.  Only when the workspace cannot be allocated
 * does the call fall back to 0 without split, and fmmhip_alg_last_pack
 * reports it).  No counterpart in the reference, whose packing
 * (plugin/ref_kernels/bli_packm_fmm_ref.c) and scatter
 * (plugin/ref_kernels/bli_gemm_fmm_ref.c) are always fused. */
int  fmmhip_alg_set_pack(fmmhip_alg_t alg, int pack);
/* The pack mode (0, 1, 2 as above) a column-major FP64 call of m x n x k
 * with this algorithm and leaf resolves to (the call's own resolution, so
 * the two agree unless the workspace allocation fails).  BF16x3 leaves
 * always stage their operand planes; they report 2 when the products are
 * materialised (pack 2, or -1 with several products and the slots in half
 * of the device's memory), else 1. */
int  fmmhip_alg_pack_for(fmmhip_alg_t alg, fmmhip_leaf_t leaf, long m, long n, long k, int* pack);
/* The pack mode the last FP64/FP32-leaf call with this algorithm took (-1
 * before the first): differs from fmmhip_alg_pack_for only when an
 * automatic call fell back to the fused mode for lack of device memory. */
int  fmmhip_alg_last_pack(fmmhip_alg_t alg, int* pack);
/* Split: 1 = each 128 x 128 tile of C is one workgroup looping over all R
 * products; 2, 4, 8 = that many workgroups per tile, each taking a
 * contiguous range of products (the classical algorithm, R = 1: a
 * contiguous range of at least 8 k-steps of 16) and scattering into its own
 * zeroed copy of C, summed into C in group order by a second kernel
 * (deterministic, but rounded differently from 1); -1 = chosen per call
 * (default) by a time model of the wave fill and the copies' traffic.  No
 * counterpart in the reference: its CPU loop over products is serial per
 * thread (frame/base/bli_strassen_ab.c:248); this is the GPU's occupancy knob. */
int  fmmhip_alg_set_split(fmmhip_alg_t alg, int groups);
/* Frees the pre-summed mode's workspaces (synchronises the device). */
int  fmmhip_workspace_release(void);

/* Launch timing: while enabled, every FMM/GEMM call records HIP events on
 * its stream around the summation stage and the MFMA kernel.  _read waits
 * for them, returns the number of timed calls, the summed kernel and
 * summation times and the longest kernel (ms), and clears the record. */
int fmmhip_timing_enable(int on);
int fmmhip_timing_read(long* launches, double* kernel_ms, double* presum_ms, double* max_kernel_ms);

/* C := beta*C + alpha*A*B on device pointers, asynchronous on `stream`
 * (NULL = the library's stream).  Element (i,j) of X is X[i*rsX + j*csX]
 * for any non-negative strides; an A or B with neither stride 1 is first
 * gathered column-major into a per-stream buffer by one bandwidth pass.
 * The leaf `leaf` applies to the double entry; the float entry always
 * computes FP32 leaves. */
int fmmhip_dgemm_fmm(fmmhip_alg_t alg, fmmhip_leaf_t leaf,
                     long m, long n, long k, double alpha,
                     const double* A, long rsA, long csA,
                     const double* B, long rsB, long csB, double beta,
                     double* C, long rsC, long csC, void* stream);
int fmmhip_sgemm_fmm(fmmhip_alg_t alg, long m, long n, long k, float alpha,
                     const float* A, long rsA, long csA,
                     const float* B, long rsB, long csB, float beta,
                     float* C, long rsC, long csC, void* stream);
/* fmmhip_dgemm_fmm with the FP64 leaf in the pre-summed mode without
 * product split (the same kernels and the same C as fmmhip_dgemm_fmm with
 * pack 1 and split 1), calling no hipMalloc, hipFree or other allocating
 * routine: its lists and S_r / T_r leaves live in a static device scratch of
 * FMMHIP_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends
 * (8 MiB of lists, the rest leaves: R*(leaf(A) + leaf(B)) doubles, e.g.
 * one-level Strassen up to 2048^3).  Larger problems return
 * FMMHIP_ERR_TOO_LARGE; an A or B with neither stride 1 returns
 * FMMHIP_ERR_UNSUPPORTED.  Every call on one GPU shares its scratch: calls
 * on one stream are ordered, calls that overlap on one GPU (different
 * streams) race.  No counterpart in the reference: for callers that
 * cannot allocate at call time (HIP graph capture, tight serving loops,
 * harnesses that forbid device allocation). */
#define FMMHIP_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int fmmhip_dgemm_fmm_noalloc(fmmhip_alg_t alg, long m, long n, long k, double alpha,
                             const double* A, long rsA, long csA, const double* B, long rsB, long csB,
                             double beta, double* C, long rsC, long csC, void* stream);
/* The project's classical FP64 / FP32 MFMA GEMMs (same kernels, R = 1). */
int fmmhip_dgemm(long m, long n, long k, double alpha,
                 const double* A, long rsA, long csA,
                 const double* B, long rsB, long csB, double beta,
                 double* C, long rsC, long csC, void* stream);
int fmmhip_sgemm(long m, long n, long k, float alpha,
                 const float* A, long rsA, long csA,
                 const float* B, long rsB, long csB, float beta,
                 float* C, long rsC, long csC, void* stream);

/* ---- the pre-summed mode in two stages: the registry's kernels ---- */

/* Pack (stage 1): forms every S_r = sum U[.][r] A_part and T_r = sum
 * V[.][r] B_part of the call into `ws` (HBM-bandwidth pass; alpha, beta and
 * C are not read).  Gemm (stage 2): the R leaf products S_r*T_r on the
 * matrix cores from `ws`, each added into every C part with alpha*W (A and
 * B are not read, only their layouts).  Both take the same arguments, one
 * unit stride per operand, device pointers, asynchronous on `stream`; stage
 * 1 then stage 2 on one stream equals fmmhip_dgemm_fmm with pack 1 and
 * split 1 bit for bit.  ws must hold fmmhip_fmm_workspace_bytes bytes
 * (16-byte aligned; hipMalloc memory is).  The reference's counterparts are
 * the pack and gemm micro-kernels the plugin registers:
 * plugin/ref_kernels/bli_packm_fmm_ref.c:50-148 and
 * plugin/ref_kernels/bli_gemm_fmm_ref.c:39-121. */
typedef int (*fmm_ukr_d_ft)(fmmhip_alg_t alg, long m, long n, long k, double alpha,
                            const double* A, long rsA, long csA, const double* B, long rsB, long csB,
                            double beta, double* C, long rsC, long csC, void* ws, size_t ws_bytes,
                            void* stream);
typedef int (*fmm_ukr_s_ft)(fmmhip_alg_t alg, long m, long n, long k, float alpha,
                            const float* A, long rsA, long csA, const float* B, long rsB, long csB,
                            float beta, float* C, long rsC, long csC, void* ws, size_t ws_bytes,
                            void* stream);
int fmmhip_fmm_workspace_bytes(fmmhip_alg_t alg, int is_double, long m, long n, long k, long rsA, long csA,
                               long rsB, long csB, long rsC, long csC, size_t* bytes);
int fmmhip_dpack_fmm(fmmhip_alg_t alg, long m, long n, long k, double alpha, const double* A, long rsA,
                     long csA, const double* B, long rsB, long csB, double beta, double* C, long rsC,
                     long csC, void* ws, size_t ws_bytes, void* stream);
int fmmhip_dgemm_fmm_packed(fmmhip_alg_t alg, long m, long n, long k, double alpha, const double* A,
                            long rsA, long csA, const double* B, long rsB, long csB, double beta,
                            double* C, long rsC, long csC, void* ws, size_t ws_bytes, void* stream);
int fmmhip_spack_fmm(fmmhip_alg_t alg, long m, long n, long k, float alpha, const float* A, long rsA,
                     long csA, const float* B, long rsB, long csB, float beta, float* C, long rsC,
                     long csC, void* ws, size_t ws_bytes, void* stream);
int fmmhip_sgemm_fmm_packed(fmmhip_alg_t alg, long m, long n, long k, float alpha, const float* A,
                            long rsA, long csA, const float* B, long rsB, long csB, float beta, float* C,
                            long rsC, long csC, void* ws, size_t ws_bytes, void* stream);

/* ---- batching form: many independent column-major GEMMs, one graph ---- */

/* Candidates the tuner chooses from (also accepted as a forced choice):
 * 0 classical, 1 one-level Strassen, 2 two-level Strassen (pre-summed). */
typedef struct fmmhip_batch_s* fmmhip_batch_t;
int  fmmhip_batch_create(fmmhip_batch_t* out);
void fmmhip_batch_destroy(fmmhip_batch_t b);
/* Adds C := beta*C + alpha*A*B (column-major, leading dimensions lda, ldb,
 * ldc; device pointers).  cand = -1 lets the tuner choose.  Invalidates a
 * previous plan. */
int  fmmhip_batch_add_d(fmmhip_batch_t b, long m, long n, long k, double alpha, const double* A, long lda,
                        const double* B, long ldb, double beta, double* C, long ldc, int cand);
int  fmmhip_batch_add_s(fmmhip_batch_t b, long m, long n, long k, float alpha, const float* A, long lda,
                        const float* B, long ldb, float beta, float* C, long ldc, int cand);
/* Chooses every problem's candidate and tile (128 x 128 or 64 x 64: the
 * smaller tile wastes less on padding where a leaf dimension is an odd
 * multiple of 64, the larger runs faster on whole tiles), prepares
 * descriptors and the workspace, and captures one HIP graph: per datatype
 * one batched summation launch, then per tile one persistent MFMA launch
 * over (problem, tile) items, costliest first, taken from a device queue.
 * The tile does not change the result: each element of C is accumulated in
 * the same order on either. */
int  fmmhip_batch_plan(fmmhip_batch_t b);
/* Launches the captured graph on `stream` (asynchronous; may be replayed). */
int  fmmhip_batch_run(fmmhip_batch_t b, void* stream);
int  fmmhip_batch_info(fmmhip_batch_t b, long* nprob, long* items_f64, long* items_f32, double* est_ms,
                       size_t* ws_bytes);
/* Of the planned work items (problem, tile), how many run on the fast
 * kernel (whole problems whose leaves it takes, and the interior tiles of
 * problems on 128 x 128 tiles). */
int  fmmhip_batch_fast_items(fmmhip_batch_t b, long* fast_f64, long* fast_f32);
int  fmmhip_batch_choice(fmmhip_batch_t b, long i, int* cand);
/* Forces problem i's tile edge (64 or 128; 0 = the tuner's choice);
 * invalidates a previous plan.  fmmhip_batch_tile reports the planned edge
 * (the 128 x 128 kernel takes what the 64 x 64 one cannot). */
int  fmmhip_batch_set_tile(fmmhip_batch_t b, long i, int tile);
int  fmmhip_batch_tile(fmmhip_batch_t b, long i, int* tile);

/* Tuner: the cost model's choice (and estimate, ms) for one problem, alone
 * on the device (lone = 1) or inside a batch; the estimate of a given
 * candidate (on its faster tile); calibration by timing each candidate on the device; the
 * calibrated table saved to / loaded from a text file. */
int fmmhip_tuner_choose(int is_double, long m, long n, long k, int lone, int* cand, double* est_ms);
/* The same choice with its tile edge (64 or 128). */
int fmmhip_tuner_choose_tile(int is_double, long m, long n, long k, int lone, int* cand, int* tile,
                             double* est_ms);
int fmmhip_tuner_estimate(int is_double, int cand, long m, long n, long k, int lone, double* est_ms);
int fmmhip_tuner_calibrate(void* stream);
int fmmhip_tuner_save(const char* path);
int fmmhip_tuner_load(const char* path);

/* ---- serving form: a request loop with continuous batching ---- */

/* One GEMM request, C := beta*C + alpha*A*B, column-major, operands resident
 * on the device.  C == NULL: the server allocates C (ldc = m) in its arena
 * (beta must be 0); keep = 1 keeps it until fmmhip_server_release, keep = 0
 * drops it at completion (latency/throughput only).  cand: -1 tuner, 0
 * classical, 1/2 levels of Strassen (pre-summed, no product split; the
 * classical k-split as for a single call: a request's result does not
 * depend on what else is in flight).  arrival_s:
 * the time, on the server's
 * clock (fmmhip_server_now), at which the request arrives; earlier means now. */
typedef struct {
    int is_double;
    long m, n, k;
    double alpha, beta;
    const void* A; long lda;
    const void* B; long ldb;
    void* C; long ldc;
    int cand;
    int keep;
    double arrival_s;
} fmmhip_request_t;

typedef struct {
    long submitted, completed, failed;
    double p50_ms, p99_ms, mean_ms, max_ms;   /* arrival -> observed completion */
    double span_s;                            /* first arrival -> last completion */
    size_t arena_bytes, arena_peak;
    int max_inflight;
} fmmhip_server_stats_t;

typedef struct fmmhip_server_s* fmmhip_server_t;
/* A server with an HBM budget of `hbm_budget` bytes (one arena, allocated
 * here) and `slots` streams (requests in flight at once). */
int    fmmhip_server_create(fmmhip_server_t* out, size_t hbm_budget, int slots);
double fmmhip_server_now(fmmhip_server_t s);
/* Queues a request (thread-safe, returns at once). */
int    fmmhip_server_submit(fmmhip_server_t s, const fmmhip_request_t* req, long* id);
/* Blocks until request `id` is done; *status = 0 or its FMMHIP_ERR_* code. */
int    fmmhip_server_wait(fmmhip_server_t s, long id, int* status);
/* A done request's C, ldc, latency and candidate; returns its status. */
int    fmmhip_server_result(fmmhip_server_t s, long id, void** C, long* ldc, double* latency_ms, int* cand);
/* Forgets a done request (frees a kept server-allocated C). */
int    fmmhip_server_release(fmmhip_server_t s, long id);
/* Blocks until every submitted request is done. */
int    fmmhip_server_drain(fmmhip_server_t s);
int    fmmhip_server_stats(fmmhip_server_t s, fmmhip_server_stats_t* st);
void   fmmhip_server_destroy(fmmhip_server_t s);

/* Device memory, streams, events (thin wrappers so that C and ctypes
 * callers need no HIP headers). */
int fmmhip_malloc(void** p, size_t bytes);
int fmmhip_free(void* p);
int fmmhip_memcpy_h2d(void* dst, const void* src, size_t bytes);
int fmmhip_memcpy_d2h(void* dst, const void* src, size_t bytes);
int fmmhip_memset(void* p, int value, size_t bytes);
/* Device-to-device copy, asynchronous on `stream`. */
int fmmhip_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
/* Device-to-device copy by a plain streaming kernel (bytes and both
 * pointers multiples of 16): a workload of known size for calibrating the
 * HBM byte counters of a profiler. */
int fmmhip_copy_kernel(void* dst, const void* src, size_t bytes, void* stream);
/* Measured peaks (no counterpart in the reference; for rooflines): the FP64
 * MFMA rate of back-to-back independent v_mfma_f64_16x16x4_f64 chains on
 * every SIMD (TFLOP/s), and the read+write rate of a 4 GiB copy (GB/s, the
 * faster of the plain and the streaming copy of fmmhip_measure_hbm).  Takes
 * about two seconds and 8 GiB of device memory. */
int fmmhip_measure_peaks(double* fp64_mfma_tflops, double* copy_gbs);
/* The device's practical HBM ceilings (GB/s, best of grids of 1, 2 and 4
 * blocks per CU over 4 GiB buffers, 16-byte non-temporal accesses): a copy
 * with four loads in flight per thread (read + write bytes), a read-only and
 * a write-only stream.  Takes under a second and 8 GiB of device memory. */
int fmmhip_measure_hbm(double* copy_gbs, double* read_gbs, double* write_gbs);
int fmmhip_stream_create(void** stream);
int fmmhip_stream_destroy(void* stream);
int fmmhip_stream_sync(void* stream);
int fmmhip_device_sync(void);
int fmmhip_event_create(void** ev);
int fmmhip_event_destroy(void* ev);
int fmmhip_event_record(void* ev, void* stream);
int fmmhip_event_elapsed_ms(float* ms, void* ev_start, void* ev_stop);
/* Device memory currently free / total, in bytes. */
int fmmhip_mem_info(size_t* free_bytes, size_t* total_bytes);

/* Seeded inputs on the device.  Element (i,j) of an m x n column-major
 * matrix (leading dimension ld) gets the value of counter j*m+i of stream
 * `stream_id` under `seed` (SplitMix64 -> 2u-1, u = 53-bit uniform), so the
 * bits equal a CPU fill with the same rule.  fill_int draws integers in
 * [-amp, amp] from the same counters. */
int fmmhip_fill_uniform(double* X, long m, long n, long ld, uint64_t seed,
                        uint64_t stream_id, void* stream);
int fmmhip_fill_uniform_f32(float* X, long m, long n, long ld, uint64_t seed,
                            uint64_t stream_id, void* stream);
int fmmhip_fill_int(double* X, long m, long n, long ld, uint64_t seed,
                    uint64_t stream_id, int amp, void* stream);
/* Scales row i of X by rowscale[i] (if non-NULL) and column j by
 * colscale[j] (if non-NULL); scales are device arrays. */
int fmmhip_scale_rows_cols(double* X, long m, long n, long ld,
                           const double* rowscale, const double* colscale,
                           void* stream);
/* ||X - Y||_F, ||X||_F computed on the device in FP64 (result to host). */
int fmmhip_fro_diff(const double* X, const double* Y, long m, long n,
                    long ldx, long ldy, double* out);
int fmmhip_fro(const double* X, long m, long n, long ld, double* out);

#ifdef __cplusplus
}
#endif
#endif /* FMM_BLIS_AMD_H */
