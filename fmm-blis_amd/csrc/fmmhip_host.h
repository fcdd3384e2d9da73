// fmmhip_host.h -- host-side call preparation shared by the one-problem
// entry points (fmmhip_gemm.hip) and the batched form (fmmhip_batch.hip).
#ifndef FMMHIP_HOST_H
#define FMMHIP_HOST_H

#include <hip/hip_runtime.h>

#include <mutex>

#include "fmmhip_kernel.h"

enum class Kind { D64, D_F32LEAF, S32 };

// Two-stage summation of a two-level algorithm's operands (device tables):
// stage 1 forms, per outer block, the inner level's sums of the block's
// parts; stage 2 forms, per inner product, the outer level's sums of those.
struct KronView {
    bool useA, useB;
    const fmmhip_stage_seg_t *a1, *a2, *b1, *b2;   // stage 1 and 2 segments per operand
    int na1, na2, nb1, nb2;
    long x1_elems;               // stage-1 leaves (the larger operand's)
};

// Two-stage W stage of the products-materialised mode (device tables):
// stage 1 forms, per outer product and inner C part, the inner level's
// W-weighted sum of the slots (the Y leaves); stage 2 forms each C part
// from the Y leaves with the outer level's W.
struct WKronView {
    const fmm_entry_t *parts1, *terms1;   // Y leaves (R0 x inner parts) and their terms
    const fmm_entry_t *parts2, *terms2;   // C parts and their terms over the Y leaves
    int np1, np2;
    long y_leaves;
    // all-parts groups: stage 1 one per outer product (its inner C parts),
    // stage 2 one per inner C part (the outer C parts that hold it)
    const int32_t *groups1, *groups2;
    int ngroups1, gparts1, nslots1, ngroups2, gparts2, nslots2;
};

struct Prepared {
    Kind kind;
    hipStream_t stream;
    bool nothing;                // empty problem
    bool scale_only;             // k == 0 or alpha == 0: C := beta*C
    void* C;
    long m, n, rsC, csC;
    double beta;
    fmmk::KArgs ka;              // kernel arguments (A/B are workspace offsets until bound)
    int shape;
    bool a_mn, b_mn, c_col;
    bool single_part;            // one part per product and operand (classical, pre-summed)
    long nwg;
    int mode;                    // pack mode taken: 0 fused, 1 pre-summed, 2 products materialised
    bool untimed;                // never recorded by the launch timing
    bool presum;                 // pre-summed pack mode
    size_t ws_bytes;             // workspace: summation leaves, then the split temporaries
    fmmhip_presum_seg_t segA, segB;
    size_t tmp_off;              // split temporaries: workspace offset (ka.groups > 1)
    KronView kron;               // two-stage summation (pre-summed modes, two-level algorithms)
    size_t x1_off;               // its stage-1 leaves: workspace offset
    void* x1;
    bool wmat;                   // products materialised: kernel into slots, then the W stage
    size_t m_off;                // slots: workspace offset
    fmmhip_wsum_t wsum;          // the W stage's arguments (slots bound with the workspace)
    bool w2;                     // two-stage W: wsum into the Y leaves, then wsum2 into C
    bool f32_leaves;             // FP32-leaf kind: the pre-summed leaves are FP32 (column-major operands)
    size_t y_off;                // the Y leaves: workspace offset
    fmmhip_wsum_t wsum2;
};

// Device coefficient lists of one call: the fused list (every nonzero part
// of every product: CSR over U|V|W, then the entries), a host copy of its
// entries, and in the pre-summed mode the one-part-per-product list.
struct DescView {
    const int* fused_ptr;
    const fmm_entry_t* fused_ent;
    const fmm_entry_t* fused_ent_h;
    const int* presum_ptr;
    const fmm_entry_t* presum_ent;
    const int* mat_ptr;          // products-materialised list (one-problem path only)
    const fmm_entry_t* mat_ent;
    KronView kron;               // two-stage summation tables (one-problem path only)
    WKronView wk;                // two-stage W stage tables (materialised, one-problem path only)
};

// Where prepare_call takes the lists from; without one, from the
// algorithm's per-geometry cache.
struct DescSource {
    // mat: also the products-materialised list (only asked of sources whose
    // can_mat() is true)
    virtual int get(fmmhip_alg_t alg, const fmm_geom_t& g, bool presum, bool mat, bool a_mn, bool b_mn,
                    hipStream_t st, DescView* v) = 0;
    virtual bool can_mat() const { return false; }
    virtual ~DescSource() {}
};

// Upper bound of the device bytes of one call's lists, and the packing of
// them into `host` in their device layout with pointers bound to `dev`
// (the caller copies host -> dev before the launch).
size_t desc_bound(fmmhip_alg_t alg);
int desc_pack(fmmhip_alg_t alg, const fmm_geom_t& g, bool presum, bool mat, bool a_mn, bool b_mn, void* host,
              const void* dev, size_t cap, size_t* used, DescView* v);

// The fused and pre-summed lists of one call built directly into `host`
// (cap bytes; no allocation), in the device layout of desc_pack with
// pointers bound to `dev`; *used = the bytes to copy to dev.
int desc_pack_presum_into(fmmhip_alg_t alg, const fmm_geom_t& g, bool a_mn, bool b_mn, void* host, const void* dev,
                          size_t cap, size_t* used, DescView* v);

// Resolves one call (pack: 0 fused, 1 pre-summed, 2 pre-summed with the
// products materialised, -1 automatic; split: product groups per tile, 1,
// 2, 4 or 8, -1 chosen from the tile count; with pack 2 any count up to R).
int prepare_call(fmmhip_alg_t alg, Kind kind, long m, long n, long k, double alpha, const void* A, long rsA,
                 long csA, const void* B, long rsB, long csB, double beta, void* C, long rsC, long csC, int pack,
                 int split, hipStream_t st, Prepared* P, DescSource* src = nullptr);
// Rebases a pre-summed call's workspace offsets onto `ws`.
void bind_workspace(Prepared* P, void* ws);
// Launches a prepared, bound call on its stream: stages & 1 the summation
// stage, stages & 2 the MFMA kernel and what follows it (split reduction, W
// stage); the launch timing records only whole calls.
int launch_prepared(const Prepared* P, int stages = 3);
// Groups per C tile (1, 2, 4, 8; at most max_groups) for a grid of `tiles`
// workgroups: `requested` when > 0, else chosen by a time model from the
// leaf flops, the bytes of C, the leaf's MFMA peak (0: FP64 or FP32 by
// `fp64`) and the kernel's resident workgroups per CU.
int split_groups(int requested, int max_groups, long tiles, double leaf_flops, double c_bytes, bool fp64, double peak,
                 int wg_per_cu);
// The shared classical (1x1x1, R = 1) algorithm.
fmmhip_alg_t classical_alg();

// The cached fused descriptor of geometry g (device CSR U|V|W and entries).
int fused_desc(fmmhip_alg_t alg, const fmm_geom_t& g, hipStream_t st, const int** dptr,
               const fmm_entry_t** dent);
// The products-materialised lists of geometry g (cached like the others):
// the kernel's W list (one slot entry per product) and the W stage's part
// table and terms, with the slot layout.
struct MatView {
    const int* wptr;             // CSR over products into ent
    const fmm_entry_t* ent;
    const fmm_entry_t* parts;
    const fmm_entry_t* terms;
    int np;
    int swap, Lc, Ls;
    long ldm, slot;
};
int mat_view(fmmhip_alg_t alg, const fmm_geom_t& g, hipStream_t st, MatView* v);
// Groups per tile of the products-materialised mode for a grid of `tiles`
// workgroups on `slots` resident workgroup slots.
int mat_groups(int R, long tiles, int slots);
// The stream's grow-only workspace of at least `bytes` (nullptr: no memory).
void* workspace_for(hipStream_t st, size_t bytes);
// Held by a call from its workspace lookup to its last launch (host-side
// enqueue only), so that concurrent calls on one stream keep their
// workspace while they enqueue.
std::unique_lock<std::mutex> enqueue_lock();
// Launch timing hooks (fmmhip_timing_*): whether timing is on, an event
// recorded on `st`, and one timed call (summation/split from e0 to e1,
// MFMA kernel from e1 to e2).
bool timing_active();
void timing_record(hipEvent_t* ev, hipStream_t st);
void timing_push(hipEvent_t e0, hipEvent_t e1, hipEvent_t e2, bool presum);

// The BF16x3 leaf (fmmhip_bf16x3.hip): split stage and bf16 MFMA GEMM with
// `terms` (6 or 8) plane products.
int bf16x3_gemm(fmmhip_alg_t alg, int terms, long m, long n, long k, double alpha, const double* A, long rsA, long csA,
                const double* B, long rsB, long csB, double beta, double* C, long rsC, long csC, hipStream_t st);
// The INT8x4 leaf (fmmhip_int8x4.hip): row maxima, int8 slices and the int8
// MFMA GEMM with ten slice products.
int int8x4_gemm(fmmhip_alg_t alg, long m, long n, long k, double alpha, const double* A, long rsA, long csA,
                const double* B, long rsB, long csB, double beta, double* C, long rsC, long csC, hipStream_t st);

#endif
