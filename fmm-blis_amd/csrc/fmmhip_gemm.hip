// fmmhip_gemm.hip -- host side of the fused FMM kernel: per-geometry
// coefficient descriptors on the device, launch-shape selection, and the
// fmmhip_{d,s}gemm{,_fmm} entry points of the C ABI.  The kernel itself is
// in fmmhip_kernel.h, instantiated per datatype in fmmhip_gemm_*.hip.
#include <hip/hip_runtime.h>
#include <cmath>
#include <mutex>
#include <algorithm>
#include <vector>
#include <string.h>
#include <stdlib.h>

#include "fmmhip_kernel.h"
#include "fmmhip_host.h"

using fmmk::KArgs;
using fmmk::BM;
using fmmk::BN;

namespace {

// ---- host side -------------------------------------------------------------

// Device copy of an entry list: [ptr (3*(R+1) int32), pad, entries].
struct DevList {
    int32_t* ptr_h = nullptr;
    fmm_entry_t* ent_h = nullptr;
    long nent = 0;
    void* dev = nullptr;
    size_t ent_off = 0;
    const int* dptr() const { return (const int*)dev; }
    const fmm_entry_t* dent() const { return (const fmm_entry_t*)((char*)dev + ent_off); }
};

struct DevDesc {
    fmm_geom_t geo;
    DevList fused;               // every nonzero part of every product
    DevList presum;              // one part per product (the workspace leaf), W as in fused
    bool has_presum = false;
    DevList mat;                 // as presum, W = one full-leaf slot per product, then the W-stage tables
    bool has_mat = false;
    void* kron_dev = nullptr;    // two-stage summation tables (one allocation)
    std::vector<char> kron_host; // their host image (source of the upload)
    KronView kron{};
    bool has_kron = false;
    void* wk_dev = nullptr;      // two-stage W stage tables (one allocation)
    std::vector<char> wk_host;
    WKronView wk{};
    bool has_wk = false;
    unsigned long last_use = 0;
};

struct DevCache {
    std::mutex mu;
    std::vector<DevDesc*> items;
    unsigned long clock = 0;
};

constexpr size_t kCacheCap = 64;

bool same_geo(const fmm_geom_t& a, const fmm_geom_t& b)
{
    return a.m == b.m && a.n == b.n && a.k == b.k && a.rsA == b.rsA && a.csA == b.csA &&
           a.rsB == b.rsB && a.csB == b.csB && a.rsC == b.rsC && a.csC == b.csC;
}

void free_list(DevList& l)
{
    if (l.dev) (void)hipFree(l.dev);
    free(l.ptr_h);
    free(l.ent_h);
    l = DevList();
}

void free_desc(DevDesc* d)
{
    free_list(d->fused);
    free_list(d->presum);
    free_list(d->mat);
    if (d->kron_dev) (void)hipFree(d->kron_dev);
    if (d->wk_dev) (void)hipFree(d->wk_dev);
    delete d;
}

int upload_list(DevList& l, int R, hipStream_t stream)
{
    size_t ptr_bytes = sizeof(int32_t) * 3 * (size_t)(R + 1);
    l.ent_off = (ptr_bytes + 255) & ~(size_t)255;
    size_t total = l.ent_off + sizeof(fmm_entry_t) * (size_t)(l.nent + 1);
    if (hipMalloc(&l.dev, total) != hipSuccess) { (void)hipGetLastError(); l.dev = nullptr; return FMMHIP_ERR_NOMEM; }
    if (hipMemcpyAsync(l.dev, l.ptr_h, ptr_bytes, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync((char*)l.dev + l.ent_off, l.ent_h, sizeof(fmm_entry_t) * (size_t)l.nent,
                       hipMemcpyHostToDevice, stream) != hipSuccess)
        return FMMHIP_ERR_HIP;
    return FMMHIP_SUCCESS;
}

// Whether the pre-summed leaves and the product slots of geometry g are laid
// out for whole 128 x 128 tiles and whole FP64 fast-kernel k-steps (16):
// column-major A, B and C whose leaf m or n is not a multiple of 128 or
// whose leaf k is not a multiple of 16.  A's leaves then take 128-aligned
// columns and up16(k) of them, B's leaves up16(k) rows and a tail after the
// last one, the slots 128-aligned rows and columns, so that
// a kernel running whole tiles and k-steps (the fast kernel) reads and
// writes inside the workspace.  The summation runs over the padded k (its
// loads are clamped into the parts and zero outside them), so the k
// padding holds zeros and adds nothing to a product; the padding rows of A
// (never written) and columns of B (the next leaf's first columns) meet
// only the padding rows and columns of the slots, which nothing reads.
bool tile_padded(const fmm_geom_t& g)
{
    return g.rsA == 1 && g.rsB == 1 && g.rsC == 1 && g.lk > 0 &&
           (g.lm % fmmk::BM != 0 || g.ln % fmmk::BN != 0 || g.lk % fmmk::BK != 0);
}
long up_tile(long x) { return (x + fmmk::BM - 1) / fmmk::BM * fmmk::BM; }
long up_kstep(long k) { return (k + fmmk::BK - 1) / fmmk::BK * fmmk::BK; }

// Workspace leaf layout of the pre-summed mode for one operand: R leaves of
// Lc x Ls (contiguous dim c), leading dimension ldo = Lc rounded up to even
// (tile-padded geometries: A's ldo up to 128 and its Ls, B's Lc = ldo, to
// whole k-steps), then `tail` elements.
struct WsLayout {
    int Lc, Ls, ldo;
    long tail;
    size_t elems(int R) const { return (size_t)ldo * Ls * R + (size_t)tail; }
};

WsLayout ws_layout_a(const fmm_geom_t& g, bool a_mn)
{
    WsLayout w;
    w.Lc = (int)(a_mn ? g.lm : g.lk);
    w.Ls = (int)(a_mn ? g.lk : g.lm);
    w.ldo = (w.Lc + 1) & ~1;
    w.tail = 0;
    if (a_mn && tile_padded(g)) {
        w.ldo = (int)up_tile(g.lm);
        w.Ls = (int)up_kstep(g.lk);
    }
    return w;
}

WsLayout ws_layout_b(const fmm_geom_t& g, bool b_mn)
{
    WsLayout w;
    w.Lc = (int)(b_mn ? g.ln : g.lk);
    w.Ls = (int)(b_mn ? g.lk : g.ln);
    w.ldo = (w.Lc + 1) & ~1;
    w.tail = 0;
    if (!b_mn && tile_padded(g)) {
        w.Lc = w.ldo = (int)up_kstep(g.lk);
        w.tail = (up_tile(g.ln) - g.ln) * w.ldo;
    }
    return w;
}

// The pre-summed list: product r reads leaf r of the A and B workspaces
// with coefficient 1 over the full leaf; its W entries are the fused ones.
// Entries of the pre-summed list: R A leaves, R B leaves, the W entries,
// then the part tables of A (mt*kt) and B (kt*nt) for the part-staged
// summation.
long presum_list_entries(const struct fmmhip_alg_s* alg)
{
    return 2L * alg->R + alg->nnz_w + (long)alg->mt * alg->kt + (long)alg->kt * alg->nt;
}

// The pre-summed list into caller memory (ptr[3*(R+1)], ent zeroed) from
// the fused list's host arrays.
void build_presum_list_into(const struct fmmhip_alg_s* alg, const fmm_geom_t& g, const int32_t* fptr,
                            const fmm_entry_t* fent, bool a_mn, bool b_mn, int32_t* ptr, fmm_entry_t* ent)
{
    const int R = alg->R;
    const WsLayout wa = ws_layout_a(g, a_mn), wb = ws_layout_b(g, b_mn);
    const int npa = alg->mt * alg->kt;
    for (int r = 0; r <= R; r++) {
        ptr[r] = r;
        ptr[(R + 1) + r] = R + r;
        ptr[2 * (R + 1) + r] = 2 * R + (fptr[2 * (R + 1) + r] - (int32_t)(alg->nnz_u + alg->nnz_v));
    }
    for (int r = 0; r < R; r++) {
        fmm_entry_t& ea = ent[r];
        ea.off = (int64_t)r * wa.ldo * wa.Ls;
        ea.rows = (int32_t)g.lm;
        ea.cols = (int32_t)g.lk;
        ea.coef = 1.0;
        ea.part = -1;
        fmm_entry_t& eb = ent[R + r];
        eb.off = (int64_t)r * wb.ldo * wb.Ls;
        eb.rows = (int32_t)g.lk;
        eb.cols = (int32_t)g.ln;
        eb.coef = 1.0;
        eb.part = -1;
    }
    memcpy(ent + 2 * R, fent + alg->nnz_u + alg->nnz_v, sizeof(fmm_entry_t) * (size_t)alg->nnz_w);
    // part tables (unreferenced parts keep zero extents: nothing is read)
    fmm_entry_t* pa = ent + 2 * R + alg->nnz_w;
    fmm_entry_t* pb = pa + npa;
    for (long e = 0; e < alg->nnz_u; e++) pa[fent[e].part] = fent[e];
    for (long e = alg->nnz_u; e < alg->nnz_u + alg->nnz_v; e++) pb[fent[e].part] = fent[e];
}

int build_presum_list(const struct fmmhip_alg_s* alg, const fmm_geom_t& g, const DevList& fused, bool a_mn,
                      bool b_mn, DevList& out)
{
    const int R = alg->R;
    out.nent = presum_list_entries(alg);
    out.ptr_h = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(R + 1));
    out.ent_h = (fmm_entry_t*)calloc((size_t)out.nent + 1, sizeof(fmm_entry_t));
    if (!out.ptr_h || !out.ent_h) return FMMHIP_ERR_NOMEM;
    build_presum_list_into(alg, g, fused.ptr_h, fused.ent_h, a_mn, b_mn, out.ptr_h, out.ent_h);
    return FMMHIP_SUCCESS;
}

// Slot layout of the products-materialised mode: R leaves of lm x ln,
// contiguous along C's contiguous dimension (rows unless C is row-major).
struct SlotLayout {
    int swap;                    // 1: slots row-major (C row-major)
    int Lc, Ls;
    long ldm, slot;
};

SlotLayout slot_layout(const fmm_geom_t& g)
{
    SlotLayout sl;
    sl.swap = (g.csC == 1 && g.rsC != 1) ? 1 : 0;
    sl.Lc = (int)(sl.swap ? g.ln : g.lm);
    sl.Ls = (int)(sl.swap ? g.lm : g.ln);
    const bool pad = !sl.swap && tile_padded(g);
    sl.ldm = pad ? up_tile(sl.Lc) : sl.Lc;
    sl.slot = sl.ldm * (pad ? up_tile(sl.Ls) : sl.Ls);
    return sl;
}

// The products-materialised list: the pre-summed A and B leaves, one W
// entry per product (its slot, coefficient 1, first writer), then the W
// stage's tables: the mt*nt C parts (off, rows, cols; `first` = index of
// the part's first term, `part` = its term count) and the terms (slot
// offset, W coefficient), per part in ascending product order -- the order
// in which the fused epilogue updates that part.
int build_mat_list(const struct fmmhip_alg_s* alg, const fmm_geom_t& g, const DevList& fused, const DevList& presum,
                   DevList& out)
{
    const int R = alg->R;
    const int npc = alg->mt * alg->nt;
    const SlotLayout sl = slot_layout(g);
    out.nent = 3L * R + npc + alg->nnz_w;
    out.ptr_h = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(R + 1));
    out.ent_h = (fmm_entry_t*)calloc((size_t)out.nent + 1, sizeof(fmm_entry_t));
    if (!out.ptr_h || !out.ent_h) return FMMHIP_ERR_NOMEM;
    for (int r = 0; r <= R; r++) {
        out.ptr_h[r] = r;
        out.ptr_h[(R + 1) + r] = R + r;
        out.ptr_h[2 * (R + 1) + r] = 2 * R + r;
    }
    memcpy(out.ent_h, presum.ent_h, sizeof(fmm_entry_t) * 2 * (size_t)R);
    for (int r = 0; r < R; r++) {
        fmm_entry_t& ew = out.ent_h[2 * R + r];
        ew.off = (int64_t)r * sl.slot;
        ew.rows = (int32_t)g.lm;
        ew.cols = (int32_t)g.ln;
        ew.coef = 1.0;
        ew.first = -1;           // first writer in every group: a plain store
        ew.part = r;
    }
    fmm_entry_t* parts = out.ent_h + 3 * R;
    fmm_entry_t* terms = parts + npc;
    const fmm_entry_t* wf = fused.ent_h + alg->nnz_u + alg->nnz_v;
    const int32_t* wptr = fused.ptr_h + 2 * (R + 1);
    const long w0 = alg->nnz_u + alg->nnz_v;
    std::vector<int> count(npc, 0);
    for (long e = 0; e < alg->nnz_w; e++) {
        const int c = wf[e].part;
        if (c < 0 || c >= npc) return FMMHIP_ERR_INVALID;
        count[c]++;
        parts[c].off = wf[e].off;
        parts[c].rows = wf[e].rows;
        parts[c].cols = wf[e].cols;
    }
    int at = 0;
    for (int c = 0; c < npc; c++) {
        if (count[c] == 0) return FMMHIP_ERR_INVALID;   // a C part no product reaches
        parts[c].first = at;
        parts[c].part = 0;
        at += count[c];
    }
    for (int r = 0; r < R; r++)
        for (long e = wptr[r] - w0; e < wptr[r + 1] - w0; e++) {
            fmm_entry_t& pc = parts[wf[e].part];
            fmm_entry_t& t = terms[pc.first + pc.part++];
            t.off = (int64_t)r * sl.slot;
            t.coef = wf[e].coef;
            t.part = r;
            t.cols = r;          // slot index in the single all-parts group
        }
    return FMMHIP_SUCCESS;
}

// Returns the cached descriptor for geometry g (building and uploading the
// fused list, and the pre-summed list when asked, stream-ordered).
// Whether an operand's sums take the two-stage form: two-level algorithms
// whose composed part count exceeds what the one-pass part-staged kernel
// runs at full occupancy (64 parts; its LDS holds every part of a column
// of elements), with at most 16 parts per level.
constexpr int kKronMinParts = 65;

bool kron_operand(const struct fmmhip_alg_s* alg, bool is_a)
{
    if (alg->levels != 2 || !alg->lv_U[0] || !alg->lv_U[1]) return false;
    const int np = is_a ? alg->mt * alg->kt : alg->kt * alg->nt;
    const int np0 = is_a ? alg->lv_m[0] * alg->lv_k[0] : alg->lv_k[0] * alg->lv_n[0];
    const int np1 = is_a ? alg->lv_m[1] * alg->lv_k[1] : alg->lv_k[1] * alg->lv_n[1];
    return np >= kKronMinParts && np0 <= FMMHIP_STAGE_PARTS && np1 <= FMMHIP_STAGE_PARTS &&
           alg->lv_R[0] <= FMMHIP_STAGE_MAX_R && alg->lv_R[1] <= FMMHIP_STAGE_MAX_R;
}

// Host image of one operand's two-stage tables.  Composed part (row
// a0*rows1 + a1, col b0*cols1 + b1) with coefficient X0[(a0,b0)][r0] *
// X1[(a1,b1)][r1] for product r0*R1 + r1 (the composition of
// fmm_spec.c).  Stage 1, segment per outer block b = (a0,b0): the inner
// sums Y[b][r1] = sum_(a1,b1) X1[(a1,b1)][r1] part(b; a1,b1), written as
// leaf b*R1 + r1 of the stage-1 region.  Stage 2, segment per inner
// product r1: S[r0*R1 + r1] = sum_b X0[b][r0] Y[b][r1].
struct KronImage {
    std::vector<int32_t> tptr0, tptr1;       // CSR over products per level
    std::vector<fmmhip_stage_term_t> term0, term1;
    std::vector<fmm_entry_t> parts1, parts2;
    std::vector<fmmhip_stage_seg_t> seg1, seg2;
    long x1_elems = 0;
};

void kron_image(const struct fmmhip_alg_s* alg, const fmm_geom_t& g, bool is_a, const WsLayout& wl, KronImage& k)
{
    const int rows0 = is_a ? alg->lv_m[0] : alg->lv_k[0], cols0 = is_a ? alg->lv_k[0] : alg->lv_n[0];
    const int rows1 = is_a ? alg->lv_m[1] : alg->lv_k[1], cols1 = is_a ? alg->lv_k[1] : alg->lv_n[1];
    const int R0 = alg->lv_R[0], R1 = alg->lv_R[1];
    const int* X0 = is_a ? alg->lv_U[0] : alg->lv_V[0];
    const int* X1 = is_a ? alg->lv_U[1] : alg->lv_V[1];
    const int np0 = rows0 * cols0, np1 = rows1 * cols1;
    // nonzero terms of one level per product, ascending part (numbered row*cols + col)
    auto sparse = [](const int* X, int np, int R, std::vector<int32_t>& ptr, std::vector<fmmhip_stage_term_t>& tm) {
        ptr.assign((size_t)R + 1, 0);
        for (int r = 0; r < R; r++) {
            ptr[r] = (int32_t)tm.size();
            for (int p = 0; p < np; p++)
                if (X[(size_t)p * R + r]) {
                    fmmhip_stage_term_t t;
                    memset(&t, 0, sizeof(t));
                    t.coef = (double)X[(size_t)p * R + r];
                    t.part = p;
                    tm.push_back(t);
                }
        }
        ptr[R] = (int32_t)tm.size();
    };
    sparse(X0, np0, R0, k.tptr0, k.term0);
    sparse(X1, np1, R1, k.tptr1, k.term1);
    // composed part geometry: the operand's rows and columns split directly
    const long wr = is_a ? g.m : g.k, wc = is_a ? g.k : g.n;
    const int tr = rows0 * rows1, tc = cols0 * cols1;
    std::vector<long> szr(tr), ofr(tr), szc(tc), ofc(tc);
    fmm_part_split(wr, tr, szr.data(), ofr.data());
    fmm_part_split(wc, tc, szc.data(), ofc.data());
    const long rs = is_a ? g.rsA : g.rsB, cs = is_a ? g.csA : g.csB;
    const long plane = (long)wl.ldo * wl.Ls;
    const long lr = is_a ? g.lm : g.lk, lc = is_a ? g.lk : g.ln;   // leaf rows and columns
    for (int a0 = 0; a0 < rows0; a0++)
        for (int b0 = 0; b0 < cols0; b0++)
            for (int a1 = 0; a1 < rows1; a1++)
                for (int b1 = 0; b1 < cols1; b1++) {
                    const int I = a0 * rows1 + a1, P = b0 * cols1 + b1;
                    fmm_entry_t e;
                    memset(&e, 0, sizeof(e));
                    e.off = ofr[I] * rs + ofc[P] * cs;
                    e.rows = (int32_t)szr[I];
                    e.cols = (int32_t)szc[P];
                    k.parts1.push_back(e);           // block (a0,b0), part a1*cols1 + b1
                }
    for (int r1 = 0; r1 < R1; r1++)
        for (int b = 0; b < np0; b++) {
            fmm_entry_t e;
            memset(&e, 0, sizeof(e));
            e.off = ((long)b * R1 + r1) * plane;
            e.rows = (int32_t)lr;
            e.cols = (int32_t)lc;
            k.parts2.push_back(e);
        }
    for (int b = 0; b < np0; b++) {
        fmmhip_stage_seg_t sg;
        memset(&sg, 0, sizeof(sg));
        sg.s_off = (long)b * R1 * plane;
        sg.pstride = plane;
        sg.R = R1;
        sg.np = np1;
        k.seg1.push_back(sg);
    }
    for (int r1 = 0; r1 < R1; r1++) {
        fmmhip_stage_seg_t sg;
        memset(&sg, 0, sizeof(sg));
        sg.s_off = (long)r1 * plane;
        sg.pstride = (long)R1 * plane;
        sg.R = R0;
        sg.np = np0;
        k.seg2.push_back(sg);
    }
    k.x1_elems = (long)np0 * R1 * plane;
}

// Builds and uploads the two-stage tables of the operands that take them.
int build_kron(const struct fmmhip_alg_s* alg, const fmm_geom_t& g, bool a_mn, bool b_mn, DevDesc* d,
               hipStream_t stream)
{
    KronImage im[2];
    const bool use[2] = {kron_operand(alg, true), kron_operand(alg, false)};
    const WsLayout wl[2] = {ws_layout_a(g, a_mn), ws_layout_b(g, b_mn)};
    for (int o = 0; o < 2; o++)
        if (use[o]) kron_image(alg, g, o == 0, wl[o], im[o]);
    // layout of the one allocation: per operand tptr0, tptr1, term0, term1, parts1, parts2, seg1, seg2
    std::vector<char>& h = d->kron_host;
    h.clear();
    auto place = [&](const void* src, size_t bytes) {
        const size_t off = (h.size() + 255) & ~(size_t)255;
        h.resize(off + bytes);
        if (bytes) memcpy(h.data() + off, src, bytes);
        return off;
    };
    size_t off[2][8] = {};
    for (int o = 0; o < 2; o++) {
        if (!use[o]) continue;
        KronImage& k = im[o];
        off[o][0] = place(k.tptr0.data(), k.tptr0.size() * sizeof(int32_t));
        off[o][1] = place(k.tptr1.data(), k.tptr1.size() * sizeof(int32_t));
        off[o][2] = place(k.term0.data(), k.term0.size() * sizeof(fmmhip_stage_term_t));
        off[o][3] = place(k.term1.data(), k.term1.size() * sizeof(fmmhip_stage_term_t));
        off[o][4] = place(k.parts1.data(), k.parts1.size() * sizeof(fmm_entry_t));
        off[o][5] = place(k.parts2.data(), k.parts2.size() * sizeof(fmm_entry_t));
        off[o][6] = place(k.seg1.data(), k.seg1.size() * sizeof(fmmhip_stage_seg_t));
        off[o][7] = place(k.seg2.data(), k.seg2.size() * sizeof(fmmhip_stage_seg_t));
    }
    if (h.empty()) return FMMHIP_SUCCESS;
    if (hipMalloc(&d->kron_dev, h.size()) != hipSuccess) {
        (void)hipGetLastError();
        d->kron_dev = nullptr;
        return FMMHIP_ERR_NOMEM;
    }
    char* dev = (char*)d->kron_dev;
    KronView& v = d->kron;
    memset(&v, 0, sizeof(v));
    for (int o = 0; o < 2; o++) {
        if (!use[o]) continue;
        KronImage& k = im[o];
        fmmhip_stage_seg_t* s1 = (fmmhip_stage_seg_t*)(h.data() + off[o][6]);
        fmmhip_stage_seg_t* s2 = (fmmhip_stage_seg_t*)(h.data() + off[o][7]);
        const int np1 = (int)(k.parts1.size() / k.seg1.size());
        const int np0 = (int)k.seg1.size();
        for (size_t b = 0; b < k.seg1.size(); b++) {
            s1[b].tptr = (const int*)(dev + off[o][1]);
            s1[b].terms = (const fmmhip_stage_term_t*)(dev + off[o][3]);
            s1[b].parts = (const fmm_entry_t*)(dev + off[o][4]) + b * np1;
        }
        for (size_t r = 0; r < k.seg2.size(); r++) {
            s2[r].tptr = (const int*)(dev + off[o][0]);
            s2[r].terms = (const fmmhip_stage_term_t*)(dev + off[o][2]);
            s2[r].parts = (const fmm_entry_t*)(dev + off[o][5]) + r * np0;
        }
        const fmmhip_stage_seg_t* d1 = (const fmmhip_stage_seg_t*)(dev + off[o][6]);
        const fmmhip_stage_seg_t* d2 = (const fmmhip_stage_seg_t*)(dev + off[o][7]);
        if (o == 0) {
            v.useA = true; v.a1 = d1; v.a2 = d2; v.na1 = (int)k.seg1.size(); v.na2 = (int)k.seg2.size();
        } else {
            v.useB = true; v.b1 = d1; v.b2 = d2; v.nb1 = (int)k.seg1.size(); v.nb2 = (int)k.seg2.size();
        }
        v.x1_elems = std::max(v.x1_elems, k.x1_elems);
    }
    if (hipMemcpyAsync(d->kron_dev, h.data(), h.size(), hipMemcpyHostToDevice, stream) != hipSuccess)
        return FMMHIP_ERR_HIP;
    return FMMHIP_SUCCESS;
}

// Two-stage W stage of the products-materialised mode (two-level
// algorithms): stage 1 forms, for every outer product r0 and inner C part
// c1, Y[r0][c1] = sum_r1 W1[c1][r1] M[r0*R1 + r1] (a full leaf, r1
// ascending); stage 2 forms composed C part (c0, c1) := beta C + alpha
// sum_r0 W0[c0][r0] Y[r0][c1] (r0 ascending).  Taken when it moves fewer
// leaves than the one-pass stage: R0*nnz(W1) slot reads + R0*np1 Y writes
// + nnz(W0)*np1 Y reads against nnz(W) slot reads (C's read and write are
// common to both).
bool wkron_pays(const struct fmmhip_alg_s* alg)
{
    if (alg->levels != 2 || !alg->lv_W[0] || !alg->lv_W[1]) return false;
    long nnz[2] = {0, 0};
    for (int l = 0; l < 2; l++) {
        const long cells = (long)alg->lv_m[l] * alg->lv_n[l] * alg->lv_R[l];
        for (long i = 0; i < cells; i++) nnz[l] += alg->lv_W[l][i] != 0;
    }
    const long np1 = (long)alg->lv_m[1] * alg->lv_n[1];
    const long two = (long)alg->lv_R[0] * nnz[1] + (long)alg->lv_R[0] * np1 + nnz[0] * np1;
    return two < alg->nnz_w;
}

// Host image and upload of the two-stage W tables, from the materialised
// list's C part table (off, rows, cols).
int build_wkron(const struct fmmhip_alg_s* alg, const fmm_geom_t& g, DevDesc* d, hipStream_t stream)
{
    const int R = alg->R, R0 = alg->lv_R[0], R1 = alg->lv_R[1];
    const int rows0 = alg->lv_m[0], cols0 = alg->lv_n[0], rows1 = alg->lv_m[1], cols1 = alg->lv_n[1];
    const int np0 = rows0 * cols0, np1 = rows1 * cols1, npc = alg->mt * alg->nt, nt = alg->nt;
    const int* W0 = alg->lv_W[0];
    const int* W1 = alg->lv_W[1];
    const long slot = slot_layout(g).slot;
    const fmm_entry_t* cparts = d->mat.ent_h + 3 * R;
    std::vector<fmm_entry_t> p1((size_t)R0 * np1), t1, p2(cparts, cparts + npc), t2;
    for (int r0 = 0; r0 < R0; r0++)
        for (int c1 = 0; c1 < np1; c1++) {
            fmm_entry_t& pe = p1[(size_t)r0 * np1 + c1];
            memset(&pe, 0, sizeof(pe));
            pe.off = ((long)r0 * np1 + c1) * slot;
            pe.rows = (int32_t)g.lm;
            pe.cols = (int32_t)g.ln;
            pe.first = (int32_t)t1.size();
            for (int r1 = 0; r1 < R1; r1++) {
                const int w = W1[(size_t)c1 * R1 + r1];
                if (!w) continue;
                fmm_entry_t t;
                memset(&t, 0, sizeof(t));
                t.off = ((long)r0 * R1 + r1) * slot;
                t.coef = (double)w;
                t.part = r0 * R1 + r1;
                t.cols = r1;                     // slot index in group r0
                t1.push_back(t);
            }
            pe.part = (int32_t)t1.size() - pe.first;
        }
    for (int a0 = 0; a0 < rows0; a0++)
        for (int b0 = 0; b0 < cols0; b0++)
            for (int a1 = 0; a1 < rows1; a1++)
                for (int b1 = 0; b1 < cols1; b1++) {
                    const int c0 = a0 * cols0 + b0, c1 = a1 * cols1 + b1;
                    fmm_entry_t& pe = p2[(size_t)(a0 * rows1 + a1) * nt + (b0 * cols1 + b1)];
                    pe.first = (int32_t)t2.size();
                    for (int r0 = 0; r0 < R0; r0++) {
                        const int w = W0[(size_t)c0 * R0 + r0];
                        if (!w) continue;
                        fmm_entry_t t;
                        memset(&t, 0, sizeof(t));
                        t.off = ((long)r0 * np1 + c1) * slot;
                        t.coef = (double)w;
                        t.part = r0;
                        t.cols = r0;             // slot index in group c1
                        t2.push_back(t);
                    }
                    pe.part = (int32_t)t2.size() - pe.first;
                }
    // all-parts groups: stage 1 group r0 = Y[r0][0..np1) over slots
    // r0*R1 .. r0*R1+R1-1; stage 2 group c1 = C parts (c0, c1) over Y[0..R0)[c1]
    std::vector<int32_t> g1((size_t)R0 * np1), g2((size_t)np1 * np0);
    for (int r0 = 0; r0 < R0; r0++)
        for (int c1 = 0; c1 < np1; c1++) g1[(size_t)r0 * np1 + c1] = r0 * np1 + c1;
    for (int a0 = 0; a0 < rows0; a0++)
        for (int b0 = 0; b0 < cols0; b0++)
            for (int a1 = 0; a1 < rows1; a1++)
                for (int b1 = 0; b1 < cols1; b1++)
                    g2[(size_t)(a1 * cols1 + b1) * np0 + (a0 * cols0 + b0)] =
                        (a0 * rows1 + a1) * nt + (b0 * cols1 + b1);
    std::vector<char>& h = d->wk_host;
    h.clear();
    auto place = [&](const void* src, size_t bytes) {
        const size_t off = (h.size() + 255) & ~(size_t)255;
        h.resize(off + bytes);
        if (bytes) memcpy(h.data() + off, src, bytes);
        return off;
    };
    const size_t o1 = place(p1.data(), p1.size() * sizeof(fmm_entry_t));
    const size_t o2 = place(t1.data(), t1.size() * sizeof(fmm_entry_t));
    const size_t o3 = place(p2.data(), p2.size() * sizeof(fmm_entry_t));
    const size_t o4 = place(t2.data(), t2.size() * sizeof(fmm_entry_t));
    const size_t o5 = place(g1.data(), g1.size() * sizeof(int32_t));
    const size_t o6 = place(g2.data(), g2.size() * sizeof(int32_t));
    if (hipMalloc(&d->wk_dev, h.size()) != hipSuccess) {
        (void)hipGetLastError();
        d->wk_dev = nullptr;
        return FMMHIP_ERR_NOMEM;
    }
    char* dev = (char*)d->wk_dev;
    WKronView& v = d->wk;
    v.parts1 = (const fmm_entry_t*)(dev + o1);
    v.terms1 = (const fmm_entry_t*)(dev + o2);
    v.parts2 = (const fmm_entry_t*)(dev + o3);
    v.terms2 = (const fmm_entry_t*)(dev + o4);
    v.np1 = R0 * np1;
    v.np2 = npc;
    v.y_leaves = (long)R0 * np1;
    v.groups1 = (const int32_t*)(dev + o5);
    v.groups2 = (const int32_t*)(dev + o6);
    v.ngroups1 = R0; v.gparts1 = np1; v.nslots1 = R1;
    v.ngroups2 = np1; v.gparts2 = np0; v.nslots2 = R0;
    if (hipMemcpyAsync(d->wk_dev, h.data(), h.size(), hipMemcpyHostToDevice, stream) != hipSuccess)
        return FMMHIP_ERR_HIP;
    return FMMHIP_SUCCESS;
}

int get_desc(fmmhip_alg_t alg, const fmm_geom_t& g, bool want_presum, bool want_mat, bool a_mn, bool b_mn,
             hipStream_t stream, DevDesc** out)
{
    if (!alg->dev_cache) {
        static std::mutex create_mu;
        std::lock_guard<std::mutex> lk(create_mu);
        if (!alg->dev_cache) alg->dev_cache = new DevCache();
    }
    DevCache* cache = (DevCache*)alg->dev_cache;
    std::lock_guard<std::mutex> lk(cache->mu);
    cache->clock++;
    DevDesc* d = nullptr;
    for (DevDesc* c : cache->items)
        if (same_geo(c->geo, g)) { d = c; break; }
    if (!d) {
        if (!alg->keep_all && cache->items.size() >= kCacheCap) {
            // evict the least recently used descriptor; kernels may still read it
            size_t victim = 0;
            for (size_t i = 1; i < cache->items.size(); i++)
                if (cache->items[i]->last_use < cache->items[victim]->last_use) victim = i;
            if (hipDeviceSynchronize() != hipSuccess) return FMMHIP_ERR_HIP;
            free_desc(cache->items[victim]);
            cache->items.erase(cache->items.begin() + victim);
        }
        d = new DevDesc();
        d->geo = g;
        int rc = fmm_build_entries(alg, &g, &d->fused.ptr_h, &d->fused.ent_h, &d->fused.nent);
        if (!rc) rc = upload_list(d->fused, alg->R, stream);
        if (rc) { free_desc(d); return rc; }
        cache->items.push_back(d);
    }
    if (want_presum && !d->has_presum) {
        int rc = build_presum_list(alg, g, d->fused, a_mn, b_mn, d->presum);
        if (!rc) rc = upload_list(d->presum, alg->R, stream);
        if (rc) { free_list(d->presum); return rc; }
        d->has_presum = true;
    }
    if (want_presum && !d->has_kron && (kron_operand(alg, true) || kron_operand(alg, false))) {
        int rc = build_kron(alg, g, a_mn, b_mn, d, stream);
        if (rc) {
            if (d->kron_dev) (void)hipFree(d->kron_dev);
            d->kron_dev = nullptr;
            memset(&d->kron, 0, sizeof(d->kron));
            return rc;
        }
        d->has_kron = true;
    }
    if (want_mat && !d->has_mat) {
        int rc = build_mat_list(alg, g, d->fused, d->presum, d->mat);
        if (!rc) rc = upload_list(d->mat, alg->R, stream);
        if (rc) { free_list(d->mat); return rc; }
        d->has_mat = true;
    }
    if (want_mat && !d->has_wk && wkron_pays(alg)) {
        int rc = build_wkron(alg, g, d, stream);
        if (rc) {
            if (d->wk_dev) (void)hipFree(d->wk_dev);
            d->wk_dev = nullptr;
            memset(&d->wk, 0, sizeof(d->wk));
            return rc;
        }
        d->has_wk = true;
    }
    d->last_use = cache->clock;
    *out = d;
    return FMMHIP_SUCCESS;
}

// Per-stream workspace for the pre-summed mode (grow-only; a stream's
// kernels are ordered, so one buffer per stream is race-free).
// Tag 0: the call's workspace; 1, 2: the gathered copies of general-strided
// A and B.
struct Workspace {
    hipStream_t stream;
    int tag;
    void* p;
    size_t cap;
};
std::mutex g_ws_mu;
std::vector<Workspace> g_ws;

void* workspace_get(hipStream_t st, size_t bytes, int tag = 0)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    for (Workspace& w : g_ws) {
        if (w.stream != st || w.tag != tag) continue;
        if (w.cap >= bytes) return w.p;
        if (hipStreamSynchronize(st) != hipSuccess) return nullptr;
        (void)hipFree(w.p);
        w.p = nullptr;
        w.cap = 0;
        if (hipMalloc(&w.p, bytes) != hipSuccess) { (void)hipGetLastError(); w.p = nullptr; return nullptr; }
        w.cap = bytes;
        return w.p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    g_ws.push_back(Workspace{st, tag, p, bytes});
    return p;
}


// Pairs along the contiguous dimension can be loaded as one 16-byte vector
// when the base, the leading dimension and every part offset keep them
// 16-byte aligned (pairs start at even indices of the tile).
bool pairs_aligned(const void* base, size_t esz, long ld, const fmm_entry_t* ent, long e0, long e1)
{
    const size_t pair = 2 * esz;
    if (((uintptr_t)base) % pair) return false;
    if (ld % 2) return false;
    for (long e = e0; e < e1; e++)
        if (ent[e].off % 2) return false;
    return true;
}

// Everything one call needs, resolved on the host: the kernel arguments,
// the launch shape and, in the pre-summed mode, the two summation segments.
// Workspace pointers (segment S and the kernel's A/B) are offsets from the
// workspace base until bind_workspace().
// Optional per-launch timing (fmmhip_timing_*): HIP events recorded on the
// launch stream around the summation stage and around the MFMA kernel.
struct TimedCall {
    hipEvent_t e0, e1, e2;      // before summation, before kernel, after kernel
    hipEvent_t e3;              // after the split reduction (counted with the summation), or null
    bool presum;
};
std::mutex g_tm_mu;
bool g_tm_on = false;
std::vector<TimedCall> g_tm;

void tm_record(hipEvent_t* ev, hipStream_t st)
{
    if (hipEventCreate(ev) == hipSuccess) (void)hipEventRecord(*ev, st);
    else *ev = nullptr;
}

// Largest part count of the part-staged summation kernel.
constexpr int kPartStagedMax = 256;


}  // namespace

// Groups per tile (split of the products, or of the k-steps of a
// classical call) for a grid of `tiles` workgroups (max_groups: the products,
// or the k-steps / 8; peak: the
// leaf's MFMA peak in FLOP/s, 0 = FP64 or FP32 by `fp64`; wg_per_cu: the
// kernel's resident workgroups per CU; c_bytes: C's extent in bytes, gaps
// included).  Time model per choice G: the leaf flops at a sustained MFMA
// rate scaled by the fraction of the last wave of workgroup slots that is
// filled, plus, for G > 1, the temporaries' traffic at HBM speed (G copies
// of C zeroed, read back by the reduction, C read and written once).  G is
// a power of two <= min(8, max_groups); ties go to fewer groups.
int split_groups(int requested, int max_groups, long tiles, double leaf_flops, double c_bytes, bool fp64,
                 double peak, int wg_per_cu)
{
    if (max_groups < 2 || requested == 1) return 1;
    if (requested > 1) {
        int g = std::min(requested, 8);
        while (g > max_groups) g /= 2;
        return g;
    }
    int cus = 256;
    fmmhip_device_cu_count(&cus);
    const double slots = (double)wg_per_cu * cus;
    const double rate = (peak > 0.0 ? peak : fp64 ? 78.6e12 : 157.3e12) * 0.85;   // sustained MFMA FLOP/s
    const double hbm = 4.0e12;                                // achieved bytes/s of the fill and reduction
    int best = 1;
    double best_t = 0.0;
    for (int g = 1; g <= 8 && g <= max_groups; g *= 2) {
        const double items = (double)tiles * g;
        const double fill = items / (std::ceil(items / slots) * slots);
        double t = leaf_flops / (rate * fill);
        if (g > 1) t += c_bytes * (2.0 * g + 2.0) / hbm;
        if (g == 1 || t < 0.98 * best_t) { best = g; best_t = t; }
    }
    return best;
}

// Groups per tile of the products-materialised mode (no copies of C to
// pay for): the G <= min(R, 64) that minimises the waves of workgroups
// times the products of the longest group, ceil(tiles G / slots)
// ceil(R / G); ties go to more groups.  Blocks are group-major, so with
// G = R the blocks of one product run together and share its S_r row and
// T_r column panels in L2, as a classical GEMM's tiles share A and B (L2
// hit rate 0.50 with one group, 0.77 with 49, at m = n = k = 16384, 222 L2).
// Whether a products-materialised launch of `groups` groups of `tiles`
// tiles deals whole groups to the XCDs: several groups, each of no more
// tiles than an XCD holds workgroups at once (its CUs' share of the
// resident slots).
int xcd_groups(int groups, long tiles, int wg_per_cu, int cus)
{
    return groups > 1 && tiles <= (long)wg_per_cu * cus / 8 ? 1 : 0;
}

int mat_groups(int R, long tiles, int slots)
{
    int best = 1;
    double best_t = 0.0;
    for (int g = 1; g <= std::min(R, 64); g++) {
        const double t = std::ceil((double)tiles * g / slots) * std::ceil((double)R / g);
        if (g == 1 || t <= best_t) { best = g; best_t = t; }
    }
    return best;
}

// Whether the pre-summed mode (W epilogue fused, no product in HBM) is
// expected to be at least as fast as the products-materialised one.  Time
// model (both share the summation stage): the leaf
// flops at a sustained MFMA rate divided by the wave fill of each mode's
// grid (pre-summed: the split its own model picks, copies of C charged;
// materialised: mat_groups, no copies), plus, pre-summed, the fused W
// epilogue's read-modify-write of one C leaf per W entry at its measured
// effective rate (6.8 TB/s, partly hidden behind the MFMAs), or,
// materialised, the W stage (one slot read per W entry, C read and
// written once) at 4.9 TB/s.
bool presum_wins(fmmhip_alg_t alg, Kind kind, const fmm_geom_t& g, int split, size_t esz, double c_span)
{
    int cus = 256;
    fmmhip_device_cu_count(&cus);
    const int slots = 2 * cus;
    const long tiles = ((g.lm + BM - 1) / BM) * ((g.ln + BN - 1) / BN);
    const int R = alg->R;
    const double leaf_flops = 2.0 * R * (double)g.lm * (double)g.lk * (double)g.ln;
    const double rate = (kind == Kind::D64 ? 78.6e12 : 157.3e12) * 0.87;
    const double leaf_c = (double)esz * (double)g.lm * (double)g.ln;
    const int g1 = split_groups(split, R, tiles, leaf_flops, (double)esz * c_span, kind == Kind::D64, 0.0, 2);
    const double items1 = (double)tiles * g1;
    const double fill1 = items1 / (std::ceil(items1 / slots) * slots);
    double t1 = leaf_flops / (rate * fill1) + (double)alg->nnz_w * leaf_c * 2.0 / 6.8e12;
    if (g1 > 1) t1 += (double)esz * c_span * (2.0 * g1 + 2.0) / 4.0e12;
    const int g2 = split > 0 ? std::min(split, R) : mat_groups(R, tiles, slots);
    const double fill2 = ((double)tiles * R / slots) /
                         (std::ceil((double)tiles * g2 / slots) * std::ceil((double)R / g2));
    // the slot writes hide behind the MFMAs of a long k loop: counted in
    // the share min(1, 512 / lk), fitted to FP64 sweeps at 13824^3 and
    // 13824^2 x 1152 (every algorithm, 1 and 2 levels) and to
    // 16384^3 222 L2 (materialised kernel 99.2 ms, fused-epilogue 105.2 ms)
    const double slot_share = std::min(1.0, 512.0 / (double)std::max(g.lk, 1L));
    const double t2 = leaf_flops / (rate * std::min(fill2, 1.0)) +
                      (slot_share * (double)R * leaf_c + (double)alg->nnz_w * leaf_c + 2.0 * (double)esz * c_span) /
                          4.9e12;
    return t1 <= t2;
}

// The pack mode of one call: 0 fused, 1 pre-summed, 2 pre-summed with the
// products materialised.  Asked for (pack >= 0; materialising needs
// several products and the one-problem path, whose lists it builds), or
// automatic: pre-summed for multi-part algorithms when its workspace fits
// the bound (half of the device's total memory: the choice, and with it the
// rounding of the result, does not depend on what is free at the time of
// the call), materialised as well when the slots fit too and the time
// model expects it to be faster (presum_wins).  prepare_call checks the
// chosen mode's whole workspace against the bound afterwards.
int resolve_pack(fmmhip_alg_t alg, Kind kind, const fmm_geom_t& g, int pack, int split, size_t esz, size_t ws_bytes,
                 double c_span, bool one_problem)
{
    // the slots' leading dimension is a kernel stride (32-bit tile offsets)
    const bool can_mat = alg->R > 1 && one_problem && g.lm <= (1L << 23) && g.ln <= (1L << 23);
    if (pack >= 0) return pack == 2 ? (can_mat ? 2 : 1) : pack;
    if (alg->max_nnz_u <= 1 && alg->max_nnz_v <= 1) return 0;
    const double bound = fmmhip_ws_bound();
    if (!((double)ws_bytes < bound)) return 0;
    if (!can_mat) return 1;
    const size_t slots = esz * (size_t)g.lm * (size_t)g.ln * (size_t)alg->R;
    if (!((double)(ws_bytes + slots) < bound)) return 1;
    return presum_wins(alg, kind, g, split, esz, c_span) ? 1 : 2;
}

// The fast kernel (shape 3) for calls that may take it: single-part
// operands whose prepared tiles all lie inside their parts (automatic shape,
// or shape 3 asked for); otherwise the prepared shape stays.
void pick_fast(Prepared* P, const struct fmmhip_alg_s* alg, Kind kind, bool single_part)
{
    const bool want = alg->shape < 0 || alg->shape == 3;
    const bool ok = single_part && (kind == Kind::D64 ? fmmk::fast_eligible<double>(P->ka, P->a_mn, P->b_mn, P->c_col)
                                                      : fmmk::fast_eligible<float>(P->ka, P->a_mn, P->b_mn, P->c_col));
    if ((P->shape == 0 || P->shape == 3) && want && ok) P->shape = 3;
    else if (P->shape == 3) P->shape = 0;
}

int prepare_once(fmmhip_alg_t alg, Kind kind, long m, long n, long k, double alpha, const void* A, long rsA,
                 long csA, const void* B, long rsB, long csB, double beta, void* C, long rsC, long csC, int pack,
                 int split, hipStream_t st, Prepared* P, DescSource* src)
{
    memset(P, 0, sizeof(*P));
    P->kind = kind;
    P->stream = st;
    if (!alg || m < 0 || n < 0 || k < 0) return FMMHIP_ERR_INVALID;
    if (m == 0 || n == 0) { P->nothing = true; return FMMHIP_SUCCESS; }
    if (!C) return FMMHIP_ERR_INVALID;
    P->C = C; P->m = m; P->n = n; P->rsC = rsC; P->csC = csC; P->beta = beta;
    if (k == 0 || alpha == 0.0) {
        P->scale_only = true;
        return FMMHIP_SUCCESS;
    }
    if (!A || !B) return FMMHIP_ERR_INVALID;
    // operand layouts: one of the strides must be 1; when both are (a
    // dimension of 1), the column-major reading is taken
    bool a_mn, b_mn;
    if (rsA == 1) a_mn = true; else if (csA == 1) a_mn = false; else return FMMHIP_ERR_UNSUPPORTED;
    if (rsB == 1) b_mn = false; else if (csB == 1) b_mn = true; else return FMMHIP_ERR_UNSUPPORTED;
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;
    // per-thread tile offsets are 32-bit: (tile extent) x (stride) must fit
    const long kStrideMax = 1L << 23;
    if ((a_mn ? csA : rsA) > kStrideMax || (b_mn ? rsB : csB) > kStrideMax || rsC > kStrideMax ||
        csC > kStrideMax || rsC < 0 || csC < 0)
        return FMMHIP_ERR_UNSUPPORTED;

    fmm_geom_t g;
    memset(&g, 0, sizeof(g));
    g.m = m; g.n = n; g.k = k;
    g.lm = (m + alg->mt - 1) / alg->mt;
    g.lk = (k + alg->kt - 1) / alg->kt;
    g.ln = (n + alg->nt - 1) / alg->nt;
    g.rsA = rsA; g.csA = csA; g.rsB = rsB; g.csB = csB; g.rsC = rsC; g.csC = csC;
    const size_t esz = kind == Kind::S32 ? 4 : 8;
    const bool multi = alg->max_nnz_u > 1 || alg->max_nnz_v > 1;
    // FP32-leaf kind with column-major operands whose sums run the one-pass
    // part-staged kernel: the pre-summed leaves may be FP32 (each FP64 sum
    // rounded once, as the kernel rounds it while staging an FP64 leaf), so
    // the kernel and the summation move half the leaf bytes; taken below
    // where the call runs on the fast kernel (the general kernel measured
    // slower over FP32 leaves: 222 L2 at 10000^3 17.7 -> 19.5 ms).  The
    // workspace keeps its FP64 size, so the pack mode does not depend on it.
    const bool f32_possible = kind == Kind::D_F32LEAF && a_mn && !b_mn && rsC == 1 &&
                              !kron_operand(alg, true) && !kron_operand(alg, false) &&
                              (long)alg->mt * alg->kt <= kPartStagedMax && (long)alg->kt * alg->nt <= kPartStagedMax;

    // pack mode: pre-summed when asked, or (automatic) for multi-part
    // algorithms when the workspace fits in half of the free device memory
    const WsLayout wa = ws_layout_a(g, a_mn), wb = ws_layout_b(g, b_mn);
    const size_t bytesA = esz * wa.elems(alg->R);
    const size_t offB = (bytesA + 255) & ~(size_t)255;
    const size_t ws_bytes = offB + esz * wb.elems(alg->R) + 256;
    const int mode = resolve_pack(alg, kind, g, pack, split, esz, ws_bytes,
                                  (double)(m - 1) * rsC + (double)(n - 1) * csC + 1.0, !src || src->can_mat());
    const bool wmat = mode == 2;
    const bool presum = mode >= 1;
    P->mode = mode;

    DescView dv;
    memset(&dv, 0, sizeof(dv));
    int rc;
    if (src) {
        rc = src->get(alg, g, presum, wmat, a_mn, b_mn, st, &dv);
    } else {
        DevDesc* d = nullptr;
        rc = get_desc(alg, g, presum, wmat, a_mn, b_mn, st, &d);
        if (!rc) {
            dv.fused_ptr = d->fused.dptr();
            dv.fused_ent = d->fused.dent();
            dv.fused_ent_h = d->fused.ent_h;
            if (presum) {
                dv.presum_ptr = d->presum.dptr();
                dv.presum_ent = d->presum.dent();
            }
            if (wmat) {
                dv.mat_ptr = d->mat.dptr();
                dv.mat_ent = d->mat.dent();
                dv.wk = d->wk;
            }
            if (presum) dv.kron = d->kron;
        }
    }
    if (rc) return rc;

    KArgs& ka = P->ka;
    ka.C = C;
    ka.rsC = rsC; ka.csC = csC;
    ka.lm = (int)g.lm; ka.lk = (int)g.lk; ka.ln = (int)g.ln;
    ka.min_cm = (int)(m / alg->mt); ka.min_cn = (int)(n / alg->nt);
    ka.R = alg->R;
    ka.alpha = alpha;
    ka.beta = beta;
    int shape = alg->shape;
    if (presum) {
        P->presum = true;
        P->ws_bytes = ws_bytes;
        const int R = alg->R;
        fmmhip_presum_seg_t& sa = P->segA;
        sa.X = A; sa.S = (void*)(uintptr_t)0; sa.ld = a_mn ? csA : rsA;
        sa.ptr = dv.fused_ptr; sa.ent = dv.fused_ent;
        sa.R = R; sa.swap = a_mn ? 0 : 1; sa.Lc = wa.Lc; sa.Ls = wa.Ls; sa.ldo = wa.ldo;
        sa.np = alg->mt * alg->kt;
        sa.parts = dv.presum_ent + 2 * R + alg->nnz_w;
        fmmhip_presum_seg_t& sb = P->segB;
        sb.X = B; sb.S = (void*)(uintptr_t)offB; sb.ld = b_mn ? rsB : csB;
        sb.ptr = dv.fused_ptr + (R + 1); sb.ent = dv.fused_ent;
        sb.R = R; sb.swap = b_mn ? 1 : 0; sb.Lc = wb.Lc; sb.Ls = wb.Ls; sb.ldo = wb.ldo;
        sb.np = alg->kt * alg->nt;
        sb.parts = sa.parts + sa.np;
        ka.A = (const void*)(uintptr_t)0;
        ka.B = (const void*)(uintptr_t)offB;
        ka.lda = wa.ldo;
        ka.ldb = wb.ldo;
        ka.min_m = (int)g.lm; ka.min_k = (int)g.lk; ka.min_n = (int)g.ln;   // full leaves
        ka.ptr = dv.presum_ptr;
        ka.ent = dv.presum_ent;
        ka.va = ka.vb = 1;           // workspace leaves are 16-byte aligned with even ldo
        if (shape < 0) shape = 0;    // one part per product
        if (dv.kron.useA || dv.kron.useB) {
            // stage-1 leaves of the two-stage summation, after the B leaves
            P->kron = dv.kron;
            P->x1_off = (P->ws_bytes + 255) & ~(size_t)255;
            P->ws_bytes = P->x1_off + esz * (size_t)dv.kron.x1_elems;
        }
    } else {
        ka.A = A;
        ka.B = B;
        ka.lda = a_mn ? csA : rsA;
        ka.ldb = b_mn ? rsB : csB;
        ka.min_m = (int)(m / alg->mt); ka.min_k = (int)(k / alg->kt); ka.min_n = (int)(n / alg->nt);
        ka.ptr = dv.fused_ptr;
        ka.ent = dv.fused_ent;
        ka.va = pairs_aligned(A, esz, ka.lda, dv.fused_ent_h, 0, alg->nnz_u);
        ka.vb = pairs_aligned(B, esz, ka.ldb, dv.fused_ent_h, alg->nnz_u, alg->nnz_u + alg->nnz_v);
        if (shape < 0) shape = multi ? 1 : 0;
    }
    // the 64 x 64 tile shape is instantiated for column-major A, B, C only
    if (shape == 2 && !(a_mn && !b_mn && rsC == 1)) shape = 0;
    // shape 3, the fast kernel, takes single-part operands only
    const bool single_part = presum || !multi;
    if (shape == 3 && !single_part) shape = 0;
    P->shape = shape;
    P->a_mn = a_mn;
    P->b_mn = b_mn;
    P->c_col = rsC == 1;
    P->single_part = single_part;
    const int T = fmmk::tile_size(shape);
    const int wg_per_cu = shape == 2 ? 4 : 2;      // resident workgroups per CU (LDS and registers)
    ka.tiles_m = (int)((g.lm + T - 1) / T);
    ka.tiles_n = (int)((g.ln + T - 1) / T);
    const long tiles = (long)ka.tiles_m * ka.tiles_n;
    // split: group g of every tile scatters into temporary g (the
    // temporaries span C's whole extent, gaps between columns included);
    // the groups take contiguous product ranges, or for the classical
    // algorithm contiguous k-step ranges of at least 8 steps
    const double c_span = (double)(m - 1) * rsC + (double)(n - 1) * csC + 1.0;
    // classical, column-major A, B, C, and at most half the CUs with a tile:
    // a lone workgroup per CU runs at about 0.9 of two (measured), so a grid
    // of one tile per CU gains too little to pay for the temporaries
    int cus = 256;
    fmmhip_device_cu_count(&cus);
    const bool kspl = alg->R == 1 && a_mn && !b_mn && rsC == 1 && 2 * tiles <= cus && shape != 2;
    const long ksteps = (g.lk + fmmk::BK - 1) / fmmk::BK;
    if (wmat) {
        // every product writes its own slot: the groups need no copies of C
        // and are chosen for the wave fill alone
        const SlotLayout sl = slot_layout(g);
        P->wmat = true;
        P->m_off = (P->ws_bytes + 255) & ~(size_t)255;
        P->ws_bytes = P->m_off + esz * (size_t)sl.slot * (size_t)alg->R;
        ka.C = (void*)(uintptr_t)P->m_off;
        ka.rsC = sl.swap ? sl.ldm : 1;
        ka.csC = sl.swap ? 1 : sl.ldm;
        ka.min_cm = (int)g.lm; ka.min_cn = (int)g.ln;
        if (tile_padded(g) && a_mn && !b_mn && rsC == 1 && T == fmmk::BM) {
            // whole tiles and k-steps over the padded leaves and slots (same
            // tile grid; the extra k-steps multiply zeros)
            ka.lm = ka.min_m = ka.min_cm = (int)up_tile(g.lm);
            ka.ln = ka.min_n = ka.min_cn = (int)up_tile(g.ln);
            ka.lk = ka.min_k = (int)up_kstep(g.lk);
        }
        ka.alpha = 1.0;
        ka.beta = 0.0;
        ka.ptr = dv.mat_ptr;
        ka.ent = dv.mat_ent;
        P->c_col = !sl.swap;
        ka.groups = split > 0 ? std::min(split, alg->R) : mat_groups(alg->R, tiles, wg_per_cu * cus);
        ka.ksplit = false;
        ka.gbit = 0;
        ka.xcd_groups = xcd_groups(ka.groups, tiles, wg_per_cu, cus);
        ka.tmp = ka.C;           // group g's Cb: the slots (no per-group copy)
        ka.tmp_elems = 0;
        fmmhip_wsum_t& w = P->wsum;
        w.M = ka.C;
        w.ldm = sl.ldm;
        w.swap = sl.swap;
        w.Lc = sl.Lc; w.Ls = sl.Ls;
        w.C = C;
        w.rsC = rsC; w.csC = csC;
        w.alpha = alpha;
        w.beta = beta;
        w.np = alg->mt * alg->nt;
        w.parts = dv.mat_ent + 3 * alg->R;
        w.terms = w.parts + w.np;
        w.nslots = alg->R;
        w.ngroups = 1;
        w.gparts = w.np;
        w.groups = nullptr;
        if (dv.wk.parts1) {
            // two-stage W: stage 1 into the Y leaves (after the slots), stage 2 into C
            P->w2 = true;
            P->y_off = (P->ws_bytes + 255) & ~(size_t)255;
            P->ws_bytes = P->y_off + esz * (size_t)sl.slot * (size_t)dv.wk.y_leaves;
            fmmhip_wsum_t& w2 = P->wsum2;
            w2 = w;                              // stage 2: from the Y leaves into C
            w2.M = (void*)(uintptr_t)P->y_off;
            w2.np = dv.wk.np2;
            w2.parts = dv.wk.parts2;
            w2.terms = dv.wk.terms2;
            w.C = (void*)(uintptr_t)P->y_off;    // stage 1: from the slots into the Y leaves
            w.rsC = ka.rsC;
            w.csC = ka.csC;
            w.alpha = 1.0;
            w.beta = 0.0;
            w.np = dv.wk.np1;
            w.parts = dv.wk.parts1;
            w.terms = dv.wk.terms1;
            w.nslots = dv.wk.nslots1;            // all-parts groups for both stages
            w.ngroups = dv.wk.ngroups1;
            w.gparts = dv.wk.gparts1;
            w.groups = dv.wk.groups1;
            w2.nslots = dv.wk.nslots2;
            w2.ngroups = dv.wk.ngroups2;
            w2.gparts = dv.wk.gparts2;
            w2.groups = dv.wk.groups2;
        }
        P->nwg = tiles * ka.groups;
        if (P->nwg > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;
        if (tiles <= 0) P->nothing = true;
        pick_fast(P, alg, kind, single_part);
        P->f32_leaves = f32_possible && P->shape == 3;
        return FMMHIP_SUCCESS;
    }
    ka.groups = split_groups(split, alg->R == 1 ? (kspl ? (int)std::min(ksteps / 8, 8L) : 1) : alg->R, tiles,
                             2.0 * alg->R * (double)g.lm * (double)g.lk * (double)g.ln, (double)esz * c_span,
                             kind == Kind::D64, 0.0, wg_per_cu);
    ka.ksplit = kspl && ka.groups > 1;
    ka.gbit = ka.groups == 8 ? 3 : ka.groups == 4 ? 2 : ka.groups == 2 ? 1 : 0;
    // groups scattering into copies of C keep the tile order that spreads a
    // group over the XCDs (whole groups per XCD measured slower there: 444
    // L2 at 13824^3 with the W epilogue fused 22.4 -> 16.8 TFLOPS)
    ka.xcd_groups = 0;
    ka.tmp = nullptr;
    ka.tmp_elems = 0;
    if (ka.groups > 1) {
        ka.tmp_elems = ((m - 1) * rsC + (n - 1) * csC + 1 + 31) & ~31L;
        const size_t base = P->presum ? ((P->ws_bytes + 255) & ~(size_t)255) : 0;
        P->tmp_off = base;
        P->ws_bytes = base + esz * (size_t)ka.tmp_elems * (size_t)ka.groups;
        ka.tmp = (void*)(uintptr_t)base;
    }
    P->nwg = tiles * ka.groups;
    if (P->nwg > INT32_MAX) return FMMHIP_ERR_UNSUPPORTED;
    if (tiles <= 0) P->nothing = true;
    pick_fast(P, alg, kind, single_part);
    P->f32_leaves = presum && f32_possible && P->shape == 3;
    return FMMHIP_SUCCESS;
}

int prepare_call(fmmhip_alg_t alg, Kind kind, long m, long n, long k, double alpha, const void* A, long rsA,
                 long csA, const void* B, long rsB, long csB, double beta, void* C, long rsC, long csC, int pack,
                 int split, hipStream_t st, Prepared* P, DescSource* src)
{
    int rc = prepare_once(alg, kind, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, pack, split, st, P,
                          src);
    // automatic mode: the chosen mode's whole workspace (summation leaves,
    // two-stage leaves, slots, Y leaves, split temporaries) must fit the
    // bound, else the next simpler mode is taken
    while (!rc && pack < 0 && P->mode > 0 && (double)P->ws_bytes > fmmhip_ws_bound())
        rc = prepare_once(alg, kind, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, P->mode - 1, split,
                          st, P, src);
    return rc;
}

void bind_workspace(Prepared* P, void* ws)
{
    char* base = (char*)ws;
    if (P->kron.useA || P->kron.useB) P->x1 = base + P->x1_off;
    if (P->wmat) {
        P->ka.C = base + P->m_off;
        P->ka.tmp = P->ka.C;
        P->wsum.M = P->ka.C;
        if (P->w2) {
            P->wsum.C = base + P->y_off;
            P->wsum2.M = base + P->y_off;
        }
    } else if (P->ka.groups > 1) {
        P->ka.tmp = base + P->tmp_off;
    }
    if (!P->presum) return;
    P->segA.S = base + (uintptr_t)P->segA.S;
    P->segB.S = base + (uintptr_t)P->segB.S;
    P->ka.A = P->segA.S;
    P->ka.B = P->segB.S;
}

// Launches the summation stage (if any) and the MFMA kernel of a prepared,
// workspace-bound call on its stream.
int launch_prepared(const Prepared* P, int stages)
{
    hipStream_t st = P->stream;
    if (P->nothing) return FMMHIP_SUCCESS;
    if (P->scale_only) {
        if (!(stages & 2)) return FMMHIP_SUCCESS;
        if (P->kind == Kind::S32) return fmmhip_scale_s((float*)P->C, P->m, P->n, P->rsC, P->csC, (float)P->beta, st);
        return fmmhip_scale_d((double*)P->C, P->m, P->n, P->rsC, P->csC, P->beta, st);
    }
    bool timed = false;
    if (stages == 3 && !P->untimed) {
        std::lock_guard<std::mutex> lk(g_tm_mu);
        timed = g_tm_on;
    }
    TimedCall tc{nullptr, nullptr, nullptr, nullptr, P->presum};
    if (timed) tm_record(&tc.e0, st);
    int rc = FMMHIP_SUCCESS;
    const bool dbl = P->kind != Kind::S32;
    const bool split = P->ka.groups > 1 && !P->wmat;
    if (split && (stages & 2)) {
        rc = fmmhip_split_zero(dbl, P->ka.tmp, P->ka.tmp_elems, P->ka.groups, st);
        if (rc) return rc;
    }
    if (P->presum && (stages & 1)) {
        // per operand: the two-stage form, the part-staged one-pass form, or
        // the entry-driven one (part counts beyond the part-staged kernel's)
        auto sum_operand = [&](const fmmhip_presum_seg_t& sg, bool two_stage, const fmmhip_stage_seg_t* s1, int n1,
                               const fmmhip_stage_seg_t* s2, int n2) -> int {
            if (two_stage) {
                int r = fmmhip_presum_stage(dbl, sg.X, sg.ld, P->x1, sg.swap, sg.Lc, sg.Ls, sg.ldo, s1, n1, st);
                if (!r) r = fmmhip_presum_stage(dbl, P->x1, sg.ldo, sg.S, sg.swap, sg.Lc, sg.Ls, sg.ldo, s2, n2, st);
                return r;
            }
            if (sg.parts && sg.np <= kPartStagedMax)
                return P->f32_leaves ? fmmhip_presum_seg_d2s(&sg, st) : fmmhip_presum_seg(dbl, &sg, st);
            if (P->kind == Kind::S32)
                return fmmhip_presum_s((const float*)sg.X, sg.ld, sg.ptr, sg.ent, sg.R, sg.swap, sg.Lc, sg.Ls, sg.ldo,
                                       (float*)sg.S, st);
            return fmmhip_presum_d((const double*)sg.X, sg.ld, sg.ptr, sg.ent, sg.R, sg.swap, sg.Lc, sg.Ls, sg.ldo,
                                   (double*)sg.S, st);
        };
        const KronView& kv = P->kron;
        rc = sum_operand(P->segA, kv.useA, kv.a1, kv.na1, kv.a2, kv.na2);
        if (!rc) rc = sum_operand(P->segB, kv.useB, kv.b1, kv.nb1, kv.b2, kv.nb2);
        if (rc) return rc;
    }
    if (timed) tm_record(&tc.e1, st);
    if (!(stages & 2)) return FMMHIP_SUCCESS;
    switch (P->kind) {
    case Kind::D64: rc = fmmk_launch_d64(P->ka, P->shape, P->a_mn, P->b_mn, P->c_col, (int)P->nwg, st); break;
    case Kind::D_F32LEAF:
        rc = P->f32_leaves ? fmmk_launch_f32leaves_d(P->ka, P->shape, (int)P->nwg, st)
                           : fmmk_launch_d_f32leaf(P->ka, P->shape, P->a_mn, P->b_mn, P->c_col, (int)P->nwg, st);
        break;
    case Kind::S32: rc = fmmk_launch_s32(P->ka, P->shape, P->a_mn, P->b_mn, P->c_col, (int)P->nwg, st); break;
    }
    if (timed) tm_record(&tc.e2, st);
    if (!rc && split)
        rc = fmmhip_split_reduce(dbl, P->C, P->m, P->n, P->rsC, P->csC, P->beta, P->ka.tmp, P->ka.tmp_elems,
                                 P->ka.groups, st);
    if (!rc && P->wmat) rc = fmmhip_wsum(dbl, &P->wsum, st);
    if (!rc && P->w2) rc = fmmhip_wsum(dbl, &P->wsum2, st);
    if (timed) {
        if (split || P->wmat) tm_record(&tc.e3, st);
        std::lock_guard<std::mutex> lk(g_tm_mu);
        g_tm.push_back(tc);
    }
    return rc;
}

static int gemm_common(fmmhip_alg_t alg, Kind kind, long m, long n, long k, double alpha,
                const void* A, long rsA, long csA, const void* B, long rsB, long csB, double beta,
                void* C, long rsC, long csC, void* stream)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)fmmhip_default_stream();
    Prepared P;
    if (!alg) return FMMHIP_ERR_INVALID;
    // from the workspace lookup to the last launch: a call from another host
    // thread on the same stream cannot regrow the workspace in between
    std::unique_lock<std::mutex> lk = enqueue_lock();
    // general-strided operands (neither stride 1, e.g. a strided sub-view):
    // gathered column-major into the stream's gather buffers by one
    // bandwidth pass, and the call proceeds on the copies
    const size_t esz = kind == Kind::S32 ? 4 : 8;
    const bool work = m > 0 && n > 0 && k > 0 && alpha != 0.0 && A && B;
    if (work && rsA != 1 && csA != 1) {
        if (rsA < 0 || csA < 0) return FMMHIP_ERR_UNSUPPORTED;
        void* g = workspace_get(st, esz * (size_t)m * (size_t)k, 1);
        if (!g) return FMMHIP_ERR_NOMEM;
        if ((rc = fmmhip_gather(kind != Kind::S32, A, m, k, rsA, csA, g, st))) return rc;
        A = g; rsA = 1; csA = m;
    }
    if (work && rsB != 1 && csB != 1) {
        if (rsB < 0 || csB < 0) return FMMHIP_ERR_UNSUPPORTED;
        void* g = workspace_get(st, esz * (size_t)k * (size_t)n, 2);
        if (!g) return FMMHIP_ERR_NOMEM;
        if ((rc = fmmhip_gather(kind != Kind::S32, B, k, n, rsB, csB, g, st))) return rc;
        B = g; rsB = 1; csB = k;
    }
    const int pack = alg->pack;
    rc = prepare_call(alg, kind, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, pack, alg->split, st,
                      &P);
    if (rc) return rc;
    if (P.ws_bytes) {
        void* ws = workspace_get(st, P.ws_bytes);
        if (!ws) {
            if (pack >= 1 || alg->split > 1) return FMMHIP_ERR_NOMEM;
            // automatic choices fall back to the fused mode without split
            // (reported by fmmhip_alg_last_pack)
            rc = prepare_call(alg, kind, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, 0, 1, st, &P);
            if (rc) return rc;
        } else {
            bind_workspace(&P, ws);
        }
    }
    rc = launch_prepared(&P);
    if (!rc && !P.nothing && !P.scale_only) {
        alg->last_pack = P.mode;
        alg->last_shape = P.shape;
    }
    return rc;
}

int fused_desc(fmmhip_alg_t alg, const fmm_geom_t& g, hipStream_t st, const int** dptr, const fmm_entry_t** dent)
{
    DevDesc* d = nullptr;
    int rc = get_desc(alg, g, false, false, true, false, st, &d);
    if (rc) return rc;
    *dptr = d->fused.dptr();
    *dent = d->fused.dent();
    return FMMHIP_SUCCESS;
}

int mat_view(fmmhip_alg_t alg, const fmm_geom_t& g, hipStream_t st, MatView* v)
{
    DevDesc* d = nullptr;
    int rc = get_desc(alg, g, true, true, g.rsA == 1, g.rsB != 1, st, &d);
    if (rc) return rc;
    const int R = alg->R;
    const SlotLayout sl = slot_layout(g);
    v->wptr = d->mat.dptr() + 2 * (R + 1);
    v->ent = d->mat.dent();
    v->np = alg->mt * alg->nt;
    v->parts = d->mat.dent() + 3 * R;
    v->terms = v->parts + v->np;
    v->swap = sl.swap;
    v->Lc = sl.Lc;
    v->Ls = sl.Ls;
    v->ldm = sl.ldm;
    v->slot = sl.slot;
    return FMMHIP_SUCCESS;
}

void* workspace_for(hipStream_t st, size_t bytes) { return workspace_get(st, bytes); }

std::mutex g_enqueue_mu;
std::unique_lock<std::mutex> enqueue_lock() { return std::unique_lock<std::mutex>(g_enqueue_mu); }

namespace {
size_t list_bytes(int R, long nent)
{
    return ((sizeof(int32_t) * 3 * (size_t)(R + 1) + 255) & ~(size_t)255) + sizeof(fmm_entry_t) * (size_t)(nent + 1);
}

// Copies a host list into `host` at byte offset `off` in the device layout
// of upload_list (CSR, then entries at the next 256-byte boundary).
size_t pack_list(const DevList& l, int R, char* host, size_t off)
{
    const size_t ptr_bytes = sizeof(int32_t) * 3 * (size_t)(R + 1);
    const size_t ent_off = (ptr_bytes + 255) & ~(size_t)255;
    memcpy(host + off, l.ptr_h, ptr_bytes);
    memcpy(host + off + ent_off, l.ent_h, sizeof(fmm_entry_t) * (size_t)l.nent);
    return ent_off;
}
}  // namespace

size_t desc_bound(fmmhip_alg_t alg)
{
    const long fused = alg->nnz_u + alg->nnz_v + alg->nnz_w;
    const long pres = 2L * alg->R + alg->nnz_w + alg->mt * alg->kt + alg->kt * alg->nt;
    const long mat = 3L * alg->R + alg->mt * alg->nt + alg->nnz_w;
    return ((list_bytes(alg->R, fused) + 255) & ~(size_t)255) + ((list_bytes(alg->R, pres) + 255) & ~(size_t)255) +
           list_bytes(alg->R, mat) + 256;
}

int desc_pack(fmmhip_alg_t alg, const fmm_geom_t& g, bool presum, bool mat, bool a_mn, bool b_mn, void* host,
              const void* dev, size_t cap, size_t* used, DescView* v)
{
    DevList fused, pres, matl;
    mat = mat && presum;
    int rc = fmm_build_entries(alg, &g, &fused.ptr_h, &fused.ent_h, &fused.nent);
    if (!rc && presum) rc = build_presum_list(alg, g, fused, a_mn, b_mn, pres);
    if (!rc && mat) rc = build_mat_list(alg, g, fused, pres, matl);
    const size_t off2 = (list_bytes(alg->R, fused.nent) + 255) & ~(size_t)255;
    const size_t off3 = presum ? off2 + ((list_bytes(alg->R, pres.nent) + 255) & ~(size_t)255) : off2;
    size_t need = presum ? off2 + list_bytes(alg->R, pres.nent) : off2;
    if (mat) need = off3 + list_bytes(alg->R, matl.nent);
    if (!rc && need > cap) rc = FMMHIP_ERR_INVALID;
    if (!rc) {
        char* h = (char*)host;
        const char* d = (const char*)dev;
        const size_t e1 = pack_list(fused, alg->R, h, 0);
        const size_t off2 = (list_bytes(alg->R, fused.nent) + 255) & ~(size_t)255;
        v->fused_ptr = (const int*)d;
        v->fused_ent = (const fmm_entry_t*)(d + e1);
        v->fused_ent_h = (const fmm_entry_t*)(h + e1);
        if (presum) {
            const size_t e2 = pack_list(pres, alg->R, h, off2);
            v->presum_ptr = (const int*)(d + off2);
            v->presum_ent = (const fmm_entry_t*)(d + off2 + e2);
        } else {
            v->presum_ptr = nullptr;
            v->presum_ent = nullptr;
        }
        if (mat) {
            const size_t e3 = pack_list(matl, alg->R, h, off3);
            v->mat_ptr = (const int*)(d + off3);
            v->mat_ent = (const fmm_entry_t*)(d + off3 + e3);
        } else {
            v->mat_ptr = nullptr;
            v->mat_ent = nullptr;
        }
        *used = need;
    }
    free(fused.ptr_h); free(fused.ent_h);
    free(pres.ptr_h); free(pres.ent_h);
    free(matl.ptr_h); free(matl.ent_h);
    return rc;
}

int desc_pack_presum_into(fmmhip_alg_t alg, const fmm_geom_t& g, bool a_mn, bool b_mn, void* host, const void* dev,
                          size_t cap, size_t* used, DescView* v)
{
    const int R = alg->R;
    const long nf = alg->nnz_u + alg->nnz_v + alg->nnz_w, np = presum_list_entries(alg);
    const size_t ptr_bytes = sizeof(int32_t) * 3 * (size_t)(R + 1);
    const size_t ent_off = (ptr_bytes + 255) & ~(size_t)255;
    const size_t off2 = (list_bytes(R, nf) + 255) & ~(size_t)255;
    const size_t need = off2 + list_bytes(R, np);
    const size_t sizes_off = (need + 255) & ~(size_t)255;    // work space of the builder, not uploaded
    const size_t total = sizes_off + sizeof(long) * 2 * (size_t)(alg->mt + alg->kt + alg->nt);
    if (total > cap) return FMMHIP_ERR_INVALID;
    char* h = (char*)host;
    const char* d = (const char*)dev;
    memset(h, 0, need);
    int32_t* fptr = (int32_t*)h;
    fmm_entry_t* fent = (fmm_entry_t*)(h + ent_off);
    fmm_build_entries_into(alg, &g, fptr, fent, (long*)(h + sizes_off));
    int32_t* pptr = (int32_t*)(h + off2);
    fmm_entry_t* pent = (fmm_entry_t*)(h + off2 + ent_off);
    build_presum_list_into(alg, g, fptr, fent, a_mn, b_mn, pptr, pent);
    memset(v, 0, sizeof(*v));
    v->fused_ptr = (const int*)d;
    v->fused_ent = (const fmm_entry_t*)(d + ent_off);
    v->fused_ent_h = fent;
    v->presum_ptr = (const int*)(d + off2);
    v->presum_ent = (const fmm_entry_t*)(d + off2 + ent_off);
    *used = need;
    return FMMHIP_SUCCESS;
}

bool timing_active()
{
    std::lock_guard<std::mutex> lk(g_tm_mu);
    return g_tm_on;
}

void timing_record(hipEvent_t* ev, hipStream_t st) { tm_record(ev, st); }

void timing_push(hipEvent_t e0, hipEvent_t e1, hipEvent_t e2, bool presum)
{
    std::lock_guard<std::mutex> lk(g_tm_mu);
    g_tm.push_back(TimedCall{e0, e1, e2, nullptr, presum});
}

fmmhip_alg_t classical_alg()
{
    static std::once_flag once;
    static fmmhip_alg_t alg = nullptr;
    std::call_once(once, [] { fmmhip_alg_create_classical(&alg); });
    return alg;
}


extern "C" int fmmhip_alg_set_shape(fmmhip_alg_t alg, int shape)
{
    if (!alg || shape < -1 || shape > 3) return FMMHIP_ERR_INVALID;
    alg->shape = shape;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_alg_set_split(fmmhip_alg_t alg, int groups)
{
    if (!alg || !(groups == -1 || groups == 1 || groups == 2 || groups == 4 || groups == 8)) return FMMHIP_ERR_INVALID;
    alg->split = groups;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_alg_set_pack(fmmhip_alg_t alg, int pack)
{
    if (!alg || pack < -1 || pack > 2) return FMMHIP_ERR_INVALID;
    alg->pack = pack;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_alg_pack_for(fmmhip_alg_t alg, fmmhip_leaf_t leaf, long m, long n, long k, int* pack)
{
    if (!alg || !pack || m < 0 || n < 0 || k < 0) return FMMHIP_ERR_INVALID;
    int rc = fmmhip_init();
    if (rc) return rc;
    if (leaf == FMMHIP_LEAF_BF16X3 || leaf == FMMHIP_LEAF_BF16X3_6 || leaf == FMMHIP_LEAF_BF16X3_8 ||
        leaf == FMMHIP_LEAF_INT8X4) {
        // planes (slices) always staged; products materialised as
        // bf16x3_gemm and int8x4_gemm decide
        bool wmat = alg->R > 1 && (alg->pack == 2 || alg->pack < 0);
        if (wmat && alg->pack < 0) {
            const double lm = (double)((m + alg->mt - 1) / alg->mt), ln = (double)((n + alg->nt - 1) / alg->nt);
            wmat = 8.0 * alg->R * lm * ln < fmmhip_ws_bound();
        }
        *pack = wmat ? 2 : 1;
        return FMMHIP_SUCCESS;
    }
    const Kind kind = leaf == FMMHIP_LEAF_F32 ? Kind::D_F32LEAF : Kind::D64;
    if (m == 0 || n == 0 || k == 0) { *pack = 0; return FMMHIP_SUCCESS; }
    // the call's own resolution (whole workspace against the bound) on
    // placeholder, 256-byte aligned operand addresses: nothing is launched
    const void* ph = (const void*)(uintptr_t)4096;
    Prepared P;
    std::unique_lock<std::mutex> lk = enqueue_lock();
    rc = prepare_call(alg, kind, m, n, k, 1.0, ph, 1, std::max(m, 1L), ph, 1, std::max(k, 1L), 1.0, (void*)ph, 1,
                      std::max(m, 1L), alg->pack, alg->split, (hipStream_t)fmmhip_default_stream(), &P);
    if (rc) return rc;
    *pack = P.mode;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_alg_last_pack(fmmhip_alg_t alg, int* pack)
{
    if (!alg || !pack) return FMMHIP_ERR_INVALID;
    std::unique_lock<std::mutex> lk = enqueue_lock();
    *pack = alg->last_pack;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_alg_last_shape(fmmhip_alg_t alg, int* shape)
{
    if (!alg || !shape) return FMMHIP_ERR_INVALID;
    std::unique_lock<std::mutex> lk = enqueue_lock();
    *shape = alg->last_shape;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_timing_enable(int on)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_tm_mu);
    g_tm_on = on != 0;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_timing_read(long* launches, double* kernel_ms, double* presum_ms, double* max_kernel_ms)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_tm_mu);
    long n = 0;
    double k_ms = 0.0, p_ms = 0.0, k_max = 0.0;
    for (TimedCall& c : g_tm) {
        float a = 0.f, b = 0.f, d = 0.f;
        hipEvent_t last = c.e3 ? c.e3 : c.e2;
        if (c.e0 && c.e1 && c.e2 && hipEventSynchronize(last) == hipSuccess &&
            hipEventElapsedTime(&a, c.e0, c.e1) == hipSuccess && hipEventElapsedTime(&b, c.e1, c.e2) == hipSuccess &&
            (!c.e3 || hipEventElapsedTime(&d, c.e2, c.e3) == hipSuccess)) {
            n++;
            p_ms += a + d;
            k_ms += b;
            if (b > k_max) k_max = b;
        }
        if (c.e0) (void)hipEventDestroy(c.e0);
        if (c.e1) (void)hipEventDestroy(c.e1);
        if (c.e2) (void)hipEventDestroy(c.e2);
        if (c.e3) (void)hipEventDestroy(c.e3);
    }
    g_tm.clear();
    if (launches) *launches = n;
    if (kernel_ms) *kernel_ms = k_ms;
    if (presum_ms) *presum_ms = p_ms;
    if (max_kernel_ms) *max_kernel_ms = k_max;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_workspace_release(void)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    if (hipDeviceSynchronize() != hipSuccess) return FMMHIP_ERR_HIP;
    for (Workspace& w : g_ws) (void)hipFree(w.p);
    g_ws.clear();
    return FMMHIP_SUCCESS;
}

extern "C" void fmmhip_alg_release_cache(struct fmmhip_alg_s* alg)
{
    if (!alg || !alg->dev_cache) return;
    DevCache* cache = (DevCache*)alg->dev_cache;
    if (!cache->items.empty()) (void)hipDeviceSynchronize();
    for (DevDesc* d : cache->items) free_desc(d);
    delete cache;
    alg->dev_cache = nullptr;
}

extern "C" int fmmhip_dgemm_fmm(fmmhip_alg_t alg, fmmhip_leaf_t leaf, long m, long n, long k,
                                double alpha, const double* A, long rsA, long csA, const double* B,
                                long rsB, long csB, double beta, double* C, long rsC, long csC,
                                void* stream)
{
    Kind kind;
    switch (leaf) {
    case FMMHIP_LEAF_NATIVE: kind = Kind::D64; break;
    case FMMHIP_LEAF_F32: kind = Kind::D_F32LEAF; break;
    case FMMHIP_LEAF_BF16X3:
    case FMMHIP_LEAF_BF16X3_6:
    case FMMHIP_LEAF_BF16X3_8: {
        int rc = fmmhip_init();
        if (rc) return rc;
        hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)fmmhip_default_stream();
        // plane products: ten (four planes), six or eight (three planes)
        const int terms = leaf == FMMHIP_LEAF_BF16X3 ? 10 : leaf == FMMHIP_LEAF_BF16X3_6 ? 6 : 8;
        return bf16x3_gemm(alg, terms, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, st);
    }
    case FMMHIP_LEAF_INT8X4: {
        int rc = fmmhip_init();
        if (rc) return rc;
        hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)fmmhip_default_stream();
        return int8x4_gemm(alg, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, st);
    }
    default: return FMMHIP_ERR_UNSUPPORTED;
    }
    return gemm_common(alg, kind, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, stream);
}

extern "C" int fmmhip_sgemm_fmm(fmmhip_alg_t alg, long m, long n, long k, float alpha, const float* A,
                                long rsA, long csA, const float* B, long rsB, long csB, float beta,
                                float* C, long rsC, long csC, void* stream)
{
    return gemm_common(alg, Kind::S32, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, stream);
}

extern "C" int fmmhip_dgemm(long m, long n, long k, double alpha, const double* A, long rsA, long csA,
                            const double* B, long rsB, long csB, double beta, double* C, long rsC,
                            long csC, void* stream)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    return gemm_common(classical_alg(), Kind::D64, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C,
                       rsC, csC, stream);
}

extern "C" int fmmhip_sgemm(long m, long n, long k, float alpha, const float* A, long rsA, long csA,
                            const float* B, long rsB, long csB, float beta, float* C, long rsC,
                            long csC, void* stream)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    return gemm_common(classical_alg(), Kind::S32, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C,
                       rsC, csC, stream);
}

// ---- the pre-summed mode in two stages (the plugin registry's kernels) ----
//
// Stage 1 (the pack kernel, plugin/ref_kernels/bli_packm_fmm_ref.c) forms
// every S_r and T_r of A and B into the caller's workspace; stage 2 (the gemm
// kernel, plugin/ref_kernels/bli_gemm_fmm_ref.c) runs the R leaf products
// from it with the W-weighted epilogue into C.  Both resolve the call alike
// (pre-summed, no product split), so stage 1 then stage 2 on one stream is
// the one-call pre-summed product bit for bit.
static int staged_common(int stage, fmmhip_alg_t alg, Kind kind, long m, long n, long k, double alpha,
                         const void* A, long rsA, long csA, const void* B, long rsB, long csB, double beta, void* C,
                         long rsC, long csC, void* ws, size_t ws_bytes, size_t* need, void* stream)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    if (!alg) return FMMHIP_ERR_INVALID;
    hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)fmmhip_default_stream();
    std::unique_lock<std::mutex> lk = enqueue_lock();
    Prepared P;
    rc = prepare_call(alg, kind, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, 1, 1, st, &P);
    if (rc) return rc;
    if (need) { *need = P.ws_bytes; return FMMHIP_SUCCESS; }
    if (P.ws_bytes) {
        if (!ws || ws_bytes < P.ws_bytes) return FMMHIP_ERR_INVALID;
        bind_workspace(&P, ws);
    }
    return launch_prepared(&P, stage);
}

extern "C" int fmmhip_fmm_workspace_bytes(fmmhip_alg_t alg, int is_double, long m, long n, long k, long rsA,
                                          long csA, long rsB, long csB, long rsC, long csC, size_t* bytes)
{
    if (!bytes) return FMMHIP_ERR_INVALID;
    const void* ph = (const void*)(uintptr_t)4096;   // placeholder, aligned: nothing is launched
    return staged_common(0, alg, is_double ? Kind::D64 : Kind::S32, m, n, k, 1.0, ph, rsA, csA, ph, rsB, csB, 1.0,
                         (void*)ph, rsC, csC, nullptr, 0, bytes, nullptr);
}

extern "C" int fmmhip_dpack_fmm(fmmhip_alg_t alg, long m, long n, long k, double alpha, const double* A, long rsA,
                                long csA, const double* B, long rsB, long csB, double beta, double* C, long rsC,
                                long csC, void* ws, size_t ws_bytes, void* stream)
{
    return staged_common(1, alg, Kind::D64, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, ws,
                         ws_bytes, nullptr, stream);
}

extern "C" int fmmhip_dgemm_fmm_packed(fmmhip_alg_t alg, long m, long n, long k, double alpha, const double* A,
                                       long rsA, long csA, const double* B, long rsB, long csB, double beta,
                                       double* C, long rsC, long csC, void* ws, size_t ws_bytes, void* stream)
{
    return staged_common(2, alg, Kind::D64, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, ws,
                         ws_bytes, nullptr, stream);
}

extern "C" int fmmhip_spack_fmm(fmmhip_alg_t alg, long m, long n, long k, float alpha, const float* A, long rsA,
                                long csA, const float* B, long rsB, long csB, float beta, float* C, long rsC,
                                long csC, void* ws, size_t ws_bytes, void* stream)
{
    return staged_common(1, alg, Kind::S32, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, ws,
                         ws_bytes, nullptr, stream);
}

extern "C" int fmmhip_sgemm_fmm_packed(fmmhip_alg_t alg, long m, long n, long k, float alpha, const float* A,
                                       long rsA, long csA, const float* B, long rsB, long csB, float beta, float* C,
                                       long rsC, long csC, void* ws, size_t ws_bytes, void* stream)
{
    return staged_common(2, alg, Kind::S32, m, n, k, alpha, A, rsA, csA, B, rsB, csB, beta, C, rsC, csC, ws,
                         ws_bytes, nullptr, stream);
}
