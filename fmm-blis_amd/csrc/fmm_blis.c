/*
 * fmm_blis.c -- the fmm_blis drop-in: plugin registration and the object-API
 * driver over BLIS obj_t operands, dispatching to the native fmmhip_* layer.
 *
 * Reference counterparts (dnparikh/fmm-blis):
 *   bli_plugin_register_fmm_blis  plugin/bli_plugin_register.c:43-64
 *   bli_strassen_ab_ex            frame/base/bli_strassen_ab.c:123-295
 *   bli_strassen_ab               frame/base/bli_strassen_ab.c:297-300
 * What the driver keeps from the reference: the operand objects and their
 * meaning (C := beta*C + alpha*A*B, C is m x n, A m x k), the early return of
 * bli_l3_return_early_if_trivial (zero dims; k == 0 or alpha == 0 scale C),
 * the uneven part geometry, and its synchronous return.  What it does not
 * inherit: alpha applied twice and beta once per part per product
 * (bli_gemm_fmm_ref.c:85-119), R fixed at 7 (bli_strassen_ab.c:248), the
 * column-stored-C-only transposition (bli_strassen_ab.c:211-246), and the
 * non-thread-safe registration guard (bli_strassen_ab.c:125-137).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fmm_internal.h"

siz_t FMM_BLIS_PACK_UKR = 0;
siz_t FMM_BLIS_GEMM_UKR = 0;

/* ---- kernel registry ---------------------------------------------------- */

/* The counterpart of the BLIS gks kernel slots and the per-configuration
 * context (blis/frame/base/bli_gks.c bli_gks_register_ukr,
 * bli_cntx_set_ukr / bli_cntx_get_ukr_dt): slot ids are handed out in
 * order, and the one configuration (the gfx950 device) holds, per slot, a
 * kernel per real datatype (index 0 float, 1 double; complex slots stay
 * empty, as the library has no complex kernels). */
#define FMM_UKR_SLOTS 8
static void* g_ukr[FMM_UKR_SLOTS][2];
static siz_t g_ukr_count = 0;
static pthread_mutex_t g_ukr_mu = PTHREAD_MUTEX_INITIALIZER;

static err_t register_ukr(siz_t* id)
{
    pthread_mutex_lock(&g_ukr_mu);
    err_t e = BLIS_FAILURE;
    if (g_ukr_count + 1 < FMM_UKR_SLOTS) { *id = ++g_ukr_count; e = BLIS_SUCCESS; }
    pthread_mutex_unlock(&g_ukr_mu);
    return e;
}

static void set_ukr(siz_t id, void* s_fn, void* d_fn)
{
    pthread_mutex_lock(&g_ukr_mu);
    if (id > 0 && id < FMM_UKR_SLOTS) { g_ukr[id][0] = s_fn; g_ukr[id][1] = d_fn; }
    pthread_mutex_unlock(&g_ukr_mu);
}

void* fmm_blis_get_ukr_dt(num_t dt, siz_t ukr_id)
{
    if (bli_plugin_register_fmm_blis() != BLIS_SUCCESS) return NULL;
    if (ukr_id == 0 || ukr_id >= FMM_UKR_SLOTS) return NULL;
    const int t = dt == BLIS_DOUBLE ? 1 : dt == BLIS_FLOAT ? 0 : -1;
    if (t < 0) return NULL;
    pthread_mutex_lock(&g_ukr_mu);
    void* f = g_ukr[ukr_id][t];
    pthread_mutex_unlock(&g_ukr_mu);
    return f;
}

/* Per-configuration init (plugin/ref_kernels/bli_plugin_init_ref.c:61-71):
 * the pack slot gets the summation stage, the gemm slot the MFMA leaf
 * products with the W epilogue, for float and double. */
void bli_plugin_init_fmm_blis_gfx950_ref(void)
{
    set_ukr(FMM_BLIS_PACK_UKR, (void*)fmmhip_spack_fmm, (void*)fmmhip_dpack_fmm);
    set_ukr(FMM_BLIS_GEMM_UKR, (void*)fmmhip_sgemm_fmm_packed, (void*)fmmhip_dgemm_fmm_packed);
}

/* ---- registration ------------------------------------------------------- */

static pthread_once_t g_reg_once = PTHREAD_ONCE_INIT;
static err_t g_reg_err = BLIS_FAILURE;

static void do_register(void)
{
    if (fmmhip_init() != FMMHIP_SUCCESS) { g_reg_err = BLIS_FAILURE; return; }
    /* two kernel slots, then the context init of every enabled
     * configuration (bli_plugin_register.c:47-61); here one, the device */
    err_t e = register_ukr(&FMM_BLIS_PACK_UKR);
    if (e == BLIS_SUCCESS) e = register_ukr(&FMM_BLIS_GEMM_UKR);
    if (e != BLIS_SUCCESS) { g_reg_err = e; return; }
    bli_plugin_init_fmm_blis_gfx950_ref();
    g_reg_err = BLIS_SUCCESS;
}

err_t bli_plugin_register_fmm_blis(void)
{
    pthread_once(&g_reg_once, do_register);
    return g_reg_err;
}

/* ---- error policy ------------------------------------------------------- */

static int g_abort_on_error = 1;
static __thread int g_last_error = 0;

void fmm_blis_amd_set_abort_on_error(int on) { g_abort_on_error = on ? 1 : 0; }
int fmm_blis_amd_last_error(void) { return g_last_error; }

static void fail(const char* where, int status, const char* what)
{
    g_last_error = status ? status : FMMHIP_ERR_INVALID;
    fprintf(stderr, "libfmm_blis_amd: %s: %s (%s)\n", where, what, fmmhip_status_string(g_last_error));
    if (g_abort_on_error) abort();
}

/* ---- obj_t access (blis/frame/include/bli_obj_macro_defs.h semantics) --- */

typedef struct {
    num_t dt;
    long m, n;              /* dims of op(X) */
    long rs, cs;            /* strides of op(X) */
    char* p;                /* element (0,0) of op(X) */
    size_t esz;
} view_t;

static num_t obj_dt(const obj_t* o) { return o->info & FMM_OBJ_DT_MASK; }

static void obj_view(const obj_t* o, view_t* v)
{
    v->dt = obj_dt(o);
    v->esz = (size_t)o->elem_size;
    v->p = (char*)o->buffer + (long)o->elem_size * (o->off[0] * o->rs + o->off[1] * o->cs);
    if (o->info & FMM_OBJ_TRANS_BIT) {
        v->m = o->dim[1]; v->n = o->dim[0]; v->rs = o->cs; v->cs = o->rs;
    } else {
        v->m = o->dim[0]; v->n = o->dim[1]; v->rs = o->rs; v->cs = o->cs;
    }
}

/* Real value of a scalar object (BLIS_CONSTANT objects hold a constdata_t:
 * float s at offset 0, double d at offset 8). */
static int obj_scalar_value(const obj_t* o, double* out)
{
    num_t dt = obj_dt(o);
    if (dt == BLIS_CONSTANT) { *out = *(const double*)((const char*)o->buffer + 8); return 0; }
    const char* p = (const char*)o->buffer + (long)o->elem_size * (o->off[0] * o->rs + o->off[1] * o->cs);
    switch (dt) {
    case BLIS_DOUBLE: case BLIS_DCOMPLEX: *out = *(const double*)p; return 0;
    case BLIS_FLOAT: case BLIS_SCOMPLEX: *out = (double)*(const float*)p; return 0;
    default: return -1;
    }
}

/* The object's attached (internal) scalar; BLIS initialises it to 1. */
static double obj_internal_scalar(const obj_t* o)
{
    num_t sdt = o->info2 & 0x7u;
    if (sdt == BLIS_DOUBLE || sdt == BLIS_DCOMPLEX) { double d; memcpy(&d, &o->scalar, sizeof d); return d; }
    if (sdt == BLIS_FLOAT || sdt == BLIS_SCOMPLEX) { float f; memcpy(&f, &o->scalar, sizeof f); return (double)f; }
    return 1.0;
}

void fmm_obj_attach(num_t dt, dim_t m, dim_t n, void* p, inc_t rs, inc_t cs, obj_t* o)
{
    memset(o, 0, sizeof(*o));
    o->root = o;
    o->dim[0] = m;
    o->dim[1] = n;
    o->info = (dt & FMM_OBJ_DT_MASK) | (0x7u << 5) /* dense */;
    o->info2 = dt & 0x7u;
    o->elem_size = (dt == BLIS_DOUBLE) ? 8 : (dt == BLIS_FLOAT ? 4 : 16);
    o->buffer = p;
    o->rs = rs;
    o->cs = cs;
    o->is = 1;
    const float one_f = 1.0f;
    const double one_d = 1.0;
    if (dt == BLIS_FLOAT) memcpy(&o->scalar, &one_f, sizeof one_f);
    else memcpy(&o->scalar, &one_d, sizeof one_d);
    o->m_padded = m;
    o->n_padded = n;
}

void fmm_obj_scalar(double* storage, obj_t* o)
{
    fmm_obj_attach(BLIS_DOUBLE, 1, 1, storage, 1, 1, o);
}

/* ---- algorithm cache (the tables are hashed by content) ------------------ */

typedef struct {
    fmmhip_alg_t alg;
    int nlevels;
    int pins;               /* calls using alg now: a pinned slot is never evicted */
    fmm_t lv[4];            /* deep copies */
} alg_slot_t;

#define ALG_SLOTS 16
static alg_slot_t g_slots[ALG_SLOTS];
static int g_nslots = 0, g_next_victim = 0;
static pthread_mutex_t g_alg_mu = PTHREAD_MUTEX_INITIALIZER;

static size_t tab_bytes(const fmm_t* f, int which)
{
    long parts = which == 0 ? (long)f->m_tilde * f->k_tilde
               : which == 1 ? (long)f->k_tilde * f->n_tilde : (long)f->m_tilde * f->n_tilde;
    return sizeof(int) * (size_t)(parts * f->R);
}

static int same_fmm(const fmm_t* a, const fmm_t* b)
{
    return a->m_tilde == b->m_tilde && a->k_tilde == b->k_tilde && a->n_tilde == b->n_tilde &&
           a->R == b->R && !memcmp(a->U, b->U, tab_bytes(a, 0)) &&
           !memcmp(a->V, b->V, tab_bytes(a, 1)) && !memcmp(a->W, b->W, tab_bytes(a, 2));
}

static void free_slot(alg_slot_t* s)
{
    fmmhip_alg_destroy(s->alg);
    for (int l = 0; l < s->nlevels; l++) { free(s->lv[l].U); free(s->lv[l].V); free(s->lv[l].W); }
    memset(s, 0, sizeof(*s));
}

/* Returns a cached composed algorithm for these levels (or creates it),
 * pinned for the caller until release_alg.  Eviction (round robin) skips
 * pinned slots; when every slot is pinned the new algorithm is not cached
 * and *owned = 1 (release_alg destroys it). */
static int lookup_alg(const fmm_t* algs, int nlevels, fmmhip_alg_t* out, int* owned)
{
    *owned = 0;
    for (int l = 0; l < nlevels; l++)
        if (algs[l].m_tilde <= 0 || algs[l].k_tilde <= 0 || algs[l].n_tilde <= 0 || algs[l].R <= 0 ||
            !algs[l].U || !algs[l].V || !algs[l].W)
            return FMMHIP_ERR_INVALID;
    pthread_mutex_lock(&g_alg_mu);
    for (int i = 0; i < g_nslots; i++) {
        alg_slot_t* s = &g_slots[i];
        if (s->nlevels != nlevels) continue;
        int same = 1;
        for (int l = 0; l < nlevels && same; l++) same = same_fmm(&s->lv[l], &algs[l]);
        if (same) {
            s->pins++;
            *out = s->alg;
            pthread_mutex_unlock(&g_alg_mu);
            return FMMHIP_SUCCESS;
        }
    }
    fmmhip_alg_t alg;
    int rc = fmmhip_alg_create(&alg, algs, nlevels);
    if (rc) { pthread_mutex_unlock(&g_alg_mu); return rc; }
    alg_slot_t* s = NULL;
    if (g_nslots < ALG_SLOTS) {
        s = &g_slots[g_nslots++];
    } else {
        for (int t = 0; t < ALG_SLOTS && !s; t++) {
            alg_slot_t* c = &g_slots[(g_next_victim + t) % ALG_SLOTS];
            if (c->pins == 0) { s = c; g_next_victim = (g_next_victim + t + 1) % ALG_SLOTS; }
        }
        if (s) free_slot(s);
    }
    if (!s) {
        *owned = 1;
        *out = alg;
        pthread_mutex_unlock(&g_alg_mu);
        return FMMHIP_SUCCESS;
    }
    s->alg = alg;
    s->nlevels = nlevels;
    s->pins = 1;
    for (int l = 0; l < nlevels; l++) {
        s->lv[l] = algs[l];
        s->lv[l].U = malloc(tab_bytes(&algs[l], 0)); memcpy(s->lv[l].U, algs[l].U, tab_bytes(&algs[l], 0));
        s->lv[l].V = malloc(tab_bytes(&algs[l], 1)); memcpy(s->lv[l].V, algs[l].V, tab_bytes(&algs[l], 1));
        s->lv[l].W = malloc(tab_bytes(&algs[l], 2)); memcpy(s->lv[l].W, algs[l].W, tab_bytes(&algs[l], 2));
    }
    *out = alg;
    pthread_mutex_unlock(&g_alg_mu);
    return FMMHIP_SUCCESS;
}

/* Unpins an algorithm returned by lookup_alg (destroys an uncached one). */
static void release_alg(fmmhip_alg_t alg, int owned)
{
    if (!alg) return;
    if (owned) { fmmhip_alg_destroy(alg); return; }
    pthread_mutex_lock(&g_alg_mu);
    for (int i = 0; i < g_nslots; i++)
        if (g_slots[i].alg == alg && g_slots[i].pins > 0) { g_slots[i].pins--; break; }
    pthread_mutex_unlock(&g_alg_mu);
}

/* Number of cached algorithms and of pinned ones (for tests). */
void fmm_blis_amd_alg_cache_stats(int* cached, int* pinned)
{
    pthread_mutex_lock(&g_alg_mu);
    int p = 0;
    for (int i = 0; i < g_nslots; i++) p += g_slots[i].pins > 0;
    if (cached) *cached = g_nslots;
    if (pinned) *pinned = p;
    pthread_mutex_unlock(&g_alg_mu);
}

/* ---- host staging -------------------------------------------------------- */

/* Number of elements from element (0,0) to the last one, for non-negative
 * strides: the contiguous span that holds the operand. */
static long span_elems(long m, long n, long rs, long cs)
{
    if (m == 0 || n == 0) return 0;
    return (m - 1) * rs + (n - 1) * cs + 1;
}

typedef struct { void* p; size_t cap; } dbuf_t;
static dbuf_t g_stage[3];
static pthread_mutex_t g_stage_mu = PTHREAD_MUTEX_INITIALIZER;

static int stage_reserve(dbuf_t* b, size_t bytes)
{
    if (b->cap >= bytes) return 0;
    if (b->p) fmmhip_free(b->p);
    b->p = NULL; b->cap = 0;
    int rc = fmmhip_malloc(&b->p, bytes);
    if (rc) return rc;
    b->cap = bytes;
    return 0;
}

/* Packs a general-strided host operand into column-major (ld = m). */
static void* pack_colmajor(const view_t* v)
{
    char* out = (char*)malloc(v->esz * (size_t)(v->m * v->n) + 1);
    if (!out) return NULL;
    for (long j = 0; j < v->n; j++)
        for (long i = 0; i < v->m; i++)
            memcpy(out + v->esz * (size_t)(i + j * v->m), v->p + v->esz * (size_t)(i * v->rs + j * v->cs), v->esz);
    return out;
}

/* ---- the driver --------------------------------------------------------- */

static void fmm_driver(const char* who, obj_t* alpha, obj_t* A, obj_t* B, obj_t* beta, obj_t* C,
                       const fmm_t* algs, int nlevels, fmmhip_leaf_t leaf, int classical)
{
    g_last_error = 0;
    if (bli_plugin_register_fmm_blis() != BLIS_SUCCESS) {
        fail(who, FMMHIP_ERR_NODEVICE, "plugin registration failed");
        return;
    }
    if (!alpha || !A || !B || !beta || !C) { fail(who, FMMHIP_ERR_INVALID, "NULL object"); return; }
    view_t va, vb, vc;
    obj_view(A, &va);
    obj_view(B, &vb);
    obj_view(C, &vc);
    /* bli_gemm_check: datatypes and conformal dimensions */
    if (vc.dt != BLIS_DOUBLE && vc.dt != BLIS_FLOAT) {
        fail(who, FMMHIP_ERR_UNSUPPORTED, "only real float/double objects are supported");
        return;
    }
    if (va.dt != vc.dt || vb.dt != vc.dt) { fail(who, FMMHIP_ERR_UNSUPPORTED, "mixed datatypes"); return; }
    if (va.m != vc.m || vb.n != vc.n || va.n != vb.m) { fail(who, FMMHIP_ERR_INVALID, "nonconformal dimensions"); return; }
    if (va.rs < 0 || va.cs < 0 || vb.rs < 0 || vb.cs < 0 || vc.rs < 0 || vc.cs < 0) {
        fail(who, FMMHIP_ERR_UNSUPPORTED, "negative strides");
        return;
    }
    double a_val, b_val;
    if (obj_scalar_value(alpha, &a_val) || obj_scalar_value(beta, &b_val)) {
        fail(who, FMMHIP_ERR_INVALID, "alpha/beta are not numeric scalars");
        return;
    }
    a_val *= obj_internal_scalar(A) * obj_internal_scalar(B);
    const long m = vc.m, n = vc.n, k = va.n;
    /* bli_l3_return_early_if_trivial */
    if (m == 0 || n == 0) return;

    const int dbl = vc.dt == BLIS_DOUBLE;
    if (!dbl && leaf != FMMHIP_LEAF_NATIVE) { fail(who, FMMHIP_ERR_UNSUPPORTED, "reduced leaf needs double objects"); return; }
    fmmhip_alg_t alg = NULL;
    int alg_owned = 0;
    if (!classical) {
        int rc = lookup_alg(algs, nlevels, &alg, &alg_owned);
        if (rc) { fail(who, rc, "invalid algorithm (zero-sized, NULL tables or Brent equations violated)"); return; }
    }

    int dev_a = va.m && va.n ? fmmhip_pointer_is_device(va.p) : 1;
    int dev_b = vb.m && vb.n ? fmmhip_pointer_is_device(vb.p) : 1;
    int dev_c = fmmhip_pointer_is_device(vc.p);
    void* stream = fmmhip_default_stream();

    const void *pa = va.p, *pb = vb.p;
    void* pc = vc.p;
    void *tmp_a = NULL, *tmp_b = NULL;
    int locked = 0, rc = 0;
    view_t ua = va, ub = vb;

    if (!(dev_a == 1 && dev_b == 1 && dev_c == 1)) {
        /* host operands: stage through device buffers, keeping strides */
        pthread_mutex_lock(&g_stage_mu);
        locked = 1;
        if (dev_a != 1 && va.rs != 1 && va.cs != 1 && k > 0) {
            tmp_a = pack_colmajor(&va);
            if (!tmp_a) { rc = FMMHIP_ERR_NOMEM; goto out; }
            ua.p = tmp_a; ua.rs = 1; ua.cs = va.m;
        }
        if (dev_b != 1 && vb.rs != 1 && vb.cs != 1 && k > 0) {
            tmp_b = pack_colmajor(&vb);
            if (!tmp_b) { rc = FMMHIP_ERR_NOMEM; goto out; }
            ub.p = tmp_b; ub.rs = 1; ub.cs = vb.m;
        }
        if (dev_a != 1 && k > 0) {
            size_t bytes = ua.esz * (size_t)span_elems(ua.m, ua.n, ua.rs, ua.cs);
            if ((rc = stage_reserve(&g_stage[0], bytes))) goto out;
            if ((rc = fmmhip_memcpy_h2d(g_stage[0].p, ua.p, bytes))) goto out;
            pa = g_stage[0].p;
        }
        if (dev_b != 1 && k > 0) {
            size_t bytes = ub.esz * (size_t)span_elems(ub.m, ub.n, ub.rs, ub.cs);
            if ((rc = stage_reserve(&g_stage[1], bytes))) goto out;
            if ((rc = fmmhip_memcpy_h2d(g_stage[1].p, ub.p, bytes))) goto out;
            pb = g_stage[1].p;
        }
        if (dev_c != 1) {
            const long span = span_elems(vc.m, vc.n, vc.rs, vc.cs);
            size_t bytes = vc.esz * (size_t)span;
            if ((rc = stage_reserve(&g_stage[2], bytes))) goto out;
            /* beta == 0 never reads C: a C without gaps between its columns
             * (or rows) need not be uploaded; with gaps it is, so that the
             * copy back returns the gap elements unchanged */
            if (!(b_val == 0.0 && span == m * n) && (rc = fmmhip_memcpy_h2d(g_stage[2].p, vc.p, bytes))) goto out;
            pc = g_stage[2].p;
        }
    }
    /* device operands with neither stride 1 (e.g. strided sub-views) are
     * gathered on the device by the native layer */

    if (dbl) {
        if (classical)
            rc = fmmhip_dgemm(m, n, k, a_val, (const double*)pa, ua.rs, ua.cs, (const double*)pb, ub.rs, ub.cs,
                              b_val, (double*)pc, vc.rs, vc.cs, stream);
        else
            rc = fmmhip_dgemm_fmm(alg, leaf, m, n, k, a_val, (const double*)pa, ua.rs, ua.cs, (const double*)pb,
                                  ub.rs, ub.cs, b_val, (double*)pc, vc.rs, vc.cs, stream);
    } else {
        if (classical)
            rc = fmmhip_sgemm(m, n, k, (float)a_val, (const float*)pa, ua.rs, ua.cs, (const float*)pb, ub.rs, ub.cs,
                              (float)b_val, (float*)pc, vc.rs, vc.cs, stream);
        else
            rc = fmmhip_sgemm_fmm(alg, m, n, k, (float)a_val, (const float*)pa, ua.rs, ua.cs, (const float*)pb,
                                  ub.rs, ub.cs, (float)b_val, (float*)pc, vc.rs, vc.cs, stream);
    }
    if (rc) goto out;
    if ((rc = fmmhip_stream_sync(stream))) goto out;
    if (pc != vc.p) {
        size_t bytes = vc.esz * (size_t)span_elems(vc.m, vc.n, vc.rs, vc.cs);
        rc = fmmhip_memcpy_d2h(vc.p, pc, bytes);
    }
out:
    if (locked) pthread_mutex_unlock(&g_stage_mu);
    release_alg(alg, alg_owned);
    free(tmp_a);
    free(tmp_b);
    if (rc) fail(who, rc, "device computation failed");
}

void bli_strassen_ab_ex(obj_t* alpha, obj_t* A, obj_t* B, obj_t* beta, obj_t* C, fmm_t fmm)
{
    fmm_driver("bli_strassen_ab_ex", alpha, A, B, beta, C, &fmm, 1, FMMHIP_LEAF_NATIVE, 0);
}

void bli_strassen_ab(obj_t* alpha, obj_t* A, obj_t* B, obj_t* beta, obj_t* C)
{
    bli_strassen_ab_ex(alpha, A, B, beta, C, STRASSEN_FMM);
}

void bli_fmm_ab_ex(obj_t* alpha, obj_t* A, obj_t* B, obj_t* beta, obj_t* C, const fmm_t* algs, int nlevels,
                   fmmhip_leaf_t leaf)
{
    if (!algs || nlevels < 1 || nlevels > 4) {
        g_last_error = FMMHIP_ERR_INVALID;
        fail("bli_fmm_ab_ex", FMMHIP_ERR_INVALID, "nlevels must be 1..4 with a table per level");
        return;
    }
    fmm_driver("bli_fmm_ab_ex", alpha, A, B, beta, C, algs, nlevels, leaf, 0);
}

void bli_gemm_hip(obj_t* alpha, obj_t* A, obj_t* B, obj_t* beta, obj_t* C)
{
    fmm_driver("bli_gemm_hip", alpha, A, B, beta, C, NULL, 0, FMMHIP_LEAF_NATIVE, 1);
}
