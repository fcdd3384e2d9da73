/*
 * fmm_spec.c -- algorithm tables: the coefficient-file reader, the built-in
 * Strassen algorithm, L-level composition, the Brent check, the per-product
 * sparse lists the kernels walk, and their resolution against a problem's
 * part geometry.
 *
 * Reference counterparts (dnparikh/fmm-blis):
 *   new_fmm / free_fmm   frame/base/bli_read_fmm.c:24-86
 *   STRASSEN_FMM         frame/base/bli_strassen_ab.c:12-24
 *   fmm_part_split       frame/base/bli_strassen_ab.c:86-121 (init_part_offsets)
 *   per-product lists    frame/base/bli_strassen_ab.c:251-283 (the reference
 *                        hands every part, zero coefficients included, to the
 *                        pack and gemm hooks; here only nonzeros are listed)
 */
#define _GNU_SOURCE
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fmm_internal.h"

/* ---- built-in Strassen (the <2,2,2;7> table of 222.txt) ----------------- */

static int strassen_u[4 * 7] = {
    1, 0, 1, 0, 1, -1, 0,
    0, 0, 0, 0, 1, 0, 1,
    0, 1, 0, 0, 0, 1, 0,
    1, 1, 0, 1, 0, 0, -1,
};
static int strassen_v[4 * 7] = {
    1, 1, 0, -1, 0, 1, 0,
    0, 0, 1, 0, 0, 1, 0,
    0, 0, 0, 1, 0, 0, 1,
    1, 0, -1, 0, 1, 0, 1,
};
static int strassen_w[4 * 7] = {
    1, 0, 0, 1, -1, 0, 1,
    0, 0, 1, 0, 1, 0, 0,
    0, 1, 0, 1, 0, 0, 0,
    1, -1, 1, 0, 0, 1, 0,
};

fmm_t STRASSEN_FMM = {
    .m_tilde = 2, .n_tilde = 2, .k_tilde = 2, .R = 7,
    .U = strassen_u, .V = strassen_v, .W = strassen_w,
};

/* ---- coefficient-file reader -------------------------------------------- */

/* Reads whole lines of any length; returns NULL at EOF. */
static char* read_line(FILE* fp, char** buf, size_t* cap)
{
    ssize_t n = getline(buf, cap, fp);
    return n < 0 ? NULL : *buf;
}

/* Parses whitespace-separated integers of a line into out[0..max); returns
 * how many integers the line holds, or -1 on a non-integer token. */
static int scan_ints(const char* s, int* out, int max)
{
    int count = 0;
    for (;;) {
        while (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n') s++;
        if (!*s) break;
        char* end;
        errno = 0;
        long v = strtol(s, &end, 10);
        if (end == s || errno) return -1;
        if (count < max) out[count] = (int)v;
        count++;
        s = end;
    }
    return count;
}

static fmm_t empty_fmm(void)
{
    fmm_t f;
    memset(&f, 0, sizeof(f));
    return f;
}

fmm_t new_fmm(const char* file_name)
{
    fmm_t f = empty_fmm();
    if (!file_name) {
        fprintf(stderr, "new_fmm: NULL file name\n");
        return f;
    }
    FILE* fp = fopen(file_name, "r");
    if (!fp) {
        fprintf(stderr, "new_fmm: cannot open '%s': %s\n", file_name, strerror(errno));
        return f;
    }
    char* line = NULL;
    size_t cap = 0;
    int hdr[4];
    const char* why = NULL;
    int* row_buf = NULL;

    /* Header: first non-comment line, "m k n R" in file order. */
    char* s;
    while ((s = read_line(fp, &line, &cap)) && (s[0] == '#' || scan_ints(s, hdr, 4) == 0)) {}
    if (!s || scan_ints(s, hdr, 4) != 4) { why = "header must be 'm k n R'"; goto fail; }
    if (hdr[0] <= 0 || hdr[1] <= 0 || hdr[2] <= 0 || hdr[3] <= 0) { why = "non-positive header value"; goto fail; }
    f.m_tilde = hdr[0];
    f.k_tilde = hdr[1];
    f.n_tilde = hdr[2];
    f.R = hdr[3];
    {
        long na = (long)f.m_tilde * f.k_tilde, nb = (long)f.k_tilde * f.n_tilde,
             nc = (long)f.m_tilde * f.n_tilde, rows = 0;
        f.U = (int*)calloc((size_t)(na * f.R), sizeof(int));
        f.V = (int*)calloc((size_t)(nb * f.R), sizeof(int));
        f.W = (int*)calloc((size_t)(nc * f.R), sizeof(int));
        row_buf = (int*)malloc(sizeof(int) * (size_t)(f.R + 1));
        if (!f.U || !f.V || !f.W || !row_buf) { why = "out of memory"; goto fail; }
        while ((s = read_line(fp, &line, &cap))) {
            if (s[0] == '#') continue;
            int cnt = scan_ints(s, row_buf, f.R + 1);
            if (cnt == 0) continue;
            if (cnt != f.R) { why = "a coefficient row does not hold R integers"; goto fail; }
            int* dst;
            if (rows < na) dst = f.U + rows * f.R;
            else if (rows < na + nb) dst = f.V + (rows - na) * f.R;
            else if (rows < na + nb + nc) dst = f.W + (rows - na - nb) * f.R;
            else { why = "more coefficient rows than m*k + k*n + m*n"; goto fail; }
            memcpy(dst, row_buf, sizeof(int) * (size_t)f.R);
            rows++;
        }
        if (rows != na + nb + nc) { why = "fewer coefficient rows than m*k + k*n + m*n"; goto fail; }
    }
    free(row_buf);
    free(line);
    fclose(fp);
    return f;

fail:
    fprintf(stderr, "new_fmm: '%s': %s\n", file_name, why);
    free(row_buf);
    free(line);
    fclose(fp);
    free(f.U); free(f.V); free(f.W);
    return empty_fmm();
}

void free_fmm(fmm_t fmm)
{
    if (fmm.U == strassen_u) return;     /* the built-in table is static */
    free(fmm.U);
    free(fmm.V);
    free(fmm.W);
}

/* ---- composition, Brent check, sparse lists ----------------------------- */

static int fmm_valid(const fmm_t* f)
{
    return f && f->m_tilde > 0 && f->k_tilde > 0 && f->n_tilde > 0 && f->R > 0 &&
           f->U && f->V && f->W;
}

/* out[((a1*rows2+a2)*cols + (b1*cols2+b2))*R + r1*R2 + r2] = X1 * X2 */
static void kron_table(const int* X1, int R1, int rows1, int cols1,
                       const int* X2, int R2, int rows2, int cols2, int* out)
{
    const int R = R1 * R2, cols = cols1 * cols2;
    for (int a1 = 0; a1 < rows1; a1++)
    for (int a2 = 0; a2 < rows2; a2++)
    for (int b1 = 0; b1 < cols1; b1++)
    for (int b2 = 0; b2 < cols2; b2++) {
        const int* x1 = X1 + (a1 * cols1 + b1) * R1;
        const int* x2 = X2 + (a2 * cols2 + b2) * R2;
        int* o = out + (size_t)((a1 * rows2 + a2) * cols + b1 * cols2 + b2) * R;
        for (int r1 = 0; r1 < R1; r1++) {
            if (!x1[r1]) { memset(o + r1 * R2, 0, sizeof(int) * R2); continue; }
            for (int r2 = 0; r2 < R2; r2++) o[r1 * R2 + r2] = x1[r1] * x2[r2];
        }
    }
}

static long brent_violations_tables(int mt, int kt, int nt, int R,
                                    const int* U, const int* V, const int* W)
{
    /* sum_r U[(i,p)] V[(q,j)] W[(a,b)] == [i==a][p==q][j==b] */
    long bad = 0;
    for (int i = 0; i < mt; i++) for (int p = 0; p < kt; p++) {
        const int* u = U + (size_t)(i * kt + p) * R;
        for (int q = 0; q < kt; q++) for (int j = 0; j < nt; j++) {
            const int* v = V + (size_t)(q * nt + j) * R;
            for (int a = 0; a < mt; a++) for (int b = 0; b < nt; b++) {
                const int* w = W + (size_t)(a * nt + b) * R;
                long s = 0;
                for (int r = 0; r < R; r++) if (u[r] && v[r]) s += (long)u[r] * v[r] * w[r];
                if (s != ((i == a && p == q && j == b) ? 1 : 0)) bad++;
            }
        }
    }
    return bad;
}

static int build_csr(const int* X, int parts, int R, int** ptr, int** part, int** coef,
                     long* nnz, int* maxnnz)
{
    long cnt = 0;
    for (long i = 0; i < (long)parts * R; i++) cnt += X[i] != 0;
    *ptr = (int*)malloc(sizeof(int) * (size_t)(R + 1));
    *part = (int*)malloc(sizeof(int) * (size_t)(cnt + 1));
    *coef = (int*)malloc(sizeof(int) * (size_t)(cnt + 1));
    if (!*ptr || !*part || !*coef) return FMMHIP_ERR_NOMEM;
    long e = 0;
    *maxnnz = 0;
    for (int r = 0; r < R; r++) {
        (*ptr)[r] = (int)e;
        for (int s = 0; s < parts; s++) {
            int c = X[(size_t)s * R + r];
            if (c) { (*part)[e] = s; (*coef)[e] = c; e++; }
        }
        int here = (int)(e - (*ptr)[r]);
        if (here > *maxnnz) *maxnnz = here;
    }
    (*ptr)[R] = (int)e;
    *nnz = cnt;
    return FMMHIP_SUCCESS;
}

/* Removes products whose U, V or W column is all zero: they contribute
 * nothing, and the kernels assume every product has a nonzero in each. */
static void drop_empty_products(struct fmmhip_alg_s* g)
{
    const int pa = g->mt * g->kt, pb = g->kt * g->nt, pc = g->mt * g->nt, R = g->R;
    int keep = 0;
    for (int r = 0; r < R; r++) {
        int nu = 0, nv = 0, nw = 0;
        for (int s = 0; s < pa; s++) nu |= g->U[(size_t)s * R + r];
        for (int s = 0; s < pb; s++) nv |= g->V[(size_t)s * R + r];
        for (int s = 0; s < pc; s++) nw |= g->W[(size_t)s * R + r];
        if (!nu || !nv || !nw) continue;
        for (int s = 0; s < pa; s++) g->U[(size_t)s * R + keep] = g->U[(size_t)s * R + r];
        for (int s = 0; s < pb; s++) g->V[(size_t)s * R + keep] = g->V[(size_t)s * R + r];
        for (int s = 0; s < pc; s++) g->W[(size_t)s * R + keep] = g->W[(size_t)s * R + r];
        keep++;
    }
    if (keep == R) return;
    /* re-pack rows from stride R to stride keep */
    for (int s = 0; s < pa; s++) memmove(g->U + (size_t)s * keep, g->U + (size_t)s * R, sizeof(int) * keep);
    for (int s = 0; s < pb; s++) memmove(g->V + (size_t)s * keep, g->V + (size_t)s * R, sizeof(int) * keep);
    for (int s = 0; s < pc; s++) memmove(g->W + (size_t)s * keep, g->W + (size_t)s * R, sizeof(int) * keep);
    g->R = keep;
}

static int finish_alg(struct fmmhip_alg_s* g)
{
    int rc;
    drop_empty_products(g);
    if (g->R < 1) return FMMHIP_ERR_INVALID;
    if ((rc = build_csr(g->U, g->mt * g->kt, g->R, &g->u_ptr, &g->u_part, &g->u_coef, &g->nnz_u, &g->max_nnz_u))) return rc;
    if ((rc = build_csr(g->V, g->kt * g->nt, g->R, &g->v_ptr, &g->v_part, &g->v_coef, &g->nnz_v, &g->max_nnz_v))) return rc;
    if ((rc = build_csr(g->W, g->mt * g->nt, g->R, &g->w_ptr, &g->w_part, &g->w_coef, &g->nnz_w, &g->max_nnz_w))) return rc;
    /* First writer of each C part: bit b of w_first[e] is set when entry e
     * is the first write of its part among the products of its group, the
     * products being split into 2^b contiguous groups [g*R/G, (g+1)*R/G)
     * (G = 1, 2, 4, 8).  Bit 0 (one group) is the first product overall,
     * which applies beta. */
    g->w_first = (int*)calloc((size_t)(g->nnz_w + 1), sizeof(int));
    char* seen = (char*)calloc((size_t)(g->mt * g->nt), 1);
    if (!g->w_first || !seen) { free(seen); return FMMHIP_ERR_NOMEM; }
    for (int b = 0; b < 4; b++) {
        const int G = 1 << b;
        for (int grp = 0; grp < G; grp++) {
            memset(seen, 0, (size_t)(g->mt * g->nt));
            const int r0 = (int)((long)grp * g->R / G), r1 = (int)((long)(grp + 1) * g->R / G);
            for (int r = r0; r < r1; r++)
                for (int e = g->w_ptr[r]; e < g->w_ptr[r + 1]; e++) {
                    int p = g->w_part[e];
                    if (!seen[p]) { seen[p] = 1; g->w_first[e] |= 1 << b; }
                }
        }
    }
    free(seen);
    return FMMHIP_SUCCESS;
}

void fmmhip_alg_destroy(fmmhip_alg_t g)
{
    if (!g) return;
    fmmhip_alg_release_cache(g);
    free(g->U); free(g->V); free(g->W);
    free(g->u_ptr); free(g->u_part); free(g->u_coef);
    free(g->v_ptr); free(g->v_part); free(g->v_coef);
    free(g->w_ptr); free(g->w_part); free(g->w_coef); free(g->w_first);
    for (int l = 0; l < 2; l++) { free(g->lv_U[l]); free(g->lv_V[l]); free(g->lv_W[l]); }
    free(g);
}

int fmmhip_alg_create(fmmhip_alg_t* out, const fmm_t* algs, int nlevels)
{
    if (!out || !algs || nlevels < 1 || nlevels > 4) return FMMHIP_ERR_INVALID;
    *out = NULL;
    for (int l = 0; l < nlevels; l++) {
        if (!fmm_valid(&algs[l])) return FMMHIP_ERR_INVALID;
        if (brent_violations_tables(algs[l].m_tilde, algs[l].k_tilde, algs[l].n_tilde, algs[l].R,
                                    algs[l].U, algs[l].V, algs[l].W) != 0)
            return FMMHIP_ERR_INVALID;
    }
    long mt = 1, kt = 1, nt = 1, R = 1;
    for (int l = 0; l < nlevels; l++) {
        mt *= algs[l].m_tilde; kt *= algs[l].k_tilde; nt *= algs[l].n_tilde; R *= algs[l].R;
    }
    if (mt * kt * R > (1L << 26) || kt * nt * R > (1L << 26) || mt * nt * R > (1L << 26))
        return FMMHIP_ERR_UNSUPPORTED;
    struct fmmhip_alg_s* g = (struct fmmhip_alg_s*)calloc(1, sizeof(*g));
    if (!g) return FMMHIP_ERR_NOMEM;
    g->levels = nlevels;
    g->shape = -1;
    g->split = -1;
    g->pack = -1;
    g->last_pack = -1;
    g->last_shape = -1;
    /* start from level 0 and fold the next levels in as inner blocks */
    int cm = algs[0].m_tilde, ck = algs[0].k_tilde, cn = algs[0].n_tilde, cR = algs[0].R;
    int* U = (int*)malloc(sizeof(int) * (size_t)(cm * ck * cR));
    int* V = (int*)malloc(sizeof(int) * (size_t)(ck * cn * cR));
    int* W = (int*)malloc(sizeof(int) * (size_t)(cm * cn * cR));
    if (!U || !V || !W) { free(U); free(V); free(W); free(g); return FMMHIP_ERR_NOMEM; }
    memcpy(U, algs[0].U, sizeof(int) * (size_t)(cm * ck * cR));
    memcpy(V, algs[0].V, sizeof(int) * (size_t)(ck * cn * cR));
    memcpy(W, algs[0].W, sizeof(int) * (size_t)(cm * cn * cR));
    for (int l = 1; l < nlevels; l++) {
        const fmm_t* b = &algs[l];
        int nm = cm * b->m_tilde, nk = ck * b->k_tilde, nn = cn * b->n_tilde, nR = cR * b->R;
        int* U2 = (int*)malloc(sizeof(int) * (size_t)nm * nk * nR);
        int* V2 = (int*)malloc(sizeof(int) * (size_t)nk * nn * nR);
        int* W2 = (int*)malloc(sizeof(int) * (size_t)nm * nn * nR);
        if (!U2 || !V2 || !W2) {
            free(U2); free(V2); free(W2); free(U); free(V); free(W); free(g);
            return FMMHIP_ERR_NOMEM;
        }
        kron_table(U, cR, cm, ck, b->U, b->R, b->m_tilde, b->k_tilde, U2);
        kron_table(V, cR, ck, cn, b->V, b->R, b->k_tilde, b->n_tilde, V2);
        kron_table(W, cR, cm, cn, b->W, b->R, b->m_tilde, b->n_tilde, W2);
        free(U); free(V); free(W);
        U = U2; V = V2; W = W2;
        cm = nm; ck = nk; cn = nn; cR = nR;
    }
    g->mt = cm; g->kt = ck; g->nt = cn; g->R = cR;
    g->U = U; g->V = V; g->W = W;
    if (nlevels == 2) {
        for (int l = 0; l < 2; l++) {
            const fmm_t* b = &algs[l];
            const size_t nu = (size_t)b->m_tilde * b->k_tilde * b->R, nv = (size_t)b->k_tilde * b->n_tilde * b->R,
                         nw = (size_t)b->m_tilde * b->n_tilde * b->R;
            g->lv_m[l] = b->m_tilde; g->lv_k[l] = b->k_tilde; g->lv_n[l] = b->n_tilde; g->lv_R[l] = b->R;
            g->lv_U[l] = (int*)malloc(sizeof(int) * nu);
            g->lv_V[l] = (int*)malloc(sizeof(int) * nv);
            g->lv_W[l] = (int*)malloc(sizeof(int) * nw);
            if (!g->lv_U[l] || !g->lv_V[l] || !g->lv_W[l]) { fmmhip_alg_destroy(g); return FMMHIP_ERR_NOMEM; }
            memcpy(g->lv_U[l], b->U, sizeof(int) * nu);
            memcpy(g->lv_V[l], b->V, sizeof(int) * nv);
            memcpy(g->lv_W[l], b->W, sizeof(int) * nw);
        }
    }
    g->classical = (cm == 1 && ck == 1 && cn == 1 && cR == 1);
    int rc = finish_alg(g);
    if (rc) { fmmhip_alg_destroy(g); return rc; }
    *out = g;
    return FMMHIP_SUCCESS;
}

int fmmhip_alg_create_classical(fmmhip_alg_t* out)
{
    static int one = 1;
    fmm_t c = { .m_tilde = 1, .n_tilde = 1, .k_tilde = 1, .R = 1, .U = &one, .V = &one, .W = &one };
    return fmmhip_alg_create(out, &c, 1);
}

int fmmhip_alg_info(fmmhip_alg_t g, int* mt, int* kt, int* nt, int* R,
                    long* nnz_u, long* nnz_v, long* nnz_w)
{
    if (!g) return FMMHIP_ERR_INVALID;
    if (mt) *mt = g->mt;
    if (kt) *kt = g->kt;
    if (nt) *nt = g->nt;
    if (R) *R = g->R;
    if (nnz_u) *nnz_u = g->nnz_u;
    if (nnz_v) *nnz_v = g->nnz_v;
    if (nnz_w) *nnz_w = g->nnz_w;
    return FMMHIP_SUCCESS;
}

long fmmhip_alg_brent_violations(fmmhip_alg_t g)
{
    if (!g) return -1;
    return brent_violations_tables(g->mt, g->kt, g->nt, g->R, g->U, g->V, g->W);
}

/* ---- geometry ----------------------------------------------------------- */

void fmm_part_split(long whole, int tilde, long* size, long* off)
{
    long q = whole / tilde, extra = whole % tilde, at = 0;
    for (int t = 0; t < tilde; t++) {
        size[t] = q + (t < extra);
        off[t] = at;
        at += size[t];
    }
}

int fmm_build_entries(const struct fmmhip_alg_s* g, const fmm_geom_t* geo,
                      int32_t** ptr_out, fmm_entry_t** ent_out, long* nent_out)
{
    const int mt = g->mt, kt = g->kt, nt = g->nt, R = g->R;
    long* sizes = (long*)malloc(sizeof(long) * 2 * (size_t)(mt + kt + nt));
    long nent = g->nnz_u + g->nnz_v + g->nnz_w;
    int32_t* ptr = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(R + 1));
    fmm_entry_t* ent = (fmm_entry_t*)calloc((size_t)(nent + 1), sizeof(fmm_entry_t));
    if (!sizes || !ptr || !ent) { free(sizes); free(ptr); free(ent); return FMMHIP_ERR_NOMEM; }
    fmm_build_entries_into(g, geo, ptr, ent, sizes);
    free(sizes);
    *ptr_out = ptr;
    *ent_out = ent;
    *nent_out = nent;
    return FMMHIP_SUCCESS;
}

void fmm_build_entries_into(const struct fmmhip_alg_s* g, const fmm_geom_t* geo, int32_t* ptr, fmm_entry_t* ent,
                            long* sizes)
{
    const int mt = g->mt, kt = g->kt, nt = g->nt, R = g->R;
    long* szm = sizes;
    long *ofm = szm + mt, *szk = ofm + mt, *ofk = szk + kt, *szn = ofk + kt, *ofn = szn + nt;
    fmm_part_split(geo->m, mt, szm, ofm);
    fmm_part_split(geo->k, kt, szk, ofk);
    fmm_part_split(geo->n, nt, szn, ofn);

    /* pointers are absolute indices into ent[] */
    for (int r = 0; r <= R; r++) {
        ptr[r] = g->u_ptr[r];
        ptr[(R + 1) + r] = (int32_t)(g->nnz_u + g->v_ptr[r]);
        ptr[2 * (R + 1) + r] = (int32_t)(g->nnz_u + g->nnz_v + g->w_ptr[r]);
    }
    long e = 0;
    for (long i = 0; i < g->nnz_u; i++, e++) {       /* A part (i,p) = row i*kt+p */
        int s = g->u_part[i], pi = s / kt, pp = s % kt;
        ent[e].off = ofm[pi] * geo->rsA + ofk[pp] * geo->csA;
        ent[e].rows = (int32_t)szm[pi];
        ent[e].cols = (int32_t)szk[pp];
        ent[e].coef = (double)g->u_coef[i];
        ent[e].part = s;
    }
    for (long i = 0; i < g->nnz_v; i++, e++) {       /* B part (p,j) = row p*nt+j */
        int s = g->v_part[i], pp = s / nt, pj = s % nt;
        ent[e].off = ofk[pp] * geo->rsB + ofn[pj] * geo->csB;
        ent[e].rows = (int32_t)szk[pp];
        ent[e].cols = (int32_t)szn[pj];
        ent[e].coef = (double)g->v_coef[i];
        ent[e].part = s;
    }
    for (long i = 0; i < g->nnz_w; i++, e++) {       /* C part (i,j) = row i*nt+j */
        int s = g->w_part[i], pi = s / nt, pj = s % nt;
        ent[e].off = ofm[pi] * geo->rsC + ofn[pj] * geo->csC;
        ent[e].rows = (int32_t)szm[pi];
        ent[e].cols = (int32_t)szn[pj];
        ent[e].coef = (double)g->w_coef[i];
        ent[e].first = g->w_first[i];
        ent[e].part = s;
    }
}
