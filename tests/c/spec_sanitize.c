/* Host-side checks of fmm_spec.c (coefficient-file parser, level
 * composition, part geometry, descriptor lists) built with
 * -fsanitize=address,undefined: every path runs over the algorithm files,
 * malformed and oversized files, one- and two-level and mixed compositions,
 * and uneven geometries, and each result is checked for consistency.  Any
 * sanitizer report aborts the program (-fno-sanitize-recover=all).
 *
 * Usage: spec_sanitize <algorithm dir> <scratch dir>
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fmm_internal.h"

/* The library's HIP layer frees an algorithm's device descriptors here; the
 * host-only build has none. */
void fmmhip_alg_release_cache(struct fmmhip_alg_s* alg)
{
    (void)alg;
}

static int failures = 0;

#define CHECK(cond, ...)                                                                 \
    do {                                                                                 \
        if (!(cond)) {                                                                   \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);                         \
            fprintf(stderr, __VA_ARGS__);                                                \
            fprintf(stderr, "\n");                                                       \
            failures++;                                                                  \
        }                                                                                \
    } while (0)

static void write_file(const char* path, const char* text)
{
    FILE* f = fopen(path, "w");
    if (!f) { perror(path); exit(2); }
    fputs(text, f);
    fclose(f);
}

/* Every geometry: parts tile each dimension, entries stay inside the
 * operand, CSR pointers are monotone and end at the entry count. */
static void check_entries(fmmhip_alg_t g, long m, long n, long k, int col_major)
{
    fmm_geom_t geo;
    memset(&geo, 0, sizeof(geo));
    geo.m = m; geo.n = n; geo.k = k;
    geo.lm = (m + g->mt - 1) / g->mt;
    geo.lk = (k + g->kt - 1) / g->kt;
    geo.ln = (n + g->nt - 1) / g->nt;
    geo.rsA = col_major ? 1 : k;  geo.csA = col_major ? m : 1;
    geo.rsB = col_major ? 1 : n;  geo.csB = col_major ? k : 1;
    geo.rsC = col_major ? 1 : n;  geo.csC = col_major ? m : 1;
    int32_t* ptr = NULL;
    fmm_entry_t* ent = NULL;
    long nent = 0;
    int rc = fmm_build_entries(g, &geo, &ptr, &ent, &nent);
    CHECK(rc == 0, "fmm_build_entries rc %d", rc);
    if (rc) return;
    const int R = g->R;
    CHECK(nent == g->nnz_u + g->nnz_v + g->nnz_w, "entry count %ld", nent);
    for (int s = 0; s < 3; s++) {
        const int32_t* p = ptr + s * (R + 1);
        for (int r = 0; r < R; r++) CHECK(p[r] <= p[r + 1], "ptr not monotone (list %d, r %d)", s, r);
    }
    CHECK(ptr[3 * (R + 1) - 1] == nent, "last pointer %d of %ld", ptr[3 * (R + 1) - 1], nent);
    for (long e = 0; e < nent; e++) {
        const fmm_entry_t* x = &ent[e];
        long rows_total, cols_total, rs, cs;
        if (e < g->nnz_u) { rows_total = m; cols_total = k; rs = geo.rsA; cs = geo.csA; }
        else if (e < g->nnz_u + g->nnz_v) { rows_total = k; cols_total = n; rs = geo.rsB; cs = geo.csB; }
        else { rows_total = m; cols_total = n; rs = geo.rsC; cs = geo.csC; }
        CHECK(x->rows >= 0 && x->cols >= 0, "negative extent at %ld", e);
        if (x->rows > 0 && x->cols > 0) {
            const long last = x->off + (long)(x->rows - 1) * rs + (long)(x->cols - 1) * cs;
            CHECK(x->off >= 0 && last < rows_total * cols_total, "entry %ld leaves the operand", e);
        }
        CHECK(x->coef != 0.0, "zero coefficient listed at %ld", e);
    }
    free(ptr);
    free(ent);
}

static void check_alg(const char* name, const fmm_t* levels, int nlevels)
{
    fmmhip_alg_t g = NULL;
    int rc = fmmhip_alg_create(&g, levels, nlevels);
    CHECK(rc == 0 && g, "%s: fmmhip_alg_create(%d levels) rc %d", name, nlevels, rc);
    if (rc || !g) return;
    /* Brent check: ~R * (mt kt nt)^2 work, skipped for the 2401-product
     * 444 L2 table (its composition is compared with the oracle's by the
     * host tests) */
    if (g->R <= 600) CHECK(fmmhip_alg_brent_violations(g) == 0, "%s L%d: Brent equations violated", name, nlevels);
    int mt, kt, nt, R;
    long nu, nv, nw;
    CHECK(fmmhip_alg_info(g, &mt, &kt, &nt, &R, &nu, &nv, &nw) == 0, "%s: info", name);
    const long shapes[][3] = {{1, 1, 1}, {2, 3, 5}, {17, 1, 9}, {257, 250, 263}, {600, 31, 1000}, {64, 64, 64}};
    for (size_t s = 0; s < sizeof(shapes) / sizeof(shapes[0]); s++)
        for (int cm = 0; cm < 2; cm++) check_entries(g, shapes[s][0], shapes[s][1], shapes[s][2], cm);
    fmmhip_alg_destroy(g);
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <algorithm dir> <scratch dir>\n", argv[0]);
        return 2;
    }
    const char* dir = argv[1];
    const char* tmp = argv[2];
    char path[4096];

    /* part geometry: sizes differ by at most one and tile the dimension */
    for (long whole = 0; whole < 300; whole++)
        for (int tilde = 1; tilde <= 16; tilde++) {
            long size[16], off[16], sum = 0;
            fmm_part_split(whole, tilde, size, off);
            for (int t = 0; t < tilde; t++) {
                CHECK(off[t] == sum, "offset of part %d of %ld/%d", t, whole, tilde);
                CHECK(size[t] == whole / tilde || size[t] == whole / tilde + 1, "size of part %d", t);
                sum += size[t];
            }
            CHECK(sum == whole, "parts of %ld/%d sum to %ld", whole, tilde, sum);
        }

    /* every algorithm file, one and two levels, and mixed compositions */
    const char* names[] = {"222", "232", "322", "333", "444"};
    fmm_t algs[5];
    for (int i = 0; i < 5; i++) {
        snprintf(path, sizeof(path), "%s/%s.txt", dir, names[i]);
        algs[i] = new_fmm(path);
        CHECK(algs[i].R > 0, "%s: parse failed", path);
    }
    for (int i = 0; i < 5; i++) {
        if (algs[i].R <= 0) continue;
        check_alg(names[i], &algs[i], 1);
        fmm_t two[2] = {algs[i], algs[i]};
        if (i != 4) check_alg(names[i], two, 2);   /* 444 L2 (2401 products) runs last, once */
    }
    fmm_t mixed[2] = {algs[0], algs[2]};
    check_alg("222x322", mixed, 2);
    fmm_t mixed2[2] = {algs[1], algs[0]};
    check_alg("232x222", mixed2, 2);
    fmm_t deep[2] = {algs[4], algs[4]};
    check_alg("444", deep, 2);
    fmmhip_alg_t c = NULL;
    CHECK(fmmhip_alg_create_classical(&c) == 0 && c && c->R == 1, "classical");
    if (c) { check_entries(c, 5, 7, 3, 1); fmmhip_alg_destroy(c); }

    /* malformed and odd files: each gives R == 0 and leaks nothing */
    const char* bad[] = {
        "",                                       /* empty */
        "2 2 2\n",                                /* short header */
        "2 2 2 7\n1 0 1 0 1 -1 0\n",              /* truncated tables */
        "2 2 2 -7\n",                             /* negative R */
        "x y z w\n",                              /* not numbers */
        "2 2 2 1\n1 1 1 1\n1 1 1 1\n1 1 1 1 1\n", /* rows of the wrong length */
    };
    for (size_t i = 0; i < sizeof(bad) / sizeof(bad[0]); i++) {
        snprintf(path, sizeof(path), "%s/bad%zu.txt", tmp, i);
        write_file(path, bad[i]);
        fmm_t f = new_fmm(path);
        CHECK(f.R == 0, "malformed file %zu parsed with R = %d", i, f.R);
        free_fmm(f);
    }
    snprintf(path, sizeof(path), "%s/missing.txt", tmp);
    fmm_t f = new_fmm(path);
    CHECK(f.R == 0, "missing file parsed");
    free_fmm(f);
    /* a valid file with a 1 MB comment-free line of padding spaces */
    {
        size_t big = 1 << 20;
        char* text = (char*)malloc(big + 256);
        strcpy(text, "1 1 1 1\n1");
        size_t at = strlen(text);
        memset(text + at, ' ', big);
        at += big;
        strcpy(text + at, "\n1\n1\n");
        snprintf(path, sizeof(path), "%s/long.txt", tmp);
        write_file(path, text);
        free(text);
        fmm_t g = new_fmm(path);
        CHECK(g.R == 1 && g.U && g.U[0] == 1, "long line: R = %d", g.R);
        free_fmm(g);
    }

    for (int i = 0; i < 5; i++) free_fmm(algs[i]);
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("spec_sanitize: ok\n");
    return 0;
}
