/*
 * fmm_oracle.c -- CPU restatement of the fmm-blis fast-matrix-multiplication
 * path, used ONLY as test infrastructure.
 *
 * This file is the checker.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only to check or time the
 * product, never as the product path.  The product (fmm-blis_amd/) does not
 * link, include or call anything here.
 *
 * What it restates (citations are files of dnparikh/fmm-blis):
 *   orc_parse_alg       frame/base/bli_read_fmm.c:24-80 (the "m k n R" header,
 *                       U rows, '#', V rows, '#', W rows; '#' lines are
 *                       comments).  Two defects of the reference are not
 *                       restated: the header is read in file order m,k,n
 *                       (the reference reads m,n,k at :33) and lines have no
 *                       length limit (LINE_N 128 at :7 overflows on 444.txt).
 *   orc_part_geometry   frame/base/bli_strassen_ab.c:86-121 (init_part_offsets:
 *                       part i has floor(whole/tilde) + (i < whole mod tilde)
 *                       rows; offsets are prefix sums; row-major part index).
 *   orc_compose         no counterpart (the reference has one level only);
 *                       the L-level algorithm is the Kronecker product of the
 *                       per-level tables, part index (i1*mt2+i2)*(kt1*kt2)+...
 *   orc_fmm_dgemm       frame/base/bli_strassen_ab.c:248-294 (one pass per
 *                       product r) with plugin/ref_kernels/bli_packm_fmm_ref.c:
 *                       89-147 (S_r = sum_s U[s][r] * A_s, clipped to each
 *                       part's extent, zero padding) and
 *                       plugin/ref_kernels/bli_gemm_fmm_ref.c:85-120 (leaf
 *                       product, then C_s += W[s][r] * M_r clipped to the part).
 *                       Semantics are BLAS C <- beta*C + alpha*A*B; on the
 *                       reference's valid domain (alpha = beta = 1) this is
 *                       what the reference computes.
 *   orc_dgemm           classical product, the role BLIS bli_gemm plays as the
 *                       reference's oracle (test/test_strassen_oapi.c:180-181,
 *                       my_mm at :18-42).
 *   orc_fill_uniform    seeded counter-based SplitMix64 -> uniform [-1, 1).
 *
 * Pinning (see tests/test_oracle.py): Brent equations on every coefficient
 * file and composed table (exact integers); orc_dgemm against numpy.matmul
 * and a brute-force triple loop; orc_fmm_dgemm against orc_dgemm bit-exactly
 * on small-integer inputs for every algorithm, level and uneven shape, and
 * against the textbook Strassen formulas; geometry against closed forms.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define ORC_MAX_LINE (1 << 20)

typedef struct {
    int mt, kt, nt, R;        /* <m~, k~, n~, R> */
    int *U, *V, *W;           /* row-major [part][r] (bli_strassen_ab.c:6-8) */
} orc_alg_t;

/* ---- algorithm tables ------------------------------------------------- */

static int parse_ints(const char *s, int *out, int max)
{
    int n = 0;
    while (*s) {
        char *end;
        long v = strtol(s, &end, 10);
        if (end == s) { s++; continue; }
        if (n < max) out[n] = (int)v;
        n++;
        s = end;
    }
    return n;
}

/* Returns 0 on success, negative on a malformed file. */
int orc_parse_alg(const char *path, orc_alg_t *alg)
{
    memset(alg, 0, sizeof(*alg));
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char *line = (char *)malloc(ORC_MAX_LINE);
    int hdr[4], rc = 0, row = 0, aparts = 0, bparts = 0, cparts = 0;
    if (!fgets(line, ORC_MAX_LINE, fp) || parse_ints(line, hdr, 4) != 4) { rc = -2; goto out; }
    alg->mt = hdr[0]; alg->kt = hdr[1]; alg->nt = hdr[2]; alg->R = hdr[3];
    if (alg->mt <= 0 || alg->kt <= 0 || alg->nt <= 0 || alg->R <= 0) { rc = -3; goto out; }
    aparts = alg->mt * alg->kt; bparts = alg->kt * alg->nt; cparts = alg->mt * alg->nt;
    alg->U = (int *)calloc((size_t)aparts * alg->R, sizeof(int));
    alg->V = (int *)calloc((size_t)bparts * alg->R, sizeof(int));
    alg->W = (int *)calloc((size_t)cparts * alg->R, sizeof(int));
    int *tmp = (int *)malloc(sizeof(int) * (size_t)(alg->R + 1));
    while (fgets(line, ORC_MAX_LINE, fp)) {
        if (line[0] == '#') continue;
        int cnt = parse_ints(line, tmp, alg->R + 1);
        if (cnt == 0) continue;                 /* blank line */
        if (cnt != alg->R) { rc = -4; break; }
        int *dst;
        if (row < aparts) dst = alg->U + (size_t)row * alg->R;
        else if (row < aparts + bparts) dst = alg->V + (size_t)(row - aparts) * alg->R;
        else if (row < aparts + bparts + cparts) dst = alg->W + (size_t)(row - aparts - bparts) * alg->R;
        else { rc = -5; break; }
        memcpy(dst, tmp, sizeof(int) * alg->R);
        row++;
    }
    free(tmp);
    if (rc == 0 && row != aparts + bparts + cparts) rc = -6;
out:
    free(line);
    fclose(fp);
    return rc;
}

void orc_free_alg(orc_alg_t *alg)
{
    free(alg->U); free(alg->V); free(alg->W);
    memset(alg, 0, sizeof(*alg));
}

/* Kronecker composition of two levels: outer a, inner b.  The composed
 * block (i, p) of A with i = i1*b.mt + i2, p = p1*b.kt + p2 is block (i2,p2)
 * of block (i1,p1); product r = r1*b.R + r2. */
static void kron(const int *X1, int r1n, int rows1, int cols1,
                 const int *X2, int r2n, int rows2, int cols2, int *out)
{
    int R = r1n * r2n, cols = cols1 * cols2;
    for (int i1 = 0; i1 < rows1; i1++)
    for (int c1 = 0; c1 < cols1; c1++)
    for (int i2 = 0; i2 < rows2; i2++)
    for (int c2 = 0; c2 < cols2; c2++) {
        int part = (i1 * rows2 + i2) * cols + (c1 * cols2 + c2);
        int p1 = i1 * cols1 + c1, p2 = i2 * cols2 + c2;
        for (int r1 = 0; r1 < r1n; r1++)
        for (int r2 = 0; r2 < r2n; r2++)
            out[(size_t)part * R + r1 * r2n + r2] = X1[p1 * r1n + r1] * X2[p2 * r2n + r2];
    }
}

int orc_compose(const orc_alg_t *a, const orc_alg_t *b, orc_alg_t *out)
{
    memset(out, 0, sizeof(*out));
    out->mt = a->mt * b->mt; out->kt = a->kt * b->kt; out->nt = a->nt * b->nt;
    out->R = a->R * b->R;
    out->U = (int *)calloc((size_t)out->mt * out->kt * out->R, sizeof(int));
    out->V = (int *)calloc((size_t)out->kt * out->nt * out->R, sizeof(int));
    out->W = (int *)calloc((size_t)out->mt * out->nt * out->R, sizeof(int));
    if (!out->U || !out->V || !out->W) return -1;
    kron(a->U, a->R, a->mt, a->kt, b->U, b->R, b->mt, b->kt, out->U);
    kron(a->V, a->R, a->kt, a->nt, b->V, b->R, b->kt, b->nt, out->V);
    kron(a->W, a->R, a->mt, a->nt, b->W, b->R, b->mt, b->nt, out->W);
    return 0;
}

/* Brent equations: sum_r U[(i,p)][r] V[(p',j)][r] W[(i',j')][r]
 *   = [i==i'][p==p'][j==j'].  Returns the number of violated equations. */
long orc_brent_violations(const orc_alg_t *g)
{
    long bad = 0;
    int mt = g->mt, kt = g->kt, nt = g->nt, R = g->R;
    for (int i = 0; i < mt; i++) for (int p = 0; p < kt; p++)
    for (int q = 0; q < kt; q++) for (int j = 0; j < nt; j++)
    for (int ii = 0; ii < mt; ii++) for (int jj = 0; jj < nt; jj++) {
        long s = 0;
        const int *u = g->U + (size_t)(i * kt + p) * R;
        const int *v = g->V + (size_t)(q * nt + j) * R;
        const int *w = g->W + (size_t)(ii * nt + jj) * R;
        for (int r = 0; r < R; r++) s += (long)u[r] * v[r] * w[r];
        long want = (i == ii && p == q && j == jj) ? 1 : 0;
        if (s != want) bad++;
    }
    return bad;
}

/* ---- geometry (bli_strassen_ab.c:86-121) ------------------------------- */

/* Split `whole` into `tilde` contiguous parts; part t has size
 * floor(whole/tilde) + (t < whole mod tilde) and starts at the prefix sum. */
void orc_part_geometry(long whole, int tilde, long *size, long *off)
{
    long base = whole / tilde, extra = whole % tilde, o = 0;
    for (int t = 0; t < tilde; t++) {
        size[t] = base + (t < extra ? 1 : 0);
        off[t] = o;
        o += size[t];
    }
}

/* ---- classical product -------------------------------------------------- */

/* C <- beta*C + alpha*A*B, column-major, A m x k (lda), B k x n (ldb).
 * beta == 0 overwrites C without reading it (BLAS convention). */
void orc_dgemm(long m, long n, long k, double alpha, const double *A, long lda,
               const double *B, long ldb, double beta, double *C, long ldc)
{
    for (long j = 0; j < n; j++) {
        double *c = C + j * ldc;
        if (beta == 0.0) for (long i = 0; i < m; i++) c[i] = 0.0;
        else if (beta != 1.0) for (long i = 0; i < m; i++) c[i] *= beta;
    }
    if (alpha == 0.0 || k == 0) return;
    /* i-blocked j-p-i order: each column of C accumulates A's columns;
     * columns are independent, so OpenMP splits them across threads. */
    const long IB = 256;
    #pragma omp parallel
    {
        double acc[256];
        #pragma omp for schedule(static)
        for (long j = 0; j < n; j++) {
            for (long i0 = 0; i0 < m; i0 += IB) {
                long ib = (m - i0 < IB) ? m - i0 : IB;
                for (long i = 0; i < ib; i++) acc[i] = 0.0;
                for (long p = 0; p < k; p++) {
                    double b = B[p + j * ldb];
                    const double *a = A + i0 + p * lda;
                    for (long i = 0; i < ib; i++) acc[i] += a[i] * b;
                }
                double *c = C + i0 + j * ldc;
                for (long i = 0; i < ib; i++) c[i] += alpha * acc[i];
            }
        }
    }
}

int orc_threads(void)
{
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Naive triple loop with the summation order of test_strassen_oapi.c:34-40
 * (p innermost), for tiny cross-checks of orc_dgemm. */
void orc_dgemm_naive(long m, long n, long k, const double *A, long lda,
                     const double *B, long ldb, double *C, long ldc)
{
    for (long i = 0; i < m; i++)
        for (long j = 0; j < n; j++)
            for (long p = 0; p < k; p++)
                C[i + j * ldc] += A[i + p * lda] * B[p + j * ldb];
}

/* ---- FMM restatement ---------------------------------------------------- */

/* C <- beta*C + alpha*A*B with the fast algorithm g (already composed to the
 * wanted number of levels), column-major.  Per product r:
 *   S   = sum_{(i,p)} U[(i,p)][r] * A_{ip}   (leaf lm x lk, zero padded)
 *   T   = sum_{(p,j)} V[(p,j)][r] * B_{pj}   (leaf lk x ln, zero padded)
 *   M   = S * T
 *   C_{ij} += alpha * W[(i,j)][r] * M        (clipped to the part's extent)
 * Zero coefficients are skipped (the reference packs and scatters them with
 * coefficient 0, bli_strassen_ab.c:255-283, which adds exact zeros).
 * Returns 0, or -1 on allocation failure. */
int orc_fmm_dgemm(const orc_alg_t *g, long m, long n, long k, double alpha,
                  const double *A, long lda, const double *B, long ldb,
                  double beta, double *C, long ldc)
{
    int mt = g->mt, kt = g->kt, nt = g->nt, R = g->R;
    long *szm = calloc(mt, sizeof(long)), *ofm = calloc(mt, sizeof(long));
    long *szk = calloc(kt, sizeof(long)), *ofk = calloc(kt, sizeof(long));
    long *szn = calloc(nt, sizeof(long)), *ofn = calloc(nt, sizeof(long));
    orc_part_geometry(m, mt, szm, ofm);
    orc_part_geometry(k, kt, szk, ofk);
    orc_part_geometry(n, nt, szn, ofn);
    long lm = (m + mt - 1) / mt, lk = (k + kt - 1) / kt, ln = (n + nt - 1) / nt;

    for (long j = 0; j < n; j++) {
        double *c = C + j * ldc;
        if (beta == 0.0) for (long i = 0; i < m; i++) c[i] = 0.0;
        else if (beta != 1.0) for (long i = 0; i < m; i++) c[i] *= beta;
    }
    int rc = 0;
    if (alpha == 0.0 || k == 0 || m == 0 || n == 0) goto done;

    double *S = malloc(sizeof(double) * (size_t)(lm * lk + 1));
    double *T = malloc(sizeof(double) * (size_t)(lk * ln + 1));
    double *M = malloc(sizeof(double) * (size_t)(lm * ln + 1));
    if (!S || !T || !M) { free(S); free(T); free(M); rc = -1; goto done; }

    for (int r = 0; r < R; r++) {
        /* pack: S_r and T_r (bli_packm_fmm_ref.c:89-147) */
        memset(S, 0, sizeof(double) * (size_t)(lm * lk));
        for (int i = 0; i < mt; i++) for (int p = 0; p < kt; p++) {
            int u = g->U[(size_t)(i * kt + p) * R + r];
            if (!u) continue;
            for (long y = 0; y < szk[p]; y++)
                for (long x = 0; x < szm[i]; x++)
                    S[x + y * lm] += (double)u * A[(ofm[i] + x) + (ofk[p] + y) * lda];
        }
        memset(T, 0, sizeof(double) * (size_t)(lk * ln));
        for (int p = 0; p < kt; p++) for (int j = 0; j < nt; j++) {
            int v = g->V[(size_t)(p * nt + j) * R + r];
            if (!v) continue;
            for (long y = 0; y < szn[j]; y++)
                for (long x = 0; x < szk[p]; x++)
                    T[x + y * lk] += (double)v * B[(ofk[p] + x) + (ofn[j] + y) * ldb];
        }
        /* leaf product (bli_gemm_fmm_ref.c:85-97) */
        orc_dgemm(lm, ln, lk, 1.0, S, lm, T, lk, 0.0, M, lm);
        /* W scatter (bli_gemm_fmm_ref.c:101-120) */
        for (int i = 0; i < mt; i++) for (int j = 0; j < nt; j++) {
            int w = g->W[(size_t)(i * nt + j) * R + r];
            if (!w) continue;
            double lam = alpha * (double)w;
            for (long y = 0; y < szn[j]; y++)
                for (long x = 0; x < szm[i]; x++)
                    C[(ofm[i] + x) + (ofn[j] + y) * ldc] += lam * M[x + y * lm];
        }
    }
    free(S); free(T); free(M);
done:
    free(szm); free(ofm); free(szk); free(ofk); free(szn); free(ofn);
    return rc;
}

/* ---- seeded inputs ------------------------------------------------------ */

/* SplitMix64 finaliser (Steele, Lea, Flood 2014). */
static inline uint64_t splitmix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Element (i, j) of an m x n column-major matrix (leading dimension ld) is
 * drawn from the counter idx = j*m + i of stream `stream` under `seed`:
 *   base = splitmix64(seed ^ stream*0xD1B54A32D192ED03)
 *   z    = splitmix64(base + (idx+1)*0x9E3779B97F4A7C15)
 * and the value is 2u - 1 with u = (z >> 11) * 2^-53, in [-1, 1).  Every
 * step is exact, so a GPU fill produces identical bits. */
static inline uint64_t orc_bits_at(uint64_t seed, uint64_t stream, uint64_t idx)
{
    uint64_t base = splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
    return splitmix64(base + (idx + 1) * 0x9E3779B97F4A7C15ull);
}

static inline double orc_uniform_at(uint64_t seed, uint64_t stream, uint64_t idx)
{
    double u = (double)(orc_bits_at(seed, stream, idx) >> 11) * 0x1.0p-53;
    return 2.0 * u - 1.0;
}

void orc_fill_uniform(double *X, long m, long n, long ld, uint64_t seed, uint64_t stream)
{
    for (long j = 0; j < n; j++)
        for (long i = 0; i < m; i++)
            X[i + j * ld] = orc_uniform_at(seed, stream, (uint64_t)(j * m + i));
}

/* Small-integer fill in [-amp, amp] from the same counter stream
 * (value = ((z >> 11) mod (2*amp+1)) - amp), used for the exact
 * (bit-for-bit) cross-checks: with such inputs every partial sum of every
 * algorithm is an integer below 2^53, so all orders of summation agree. */
void orc_fill_int(double *X, long m, long n, long ld, uint64_t seed, uint64_t stream, int amp)
{
    uint64_t mod = (uint64_t)(2 * amp + 1);
    for (long j = 0; j < n; j++)
        for (long i = 0; i < m; i++) {
            uint64_t z = orc_bits_at(seed, stream, (uint64_t)(j * m + i)) >> 11;
            X[i + j * ld] = (double)((long)(z % mod) - amp);
        }
}

/* Single-precision copy of orc_fill_uniform's values (rounded to nearest),
 * the inputs of the FP32 forms. */
void orc_fill_uniform_f32(float *X, long m, long n, long ld, uint64_t seed, uint64_t stream)
{
    for (long j = 0; j < n; j++)
        for (long i = 0; i < m; i++)
            X[i + j * ld] = (float)orc_uniform_at(seed, stream, (uint64_t)(j * m + i));
}

/* Frobenius-norm helpers for the normalised error
 * ||C - C_ref||_F / (||A||_F * ||B||_F). */
double orc_fro_diff(const double *X, const double *Y, long m, long n, long ldx, long ldy)
{
    double s = 0.0;
    for (long j = 0; j < n; j++)
        for (long i = 0; i < m; i++) {
            double d = X[i + j * ldx] - Y[i + j * ldy];
            s += d * d;
        }
    return sqrt(s);
}

double orc_fro(const double *X, long m, long n, long ld)
{
    double s = 0.0;
    for (long j = 0; j < n; j++)
        for (long i = 0; i < m; i++) s += X[i + j * ld] * X[i + j * ld];
    return sqrt(s);
}
