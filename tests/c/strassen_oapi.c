/* strassen_oapi.c -- the drop-in driven the way the reference's harness
 * drives fmm-blis (test/test_strassen_oapi.c:43-224): objects over host
 * buffers, C += A*B through bli_strassen_ab_ex with the algorithm read from a
 * coefficient file, min-of-N timing, and the Frobenius residual against a
 * classical product; plus the normalised error
 * ||C - C_ref||_F / (||A||_F ||B||_F) and an independent check of sampled
 * entries by plain dot products on the host.
 *
 *   strassen_oapi m n k coef_file [repeats] [levels]
 *
 * Prints "m n k gflops_fmm gflops_classical resid normalised_error" and exits
 * 1 when the normalised error exceeds 1e-14 (or a sampled entry is off). */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "fmm_blis_amd.h"

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* uniform[-1,1) from a 64-bit linear congruential stream (host data only) */
static double next_uniform(uint64_t* s)
{
    *s = *s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(*s >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static double fro(const double* X, long n)
{
    double s = 0.0;
    for (long i = 0; i < n; i++) s += X[i] * X[i];
    return sqrt(s);
}

int main(int argc, char** argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s m n k coef_file [repeats] [levels]\n", argv[0]);
        return 2;
    }
    const long m = atol(argv[1]), n = atol(argv[2]), k = atol(argv[3]);
    const int reps = argc > 5 ? atoi(argv[5]) : 3;
    const int levels = argc > 6 ? atoi(argv[6]) : 1;
    fmm_t fmm = new_fmm(argv[4]);
    if (fmm.R == 0) {
        fprintf(stderr, "cannot read %s\n", argv[4]);
        return 2;
    }
    double* A = malloc(sizeof(double) * (size_t)m * k);
    double* B = malloc(sizeof(double) * (size_t)k * n);
    double* C = malloc(sizeof(double) * (size_t)m * n);
    double* C0 = malloc(sizeof(double) * (size_t)m * n);
    double* Cr = malloc(sizeof(double) * (size_t)m * n);
    if (!A || !B || !C || !C0 || !Cr) return 2;
    uint64_t seed = 12345;
    for (long i = 0; i < m * k; i++) A[i] = next_uniform(&seed);
    for (long i = 0; i < k * n; i++) B[i] = next_uniform(&seed);
    for (long i = 0; i < m * n; i++) C0[i] = next_uniform(&seed);

    /* column-major objects over the host buffers; alpha = beta = 1 */
    obj_t oA, oB, oC, oCr, one;
    double one_v = 1.0;
    fmm_obj_attach(BLIS_DOUBLE, m, k, A, 1, m, &oA);
    fmm_obj_attach(BLIS_DOUBLE, k, n, B, 1, k, &oB);
    fmm_obj_attach(BLIS_DOUBLE, m, n, C, 1, m, &oC);
    fmm_obj_attach(BLIS_DOUBLE, m, n, Cr, 1, m, &oCr);
    fmm_obj_scalar(&one_v, &one);
    if (bli_plugin_register_fmm_blis() != BLIS_SUCCESS) {
        fprintf(stderr, "registration failed\n");
        return 2;
    }
    fmm_t algs[4] = {fmm, fmm, fmm, fmm};
    if (levels < 1 || levels > 4) return 2;

    double t_fmm = 1e30, t_cls = 1e30;
    for (int r = 0; r < reps; r++) {
        memcpy(C, C0, sizeof(double) * (size_t)m * n);
        double t0 = now_s();
        if (levels == 1) bli_strassen_ab_ex(&one, &oA, &oB, &one, &oC, fmm);
        else bli_fmm_ab_ex(&one, &oA, &oB, &one, &oC, algs, levels, FMMHIP_LEAF_NATIVE);
        double t = now_s() - t0;
        if (t < t_fmm) t_fmm = t;
    }
    for (int r = 0; r < reps; r++) {
        memcpy(Cr, C0, sizeof(double) * (size_t)m * n);
        double t0 = now_s();
        bli_gemm_hip(&one, &oA, &oB, &one, &oCr);
        double t = now_s() - t0;
        if (t < t_cls) t_cls = t;
    }

    double d2 = 0.0;
    for (long i = 0; i < m * n; i++) d2 += (C[i] - Cr[i]) * (C[i] - Cr[i]);
    const double resid = sqrt(d2);
    const double nerr = resid / (fro(A, m * k) * fro(B, k * n));
    /* independent host check of 64 sampled entries */
    int bad = 0;
    uint64_t ps = 99;
    for (int t = 0; t < 64 && m > 0 && n > 0; t++) {
        ps = ps * 6364136223846793005ull + 1442695040888963407ull;
        const long i = (long)((ps >> 33) % (uint64_t)m), j = (long)((ps >> 13) % (uint64_t)n);
        double s = C0[i + j * m], sa = fabs(C0[i + j * m]);
        for (long p = 0; p < k; p++) {
            s += A[i + p * m] * B[p + j * k];
            sa += fabs(A[i + p * m] * B[p + j * k]);
        }
        if (fabs(s - C[i + j * m]) > 1e-13 * (sa + 1.0)) bad = 1;
    }
    const double gf = 2.0 * (double)m * (double)n * (double)k * 1e-9;
    printf("%5ld\t %5ld\t %5ld\t %8.2lf\t %8.2lf\t %9.3g\t %9.3g\n", m, n, k, gf / t_fmm, gf / t_cls, resid, nerr);
    free(A); free(B); free(C); free(C0); free(Cr);
    free_fmm(fmm);
    return (nerr > 1e-14 || bad) ? 1 : 0;
}
