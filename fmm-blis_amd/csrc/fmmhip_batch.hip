// fmmhip_batch.hip -- the batching form and its autotuner.
//
// Many independent column-major GEMMs (FP64 and FP32) run as one captured
// HIP graph: per datatype, one batched summation launch (the pre-summed
// operand sums of every FMM problem), persistent MFMA launches (per tile
// kernel) whose workgroups take (problem, tile, product group) items from
// device queues until they are empty, and the W stages of the FMM problems
// (their products materialised).  The algorithm, level and tile of every
// problem are chosen by the tuner below.  No counterpart in the reference (its driver runs one
// product at a time, frame/base/bli_strassen_ab.c:123-295).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>
#include <stdio.h>
#include <string.h>

#include "fmmhip_host.h"

using fmmk::BatchItem;

namespace {

// Whether a problem prepared for the general 128 x 128 kernel can run its
// interior tiles on the fast kernel: the fast kernel's layouts, 16-byte
// aligned pairs, no k-split, product groups only into the materialised
// slots (no per-group copies of C), every part with the leaf's full k in
// whole k-steps of the leaf type (d: FP64).
bool fast_tiles_ok(const Prepared& P, int d)
{
    const fmmk::KArgs& a = P.ka;
    const int bk = d ? fmmk::fast_bk<double>() : fmmk::fast_bk<float>();
    return P.a_mn && !P.b_mn && P.c_col && a.va && a.vb && !a.ksplit && (a.groups == 1 || P.wmat) &&
           a.lk % bk == 0 && a.lk >= bk && a.min_k == a.lk;
}

// Whether tile (tm, tn) lies inside every part of A, B and C (the fast
// kernel's loads and its unclipped W epilogue need no bounds there).
bool fast_tile_inside(const fmmk::KArgs& a, long tm, long tn)
{
    constexpr long T = fmmk::FastCfg::T;
    return (tm + 1) * T <= std::min(a.min_m, a.min_cm) && (tn + 1) * T <= std::min(a.min_n, a.min_cn);
}

}  // namespace
using fmmk::KArgs;

namespace {

// ---- tuner -------------------------------------------------------------------

// Candidates: 0 classical, 1 one-level Strassen, 2 two-level Strassen (both
// pre-summed, products materialised), each on one of two tiles: 128 x 128
// (launch shape 0) or 64 x 64 (launch shape 2).  Cost model of one
// candidate and tile on one problem:
//   t = F_leaf * pad / (P * eff * teff * nk/(nk + ovh)) * quant + bytes_sum / bw
// F_leaf the leaf flops, pad the tile grid's padding (tiles * T^2 over the
// leaf's lm * ln: the kernel runs whole tiles), P the MFMA peak of the
// datatype, eff the 128-tile kernel's sustained fraction of peak, teff the
// tile's rate relative to it on whole tiles, nk the 16-wide k steps per
// product, ovh the per-product prologue/epilogue cost in k steps, quant the
// wave-quantisation loss of a lone launch without product split (1 inside a
// batch), bytes_sum the traffic of the summation and W stages and bw their
// bandwidth.
constexpr int kCands = 3;
constexpr int kTiles = 2;                         // 0: 128 x 128, 1: 64 x 64
const int kTileShape[kTiles] = {0, 2};            // their kernel launch shapes
const int kTileWgPerCu[kTiles] = {2, 4};          // resident workgroups per CU
struct TunerParams {
    double eff[2][kCands];      // [is_double][cand]
    double ovh[2][kCands];
    double bw;                  // bandwidth of the summation and W stages, bytes/s
    double launch_ms;           // fixed cost per launch
    int cus;
    double teff[kTiles];        // rate of each tile's kernel relative to the 128 x 128 one (whole tiles)
};

std::mutex g_tuner_mu;
TunerParams g_tp = {
    // defaults: fmmhip_tuner_calibrate on an MI355X (FP32 row, then FP64;
    // 128 x 128 tiles on the fast kernel wherever it applies, FMM products
    // materialised)
    {{0.856, 0.859, 0.861}, {0.929, 0.925, 0.921}},
    {{4.0, 4.0, 4.0}, {4.0, 4.0, 4.0}},
    4.09e12, 0.01, 256,
    {1.0, 0.88}};                // 64 x 64 tiles against the fast kernel's 128 x 128 ones, classical FP64 16384^3

const double kPeak[2] = {157.3e12, 78.6e12};   // FP32, FP64 MFMA peaks (FLOP/s)

struct CandShape {
    int mt, kt, nt, R;
};
const CandShape kShape[kCands] = {{1, 1, 1, 1}, {2, 2, 2, 7}, {4, 4, 4, 49}};

double est_ms(const TunerParams& tp, int is_double, int cand, long m, long n, long k, bool lone, int tile = 0)
{
    const CandShape& c = kShape[cand];
    const double lm = (double)((m + c.mt - 1) / c.mt), lk = (double)((k + c.kt - 1) / c.kt),
                 ln = (double)((n + c.nt - 1) / c.nt);
    const double T = (double)fmmk::tile_size(kTileShape[tile]);
    const double tiles = std::ceil(lm / T) * std::ceil(ln / T);
    const double fl = 2.0 * c.R * lk * tiles * T * T;      // the tile grid's flops, padding included
    const double nk = std::ceil(lk / fmmk::BK);
    double eff = tp.eff[is_double][cand] * tp.teff[tile] * nk / (nk + tp.ovh[is_double][cand]);
    if (lone) {
        const double slots = (double)kTileWgPerCu[tile] * tp.cus;
        const double waves = std::ceil(tiles / slots);
        eff *= tiles / (waves * slots);
    }
    double t = fl / (kPeak[is_double] * eff);
    if (cand > 0) {
        // summation (A and B read, R leaves of each written) and W stage
        // (the products materialised: R slots read, C read and written)
        const double esz = is_double ? 8.0 : 4.0;
        const double bytes =
            esz * ((double)m * k + c.R * lm * lk + (double)k * n + c.R * lk * ln + c.R * lm * ln + 2.0 * m * n);
        t += bytes / tp.bw;
    }
    return t * 1e3 + (lone ? tp.launch_ms * (cand > 0 ? 3 : 1) : 0.0);
}

// An FMM candidate must beat the best simpler one by a factor: the model's
// error is of the order of a few percent, and a wrong FMM choice costs more
// (workspace, a summation launch, long work items) than a wrong classical one.
// Where the candidate's leaves are whole 128 x 128 tiles and whole fast-kernel
// k-steps (every tile on the fast kernel, the regime the calibration times)
// the model came within 4% of the measured batches, on the cautious side for
// FMM (FP64, classical / two-level Strassen: 4 x 12288^3 204 / 176 ms
// estimated, 203 / 170 measured; 8 x 8192^3 121 / 110, 121 / 106;
// 16 x 6144^3 103 / 97, 103 / 94) and the factor is 0.97; elsewhere edge
// tiles run the general kernel, which the model does not price, and the
// factor is 0.95.
constexpr double kFmmMargin = 0.95, kFmmMarginWhole = 0.97;

bool whole_fast_leaves(int is_double, int cand, long m, long n, long k)
{
    const CandShape& s = kShape[cand];
    const long kstep = is_double ? 16 : 32;
    return m % (s.mt * fmmk::BM) == 0 && n % (s.nt * fmmk::BN) == 0 && k % (s.kt * kstep) == 0;
}

// The faster tile for candidate `cand` (ties keep the 128 x 128 tile).  A
// lone classical call with at most half the CUs covered by 128 x 128 tiles
// splits k instead (one-problem path), which fills the device better.
int best_tile(const TunerParams& tp, int is_double, int cand, long m, long n, long k, bool lone, double* ms)
{
    int best = 0;
    double bt = est_ms(tp, is_double, cand, m, n, k, lone, 0);
    const long T0 = fmmk::tile_size(kTileShape[0]);
    if (lone && cand == 0 && 2 * ((m + T0 - 1) / T0) * ((n + T0 - 1) / T0) <= tp.cus) {
        if (ms) *ms = bt;
        return 0;
    }
    for (int t = 1; t < kTiles; t++) {
        const double e = est_ms(tp, is_double, cand, m, n, k, lone, t);
        if (e < bt) { bt = e; best = t; }
    }
    if (ms) *ms = bt;
    return best;
}

int choose(int is_double, long m, long n, long k, bool lone, double* best_ms, int* tile_out = nullptr)
{
    TunerParams tp;
    {
        std::lock_guard<std::mutex> lk(g_tuner_mu);
        tp = g_tp;
    }
    int best = 0, best_t = 0;
    double bt = 1e300, tmin = 1e300;   // the chosen estimate; the least over every simpler candidate
    for (int c = 0; c < kCands; c++) {
        // a level needs leaves of at least two tiles and k of 8 BK-steps
        const CandShape& s = kShape[c];
        if (c > 0 && (m < 2L * fmmk::BM * s.mt || n < 2L * fmmk::BN * s.nt || k < 8L * fmmk::BK * s.kt)) continue;
        double t = 0.0;
        const int tile = best_tile(tp, is_double, c, m, n, k, lone, &t);
        const double margin = tile == 0 && whole_fast_leaves(is_double, c, m, n, k) ? kFmmMarginWhole : kFmmMargin;
        if (c == 0 || t < margin * tmin) { bt = t; best = c; best_t = tile; }
        tmin = std::min(tmin, t);
    }
    if (best_ms) *best_ms = bt;
    if (tile_out) *tile_out = best_t;
    return best;
}

// The candidates' algorithms (shared by all batches).
fmmhip_alg_t cand_alg(int cand)
{
    static std::once_flag once;
    static fmmhip_alg_t algs[kCands] = {nullptr, nullptr, nullptr};
    std::call_once(once, [] {
        algs[0] = classical_alg();
        fmmhip_alg_create(&algs[1], &STRASSEN_FMM, 1);
        fmm_t two[2] = {STRASSEN_FMM, STRASSEN_FMM};
        fmmhip_alg_create(&algs[2], two, 2);
    });
    return algs[cand];
}

}  // namespace

extern "C" int fmmhip_tuner_choose(int is_double, long m, long n, long k, int lone, int* cand, double* est)
{
    if (m <= 0 || n <= 0 || k <= 0 || !cand) return FMMHIP_ERR_INVALID;
    *cand = choose(is_double ? 1 : 0, m, n, k, lone != 0, est);
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_tuner_choose_tile(int is_double, long m, long n, long k, int lone, int* cand, int* tile,
                                        double* est)
{
    if (m <= 0 || n <= 0 || k <= 0 || !cand || !tile) return FMMHIP_ERR_INVALID;
    int t = 0;
    *cand = choose(is_double ? 1 : 0, m, n, k, lone != 0, est, &t);
    *tile = fmmk::tile_size(kTileShape[t]);
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_tuner_estimate(int is_double, int cand, long m, long n, long k, int lone, double* est)
{
    if (cand < 0 || cand >= kCands || !est) return FMMHIP_ERR_INVALID;
    TunerParams tp;
    {
        std::lock_guard<std::mutex> lk(g_tuner_mu);
        tp = g_tp;
    }
    (void)best_tile(tp, is_double ? 1 : 0, cand, m, n, k, lone != 0, est);   // on the candidate's faster tile
    return FMMHIP_SUCCESS;
}

// Times each candidate at 16384^3 (leaves of 16384, 8192 and 4096: 16384,
// 4096 and 1024 tiles of 128x128, whole waves of the device's workgroup
// slots) with the library's launch timing, and refits its sustained
// efficiency (at the fixed prologue cost ovh) and the bandwidth of the
// summation and W stages (FMM candidates run with their products
// materialised, as in the batch and the server).
// Large, device-filling launches give the steady-state efficiency that a
// batch (no tails) and a busy server see; smaller calibration sizes folded
// the launch tail into the classical figure and made the model favour FMM
// where it measured slower.  The 64 x 64 tile's relative rate is the
// classical FP64 kernel's time on 128 x 128 tiles over its time on
// 64 x 64 tiles at the same size (65536 tiles: whole waves too).
extern "C" int fmmhip_tuner_calibrate(void* stream)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)fmmhip_default_stream();
    int cus = 256;
    fmmhip_device_cu_count(&cus);
    const long nmax = 4096L << (kCands - 1);
    TunerParams tp;
    {
        std::lock_guard<std::mutex> lk(g_tuner_mu);
        tp = g_tp;
    }
    tp.cus = cus;
    double bw_sum = 0.0, bw_n = 0.0;
    double cls_ms = 0.0;              // classical FP64, 128 x 128 tiles, per launch
    for (int dbl = 0; dbl < 2 && !rc; dbl++) {
        const size_t esz = dbl ? 8 : 4;
        void *A = nullptr, *B = nullptr, *C = nullptr;
        if (fmmhip_malloc(&A, esz * nmax * nmax) || fmmhip_malloc(&B, esz * nmax * nmax) ||
            fmmhip_malloc(&C, esz * nmax * nmax)) {
            fmmhip_free(A); fmmhip_free(B); fmmhip_free(C);
            return FMMHIP_ERR_NOMEM;
        }
        if (dbl) {
            fmmhip_fill_uniform((double*)A, nmax, nmax, nmax, 7, 0, st);
            fmmhip_fill_uniform((double*)B, nmax, nmax, nmax, 7, 1, st);
            fmmhip_fill_uniform((double*)C, nmax, nmax, nmax, 7, 2, st);
        } else {
            fmmhip_fill_uniform_f32((float*)A, nmax, nmax, nmax, 7, 0, st);
            fmmhip_fill_uniform_f32((float*)B, nmax, nmax, nmax, 7, 1, st);
            fmmhip_fill_uniform_f32((float*)C, nmax, nmax, nmax, 7, 2, st);
        }
        for (int cand = 0; cand < kCands && !rc; cand++) {
            fmmhip_alg_t alg = cand_alg(cand);
            const int saved_pack = alg->pack, saved_split = alg->split;
            alg->pack = cand > 0 ? 2 : 0;   // as the batch and the server run them
            alg->split = cand > 0 ? -1 : 1;
            const long nn = nmax;
            for (int rep = 0; rep < 3 && !rc; rep++) {
                if (rep == 1) fmmhip_timing_enable(1);
                if (dbl)
                    rc = fmmhip_dgemm_fmm(alg, FMMHIP_LEAF_NATIVE, nn, nn, nn, 1.0, (double*)A, 1, nmax, (double*)B,
                                          1, nmax, 1.0, (double*)C, 1, nmax, st);
                else
                    rc = fmmhip_sgemm_fmm(alg, nn, nn, nn, 1.0f, (float*)A, 1, nmax, (float*)B, 1, nmax, 1.0f,
                                          (float*)C, 1, nmax, st);
            }
            fmmhip_timing_enable(0);
            alg->pack = saved_pack;
            alg->split = saved_split;
            long nl = 0;
            double kms = 0, pms = 0, mx = 0;
            fmmhip_timing_read(&nl, &kms, &pms, &mx);
            if (rc || nl <= 0) break;
            if (dbl && cand == 0) cls_ms = kms / nl;
            const CandShape& c = kShape[cand];
            const double leaf = (double)nn / c.mt;
            const double fl = 2.0 * c.R * leaf * leaf * leaf;
            const double nk = std::ceil(leaf / fmmk::BK);
            const double e = fl / (kPeak[dbl] * (kms / nl) * 1e-3);     // = eff * nk / (nk + ovh)
            double eff = e * (nk + tp.ovh[dbl][cand]) / nk;
            if (eff > 0.98) eff = 0.98;
            if (eff > 0.05) tp.eff[dbl][cand] = eff;
            if (cand > 0 && pms > 0) {
                // the timed stages: the summation (A and B read, R leaves of
                // each written) and the W stage (R slots read, C read and written)
                const double bytes = (double)esz * (2.0 * nn * nn + 2.0 * c.R * leaf * leaf + c.R * leaf * leaf +
                                                    2.0 * nn * nn);
                bw_sum += bytes / (pms / nl * 1e-3);
                bw_n += 1;
            }
        }
        if (!rc && dbl && cls_ms > 0.0) {
            fmmhip_alg_t a64 = nullptr;
            rc = fmmhip_alg_create_classical(&a64);
            if (!rc) rc = fmmhip_alg_set_shape(a64, 2);
            if (!rc) rc = fmmhip_alg_set_split(a64, 1);
            for (int rep = 0; rep < 3 && !rc; rep++) {
                if (rep == 1) fmmhip_timing_enable(1);
                rc = fmmhip_dgemm_fmm(a64, FMMHIP_LEAF_NATIVE, nmax, nmax, nmax, 1.0, (double*)A, 1, nmax,
                                      (double*)B, 1, nmax, 1.0, (double*)C, 1, nmax, st);
            }
            fmmhip_timing_enable(0);
            long nl = 0;
            double kms = 0, pms = 0, mx = 0;
            fmmhip_timing_read(&nl, &kms, &pms, &mx);
            if (!rc && nl > 0 && kms > 0) tp.teff[1] = std::min(1.5, std::max(0.3, cls_ms / (kms / nl)));
            fmmhip_alg_destroy(a64);
        }
        fmmhip_free(A); fmmhip_free(B); fmmhip_free(C);
    }
    if (rc) return rc;
    if (bw_n > 0) tp.bw = bw_sum / bw_n;
    std::lock_guard<std::mutex> lk(g_tuner_mu);
    g_tp = tp;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_tuner_save(const char* path)
{
    if (!path) return FMMHIP_ERR_INVALID;
    TunerParams tp;
    {
        std::lock_guard<std::mutex> lk(g_tuner_mu);
        tp = g_tp;
    }
    FILE* f = fopen(path, "w");
    if (!f) return FMMHIP_ERR_IO;
    fprintf(f, "# fmm_blis_amd tuner table: candidate dtype eff ovh (candidates 0 classical, 1 222x1, 2 222x2)\n");
    for (int d = 0; d < 2; d++)
        for (int c = 0; c < kCands; c++) fprintf(f, "cand %d %s %.6f %.6f\n", c, d ? "f64" : "f32", tp.eff[d][c], tp.ovh[d][c]);
    fprintf(f, "bw %.6e\nlaunch_ms %.6f\ncus %d\n", tp.bw, tp.launch_ms, tp.cus);
    fprintf(f, "# relative rate of the 128 x 128 and 64 x 64 tile kernels\ntile_eff %.6f %.6f\n", tp.teff[0],
            tp.teff[1]);
    fclose(f);
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_tuner_load(const char* path)
{
    if (!path) return FMMHIP_ERR_INVALID;
    FILE* f = fopen(path, "r");
    if (!f) return FMMHIP_ERR_IO;
    TunerParams tp;
    {
        std::lock_guard<std::mutex> lk(g_tuner_mu);
        tp = g_tp;
    }
    char line[256];
    int ok = 0;
    while (fgets(line, sizeof line, f)) {
        int c;
        char dt[8];
        double e, o;
        if (sscanf(line, "cand %d %7s %lf %lf", &c, dt, &e, &o) == 4 && c >= 0 && c < kCands && e > 0 && e <= 1 &&
            o >= 0) {
            const int d = strcmp(dt, "f64") == 0 ? 1 : 0;
            tp.eff[d][c] = e;
            tp.ovh[d][c] = o;
            ok++;
        } else if (sscanf(line, "bw %lf", &e) == 1 && e > 0) {
            tp.bw = e;
        } else if (sscanf(line, "launch_ms %lf", &e) == 1 && e >= 0) {
            tp.launch_ms = e;
        } else if (sscanf(line, "cus %d", &c) == 1 && c > 0) {
            tp.cus = c;
        } else if (sscanf(line, "tile_eff %lf %lf", &e, &o) == 2 && e > 0 && e <= 1.5 && o > 0 && o <= 1.5) {
            tp.teff[0] = e;
            tp.teff[1] = o;
        }
    }
    fclose(f);
    if (ok == 0) return FMMHIP_ERR_IO;
    std::lock_guard<std::mutex> lk(g_tuner_mu);
    g_tp = tp;
    return FMMHIP_SUCCESS;
}

// ---- batch ------------------------------------------------------------------

struct fmmhip_batch_s {
    struct Prob {
        int dbl;
        long m, n, k;
        double alpha, beta;
        const void* A;
        long lda;
        const void* B;
        long ldb;
        void* C;
        long ldc;
        int force;              // -1 tuner, else candidate
        int force_tile;         // -1 tuner, else tile index
        int cand;               // chosen
        int tile;               // chosen tile index
    };
    std::vector<Prob> probs;
    std::vector<Prepared> prep;
    // the batch's own algorithms: their descriptor caches never evict, so
    // every pointer captured in the graph stays valid for the batch's life
    fmmhip_alg_t algs[kTiles][kCands] = {};
    hipStream_t stream = nullptr;
    void* ws = nullptr;
    void* dev = nullptr;        // KArgs, items, segments, block offsets, heads
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool planned = false;
    double est_ms = 0.0;
    long items[2] = {0, 0};     // work items per datatype (both tiles)
    long fast_items[2] = {0, 0};  // of which on the fast kernel

    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        if (ws || dev) (void)hipDeviceSynchronize();
        if (ws) (void)hipFree(ws);
        if (dev) (void)hipFree(dev);
        ws = dev = nullptr;
        for (int t = 0; t < kTiles; t++)
            for (int c = 0; c < kCands; c++) {
                fmmhip_alg_destroy(algs[t][c]);
                algs[t][c] = nullptr;
            }
        planned = false;
    }

    int make_algs()
    {
        int rc = 0;
        for (int t = 0; t < kTiles && !rc; t++) {
            rc = fmmhip_alg_create_classical(&algs[t][0]);
            if (!rc) rc = fmmhip_alg_create(&algs[t][1], &STRASSEN_FMM, 1);
            fmm_t two[2] = {STRASSEN_FMM, STRASSEN_FMM};
            if (!rc) rc = fmmhip_alg_create(&algs[t][2], two, 2);
            for (int c = 0; c < kCands && !rc; c++) {
                algs[t][c]->keep_all = 1;
                // 128 x 128 tiles: the fast FP64 kernel where a problem allows it (shape 3; others 0)
                rc = fmmhip_alg_set_shape(algs[t][c], kTileShape[t] == 0 ? 3 : kTileShape[t]);
            }
        }
        return rc;
    }
};

extern "C" int fmmhip_batch_create(fmmhip_batch_t* out)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    if (!out) return FMMHIP_ERR_INVALID;
    fmmhip_batch_t b = new fmmhip_batch_s();
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return FMMHIP_ERR_HIP;
    }
    *out = b;
    return FMMHIP_SUCCESS;
}

extern "C" void fmmhip_batch_destroy(fmmhip_batch_t b)
{
    if (!b) return;
    b->release();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

static int batch_add(fmmhip_batch_t b, int dbl, long m, long n, long k, double alpha, const void* A, long lda,
                     const void* B, long ldb, double beta, void* C, long ldc, int force)
{
    if (!b || m < 0 || n < 0 || k < 0 || lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) ||
        ldc < (m > 1 ? m : 1) || force < -1 || force >= kCands)
        return FMMHIP_ERR_INVALID;
    b->release();
    fmmhip_batch_s::Prob p{dbl, m, n, k, alpha, beta, A, lda, B, ldb, C, ldc, force, -1, -1, 0};
    b->probs.push_back(p);
    return (int)b->probs.size() - 1 >= 0 ? FMMHIP_SUCCESS : FMMHIP_ERR_INVALID;
}

extern "C" int fmmhip_batch_add_d(fmmhip_batch_t b, long m, long n, long k, double alpha, const double* A, long lda,
                                  const double* B, long ldb, double beta, double* C, long ldc, int cand)
{
    return batch_add(b, 1, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cand);
}

extern "C" int fmmhip_batch_add_s(fmmhip_batch_t b, long m, long n, long k, float alpha, const float* A, long lda,
                                  const float* B, long ldb, float beta, float* C, long ldc, int cand)
{
    return batch_add(b, 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, cand);
}

extern "C" int fmmhip_batch_plan(fmmhip_batch_t b)
{
    if (!b) return FMMHIP_ERR_INVALID;
    b->release();
    int rc = b->make_algs();
    if (rc) return rc;
    const size_t np = b->probs.size();
    b->prep.assign(np, Prepared());
    b->est_ms = 0.0;
    // 1. choices and preparation (descriptors uploaded on the batch stream).
    //    FMM problems materialise their products (each product stored once
    //    into its slot by product groups, then the W stage), as single calls
    //    do where the slots fit; pre-summed with the W epilogue fused when the
    //    batch's workspace would pass the bound.  Both give the same C.
    size_t ws_total = 0;
    std::vector<size_t> ws_off(np, 0);
    // pass 1 with the products materialised; pass 2, pre-summed, only when
    // the first one's workspace passes the bound
    for (int pack_fmm = 2; pack_fmm >= 1; pack_fmm--) {
        ws_total = 0;
        b->est_ms = 0.0;
        for (size_t i = 0; i < np && !rc; i++) {
            fmmhip_batch_s::Prob& p = b->probs[i];
            double est = 0.0;
            TunerParams tp;
            {
                std::lock_guard<std::mutex> lk(g_tuner_mu);
                tp = g_tp;
            }
            const bool work = p.m > 0 && p.n > 0 && p.k > 0;
            p.cand = 0;
            p.tile = 0;
            if (work && p.force < 0 && p.force_tile < 0) {
                p.cand = choose(p.dbl, p.m, p.n, p.k, false, &est, &p.tile);
            } else if (work) {
                p.cand = p.force >= 0 ? p.force : choose(p.dbl, p.m, p.n, p.k, false, nullptr);
                if (p.force_tile >= 0) {
                    p.tile = p.force_tile;
                    est = est_ms(tp, p.dbl, p.cand, p.m, p.n, p.k, false, p.tile);
                } else {
                    p.tile = best_tile(tp, p.dbl, p.cand, p.m, p.n, p.k, false, &est);
                }
            }
            b->est_ms += est;
            b->prep[i] = Prepared();
            rc = prepare_call(b->algs[p.tile][p.cand], p.dbl ? Kind::D64 : Kind::S32, p.m, p.n, p.k, p.alpha, p.A, 1,
                              p.lda, p.B, 1, p.ldb, p.beta, p.C, 1, p.ldc, p.cand > 0 ? pack_fmm : 0,
                              p.cand > 0 ? -1 : 1, b->stream, &b->prep[i]);
            // the batched kernel is instantiated for column-major operands only
            if (!rc && !b->prep[i].nothing && !b->prep[i].scale_only &&
                !(b->prep[i].a_mn && !b->prep[i].b_mn && b->prep[i].c_col))
                rc = FMMHIP_ERR_UNSUPPORTED;
            if (!rc && b->prep[i].presum) {
                ws_off[i] = ws_total;
                ws_total += (b->prep[i].ws_bytes + 255) & ~(size_t)255;
            }
        }
        if (rc) return rc;
        if ((double)ws_total < fmmhip_ws_bound()) break;
    }
    if (ws_total) {
        if (hipMalloc(&b->ws, ws_total) != hipSuccess) {
            (void)hipGetLastError();
            b->ws = nullptr;
            return FMMHIP_ERR_NOMEM;
        }
        for (size_t i = 0; i < np; i++)
            if (b->prep[i].presum) bind_workspace(&b->prep[i], (char*)b->ws + ws_off[i]);
    }
    // 2. per datatype and launch shape (group g = 2 * s + datatype, s = 0
    //    for 128 x 128 tiles, 1 for 64 x 64, 2 for the fast kernel: whole
    //    problems, and the interior tiles of problems on 128 x 128 tiles):
    //    kernel arguments and work items (costliest first); per datatype:
    //    summation segments and their block offsets
    constexpr int kShapes = 3;
    const int kShapeOf[kShapes] = {0, 2, 3};
    constexpr int kGroups = 2 * kShapes;
    std::vector<KArgs> ka[kGroups];
    std::vector<BatchItem> items[kGroups];
    std::vector<BatchItem> qitems[kGroups][fmmk::kBatchQueues];
    std::vector<double> qcost[kGroups][fmmk::kBatchQueues];
    std::vector<fmmhip_presum_seg_t> segs[2];
    std::vector<int> boff[2];
    bool vec[kGroups];
    for (int g = 0; g < kGroups; g++) vec[g] = true;
    std::vector<size_t> scale_only;
    bool parts[2] = {true, true};     // part-staged summation for every segment of the datatype
    for (int d = 0; d < 2; d++) boff[d].push_back(0);
    for (size_t i = 0; i < np; i++) {
        const Prepared& P = b->prep[i];
        const int d = b->probs[i].dbl;
        if (P.nothing) continue;
        if (P.scale_only) { scale_only.push_back(i); continue; }
        const int g = 2 * (P.shape == 2 ? 1 : P.shape == 3 ? 2 : 0) + d;
        const int pi = (int)ka[g].size();
        ka[g].push_back(P.ka);
        vec[g] = vec[g] && P.ka.va && P.ka.vb;
        // a problem on the general 128 x 128 kernel sends the tiles that lie
        // inside all its parts to the fast kernel (same accumulation order,
        // so the same bits); the edge tiles stay on the general kernel
        const int gf = 2 * 2 + d;
        int pf = -1;
        if (g == d && P.single_part && fast_tiles_ok(P, d)) {
            pf = (int)ka[gf].size();
            ka[gf].push_back(P.ka);
        }
        const double ksteps = (double)((P.ka.lk + fmmk::BK - 1) / fmmk::BK);
        // per product group (materialised problems; else one group of all
        // products), the problem's tiles in blocks of 8 tile rows (column-
        // major inside a block), dealt to the queues in contiguous runs per
        // group, or, where the launch would deal whole groups to the XCDs
        // (xcd_groups), the (group, tile) sequence group-major in contiguous
        // runs; an item's tile index carries its group: tile + group * tiles
        const int tiles_m = P.ka.tiles_m, tiles_n = P.ka.tiles_n;
        const long ntiles = (long)tiles_m * tiles_n;
        const int G = P.ka.groups, R = P.ka.R;
        for (int gr = 0; gr < G; gr++) {
            const double c = (double)((long)(gr + 1) * R / G - (long)gr * R / G) * ksteps;
            for (long lin = 0; lin < ntiles; lin++) {
                const long per_group = 8L * tiles_n;
                const long gi = lin / per_group, first_m = gi * 8;
                const long gsz = std::min<long>(tiles_m - first_m, 8);
                const long in = lin - gi * per_group;
                const long tm = first_m + in % gsz, tn = in / gsz;
                const int q = P.ka.xcd_groups ? (int)(((long)gr * ntiles + lin) * fmmk::kBatchQueues / ((long)G * ntiles))
                                              : (int)(lin * fmmk::kBatchQueues / ntiles);
                const bool fast = pf >= 0 && fast_tile_inside(P.ka, tm, tn);
                qitems[fast ? gf : g][q].push_back(
                    BatchItem{fast ? pf : pi, (int)(tm + tn * tiles_m + (long)gr * ntiles)});
                qcost[fast ? gf : g][q].push_back(c);
            }
        }
        if (P.presum) {
            for (const fmmhip_presum_seg_t* sg : {&P.segA, &P.segB}) {
                segs[d].push_back(*sg);
                parts[d] = parts[d] && sg->parts && sg->np > 0 && sg->np <= FMMHIP_BATCH_PARTS_MAX;
            }
        }
    }
    // each queue costliest first, the queues of a group laid end to end
    b->items[0] = b->items[1] = 0;
    b->fast_items[0] = b->fast_items[1] = 0;
    std::vector<int> qoff[kGroups];
    for (int g = 0; g < kGroups; g++) {
        qoff[g].push_back(0);
        for (int q = 0; q < fmmk::kBatchQueues; q++) {
            std::vector<size_t> order(qitems[g][q].size());
            for (size_t j = 0; j < order.size(); j++) order[j] = j;
            std::stable_sort(order.begin(), order.end(),
                             [&](size_t x, size_t y) { return qcost[g][q][x] > qcost[g][q][y]; });
            for (size_t j = 0; j < order.size(); j++) items[g].push_back(qitems[g][q][order[j]]);
            qoff[g].push_back((int)items[g].size());
        }
        b->items[g % 2] += (long)items[g].size();
        if (g / 2 == 2) b->fast_items[g % 2] += (long)items[g].size();
    }
    // block ranges of the summation segments: part-staged units are 256
    // elements along c (32 per block), entry-driven ones pairs (4096 per block)
    for (int d = 0; d < 2; d++)
        for (const fmmhip_presum_seg_t& sg : segs[d]) {
            const long work = parts[d] ? (long)((sg.Lc + 255) / 256) * sg.Ls : (long)((sg.Lc + 1) / 2) * sg.Ls;
            const long per = parts[d] ? 32 : 4096;
            boff[d].push_back(boff[d].back() + (int)std::max(1L, (work + per - 1) / per));
        }
    // 3. one device block: [ka per group | items per group | segs0 | segs1 | boff0 | boff1 | heads]
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = (off + bytes + 255) & ~(size_t)255; return o; };
    size_t o_ka[kGroups], o_it[kGroups], o_qo[kGroups], o_sg[2], o_bo[2];
    for (int g = 0; g < kGroups; g++) o_ka[g] = take(sizeof(KArgs) * (ka[g].size() + 1));
    for (int g = 0; g < kGroups; g++) o_it[g] = take(sizeof(BatchItem) * (items[g].size() + 1));
    for (int g = 0; g < kGroups; g++) o_qo[g] = take(sizeof(int) * (fmmk::kBatchQueues + 1));
    for (int d = 0; d < 2; d++) o_sg[d] = take(sizeof(fmmhip_presum_seg_t) * (segs[d].size() + 1));
    for (int d = 0; d < 2; d++) o_bo[d] = take(sizeof(int) * (boff[d].size() + 1));
    const size_t o_head = take(kGroups * fmmk::kBatchQueues * sizeof(int));
    if (hipMalloc(&b->dev, off) != hipSuccess) { (void)hipGetLastError(); b->dev = nullptr; return FMMHIP_ERR_NOMEM; }
    char* dv = (char*)b->dev;
    for (int g = 0; g < kGroups; g++) {
        if (!ka[g].empty() &&
            hipMemcpy(dv + o_ka[g], ka[g].data(), sizeof(KArgs) * ka[g].size(), hipMemcpyHostToDevice) != hipSuccess)
            return FMMHIP_ERR_HIP;
        if (!items[g].empty() && hipMemcpy(dv + o_it[g], items[g].data(), sizeof(BatchItem) * items[g].size(),
                                           hipMemcpyHostToDevice) != hipSuccess)
            return FMMHIP_ERR_HIP;
        if (hipMemcpy(dv + o_qo[g], qoff[g].data(), sizeof(int) * qoff[g].size(), hipMemcpyHostToDevice) !=
            hipSuccess)
            return FMMHIP_ERR_HIP;
    }
    for (int d = 0; d < 2; d++) {
        if (!segs[d].empty() && hipMemcpy(dv + o_sg[d], segs[d].data(), sizeof(fmmhip_presum_seg_t) * segs[d].size(),
                                          hipMemcpyHostToDevice) != hipSuccess)
            return FMMHIP_ERR_HIP;
        if (hipMemcpy(dv + o_bo[d], boff[d].data(), sizeof(int) * boff[d].size(), hipMemcpyHostToDevice) != hipSuccess)
            return FMMHIP_ERR_HIP;
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) return FMMHIP_ERR_HIP;   // descriptor uploads
    // 4. capture: per datatype, the summation, then one persistent kernel per tile
    int cus = 256;
    fmmhip_device_cu_count(&cus);
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return FMMHIP_ERR_HIP;
    for (size_t idx : scale_only) {
        const Prepared& P = b->prep[idx];
        if (P.kind == Kind::S32) rc = fmmhip_scale_s((float*)P.C, P.m, P.n, P.rsC, P.csC, (float)P.beta, b->stream);
        else rc = fmmhip_scale_d((double*)P.C, P.m, P.n, P.rsC, P.csC, P.beta, b->stream);
        if (rc) break;
    }
    for (int d = 1; d >= 0 && !rc; d--) {
        if (!segs[d].empty())
            rc = fmmhip_presum_batch(d, (const fmmhip_presum_seg_t*)(dv + o_sg[d]), (int)segs[d].size(),
                                     (const int*)(dv + o_bo[d]), boff[d].back(), parts[d] ? 1 : 0, b->stream);
        for (int t = 0; t < kShapes && !rc; t++) {
            const int g = 2 * t + d;
            if (items[g].empty()) continue;
            const int shape = kShapeOf[t];
            const int nwg = (int)std::min<long>((long)items[g].size(), (shape == 2 ? 4L : 2L) * cus);
            int* heads = (int*)(dv + o_head) + g * fmmk::kBatchQueues;
            const int* qo = (const int*)(dv + o_qo[g]);
            if (d)
                rc = fmmk_launch_batch_d64((const KArgs*)(dv + o_ka[g]), (const BatchItem*)(dv + o_it[g]), qo, heads,
                                           vec[g], shape, nwg, b->stream);
            else
                rc = fmmk_launch_batch_s32((const KArgs*)(dv + o_ka[g]), (const BatchItem*)(dv + o_it[g]), qo, heads,
                                           vec[g], shape, nwg, b->stream);
        }
        // the W stages of the datatype's materialised problems
        for (size_t i = 0; i < np && !rc; i++) {
            const Prepared& P = b->prep[i];
            if (b->probs[i].dbl != d || P.nothing || P.scale_only || !P.wmat) continue;
            rc = fmmhip_wsum(d, &P.wsum, b->stream);
            if (!rc && P.w2) rc = fmmhip_wsum(d, &P.wsum2, b->stream);
        }
    }
    hipGraph_t g = nullptr;
    const bool cap_ok = hipStreamEndCapture(b->stream, &g) == hipSuccess;
    if (rc || !cap_ok) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : FMMHIP_ERR_HIP;
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) return FMMHIP_ERR_HIP;
    b->planned = true;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_batch_run(fmmhip_batch_t b, void* stream)
{
    if (!b || !b->planned) return FMMHIP_ERR_INVALID;
    hipStream_t st = stream ? (hipStream_t)stream : (hipStream_t)fmmhip_default_stream();
    return hipGraphLaunch(b->exec, st) == hipSuccess ? FMMHIP_SUCCESS : FMMHIP_ERR_HIP;
}

extern "C" int fmmhip_batch_info(fmmhip_batch_t b, long* nprob, long* items_f64, long* items_f32, double* est_ms,
                                 size_t* ws_bytes)
{
    if (!b) return FMMHIP_ERR_INVALID;
    if (nprob) *nprob = (long)b->probs.size();
    if (items_f64) *items_f64 = b->items[1];
    if (items_f32) *items_f32 = b->items[0];
    if (est_ms) *est_ms = b->est_ms;
    if (ws_bytes) {
        size_t t = 0;
        for (const Prepared& P : b->prep)
            if (P.presum) t += (P.ws_bytes + 255) & ~(size_t)255;
        *ws_bytes = t;
    }
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_batch_fast_items(fmmhip_batch_t b, long* fast_f64, long* fast_f32)
{
    if (!b || !b->planned) return FMMHIP_ERR_INVALID;
    if (fast_f64) *fast_f64 = b->fast_items[1];
    if (fast_f32) *fast_f32 = b->fast_items[0];
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_batch_choice(fmmhip_batch_t b, long i, int* cand)
{
    if (!b || i < 0 || (size_t)i >= b->probs.size() || !cand) return FMMHIP_ERR_INVALID;
    *cand = b->probs[i].cand;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_batch_set_tile(fmmhip_batch_t b, long i, int tile)
{
    if (!b || i < 0 || (size_t)i >= b->probs.size() || !(tile == 0 || tile == 64 || tile == 128))
        return FMMHIP_ERR_INVALID;
    b->release();
    b->probs[i].force_tile = tile == 0 ? -1 : tile == 64 ? 1 : 0;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_batch_tile(fmmhip_batch_t b, long i, int* tile)
{
    if (!b || i < 0 || (size_t)i >= b->probs.size() || !tile) return FMMHIP_ERR_INVALID;
    if (!b->planned) return FMMHIP_ERR_INVALID;
    const fmmhip_batch_s::Prob& p = b->probs[i];
    // a problem the 64 x 64 kernel cannot take runs on the 128 x 128 one
    *tile = fmmk::tile_size(b->prep[i].nothing || b->prep[i].scale_only ? kTileShape[p.tile] : b->prep[i].shape);
    return FMMHIP_SUCCESS;
}
