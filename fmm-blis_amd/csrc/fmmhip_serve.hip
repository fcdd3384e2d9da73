// fmmhip_serve.hip -- the serving form: a request loop with continuous
// batching over device-resident operands.
//
// A server owns one device arena (the HBM budget, allocated once and
// sub-allocated first-fit, so admission never calls hipMalloc/hipFree) and
// `slots` streams.  A host thread admits arrived requests in arrival order
// whenever a slot is free and the arena holds the request's output (when
// the server allocates it), coefficient lists and workspace; the request's
// summation and MFMA launches go to its slot's stream, so requests admitted
// while others run overlap on the device.  Completion is seen through one
// event per slot; the slot and the arena blocks are then freed and the
// latency (arrival to observed completion) recorded.  Each request's lists
// are packed on the host into the slot's pinned staging buffer and copied
// into its arena block on the slot's stream: no shared descriptor cache, no
// device-wide synchronisation on the request path.
//
// No counterpart in the reference (its driver serves one product per call,
// frame/base/bli_strassen_ab.c:123-295).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>
#include <string.h>

#include "fmmhip_host.h"

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, cap) with coalescing of free ranges.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;           // offset -> size

    void reset(char* b, size_t c)
    {
        base = b;
        cap = c;
        used = peak = 0;
        free_.clear();
        if (c) free_[0] = c;
    }
    bool alloc(size_t bytes, size_t* off)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used += bytes;
            peak = std::max(peak, used);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

constexpr int kCands = 3;
constexpr size_t kNoBlock = ~(size_t)0;

struct Req {
    fmmhip_request_t r;
    long id = 0;
    double t_arr = 0.0, t_admit = 0.0, t_done = 0.0;
    int cand = -1;
    int tile = 0;                   // 0: 128 x 128, 1: 64 x 64
    int status = -1;                // once done: 0, or the FMMHIP_ERR_* code
    bool done = false;
    size_t off1 = kNoBlock, bytes1 = 0;     // output + lists
    size_t off2 = kNoBlock, bytes2 = 0;     // workspace
    void* C = nullptr;
    long ldc = 0;
    bool own_c = false;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    void* pinned = nullptr;
    size_t pinned_cap = 0;
    long req = -1;
};

// Lists packed into the slot's pinned buffer, bound to the request's block.
struct PackSource : DescSource {
    void* host;
    const void* dev;
    size_t cap, used = 0;
    int get(fmmhip_alg_t alg, const fmm_geom_t& g, bool presum, bool mat, bool a_mn, bool b_mn, hipStream_t,
            DescView* v) override
    {
        return desc_pack(alg, g, presum, mat, a_mn, b_mn, host, dev, cap, &used, v);
    }
    bool can_mat() const override { return true; }
};

}  // namespace

struct fmmhip_server_s {
    std::mutex mu;
    std::condition_variable cv_done;
    std::thread th;
    bool stop = false;
    double t0 = 0.0;
    void* dev = nullptr;
    Arena arena;
    std::vector<Slot> slots;
    fmmhip_alg_t algs[2][kCands] = {};          // [tile: 128 x 128, 64 x 64][candidate]
    size_t list_cap = 0;
    long next_id = 1;
    std::deque<long> waiting;                    // submitted, in arrival order
    std::unordered_map<long, Req> reqs;
    std::vector<double> lat_ms;
    long submitted = 0, completed = 0, failed = 0;
    int inflight = 0, max_inflight = 0;
    double first_arr = -1.0, last_done = 0.0;

    void loop();
    bool try_admit(Req& q, Slot& sl, int si);
    void finish(Req& q, int rc);
};

// Frees a request's blocks and records it (called with mu held).
void fmmhip_server_s::finish(Req& q, int rc)
{
    if (q.off2 != kNoBlock) arena.release(q.off2, q.bytes2);
    if (q.off1 != kNoBlock && !(q.own_c && q.r.keep && rc == 0)) {
        arena.release(q.off1, q.bytes1);
        q.off1 = kNoBlock;
        if (q.own_c) q.C = nullptr;
    }
    q.off2 = kNoBlock;
    q.t_done = now_s() - t0;
    q.done = true;
    q.status = rc;
    if (rc == 0) {
        completed++;
        lat_ms.push_back((q.t_done - q.r.arrival_s) * 1e3);
        last_done = std::max(last_done, q.t_done);
    } else {
        failed++;
    }
    cv_done.notify_all();
}

// Resolves, places and launches one request on slot si; false: it does not
// fit in the arena now (nothing was launched or kept).
bool fmmhip_server_s::try_admit(Req& q, Slot& sl, int si)
{
    const fmmhip_request_t& r = q.r;
    if (q.cand < 0) {
        int edge = 128;
        if (r.cand >= 0) q.cand = r.cand;
        else if (r.m > 0 && r.n > 0 && r.k > 0)
            fmmhip_tuner_choose_tile(r.is_double, r.m, r.n, r.k, 1, &q.cand, &edge, nullptr);
        else q.cand = 0;
        q.tile = edge == 64 ? 1 : 0;
    }
    fmmhip_alg_t alg = algs[q.tile][q.cand];
    const size_t esz = r.is_double ? 8 : 4;
    const bool own_c = r.C == nullptr;
    const size_t c_bytes = own_c ? (esz * (size_t)std::max(r.m, 1L) * (size_t)std::max(r.n, 1L) + 255) & ~(size_t)255 : 0;
    const size_t b1 = c_bytes + list_cap;
    if (b1 > arena.cap) { finish(q, FMMHIP_ERR_NOMEM); return true; }
    size_t off1;
    if (!arena.alloc(b1, &off1)) return false;
    char* blk = arena.base + off1;
    void* C = own_c ? (void*)blk : r.C;
    const long ldc = own_c ? std::max(r.m, 1L) : r.ldc;
    PackSource src;
    src.host = sl.pinned;
    src.dev = blk + c_bytes;
    src.cap = list_cap;
    Prepared P;
    int rc = prepare_call(alg, r.is_double ? Kind::D64 : Kind::S32, r.m, r.n, r.k, r.alpha, r.A, 1, r.lda, r.B, 1,
                          r.ldb, r.beta, C, 1, ldc, q.cand > 0 ? 2 : 0, alg->split, sl.st, &P, &src);
    if (rc) {
        arena.release(off1, b1);
        finish(q, rc);
        return true;
    }
    size_t off2 = kNoBlock;
    if (P.ws_bytes) {
        if (((b1 + 255) & ~(size_t)255) + ((P.ws_bytes + 255) & ~(size_t)255) > arena.cap) {
            arena.release(off1, b1);
            finish(q, FMMHIP_ERR_NOMEM);
            return true;
        }
        if (!arena.alloc(P.ws_bytes, &off2)) {
            arena.release(off1, b1);
            return false;
        }
        bind_workspace(&P, arena.base + off2);
    }
    q.off1 = off1;
    q.bytes1 = b1;
    q.off2 = off2;
    q.bytes2 = P.ws_bytes;
    q.C = C;
    q.ldc = ldc;
    q.own_c = own_c;
    if (src.used && hipMemcpyAsync(blk + c_bytes, sl.pinned, src.used, hipMemcpyHostToDevice, sl.st) != hipSuccess)
        rc = FMMHIP_ERR_HIP;
    if (!rc) rc = launch_prepared(&P);
    if (!rc && hipEventRecord(sl.ev, sl.st) != hipSuccess) rc = FMMHIP_ERR_HIP;
    if (rc) {
        (void)hipStreamSynchronize(sl.st);
        finish(q, rc);
        return true;
    }
    q.t_admit = now_s() - t0;
    sl.req = q.id;
    inflight++;
    max_inflight = std::max(max_inflight, inflight);
    (void)si;
    return true;
}

void fmmhip_server_s::loop()
{
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progress = false;
        // completions
        for (Slot& sl : slots) {
            if (sl.req < 0) continue;
            const hipError_t e = hipEventQuery(sl.ev);
            if (e == hipErrorNotReady) continue;
            Req& q = reqs[sl.req];
            sl.req = -1;
            inflight--;
            finish(q, e == hipSuccess ? 0 : FMMHIP_ERR_HIP);
            progress = true;
        }
        // admissions, in arrival order, while slots are free
        const double t = now_s() - t0;
        while (!waiting.empty()) {
            Req& q = reqs[waiting.front()];
            if (q.r.arrival_s > t) break;
            int si = -1;
            for (int i = 0; i < (int)slots.size(); i++)
                if (slots[i].req < 0) { si = i; break; }
            if (si < 0) break;
            if (!try_admit(q, slots[si], si)) {
                // waits for memory held by requests in flight; with none in
                // flight only kept results hold it, and the request fails
                if (inflight > 0) break;
                finish(q, FMMHIP_ERR_NOMEM);
            }
            waiting.pop_front();
            progress = true;
        }
        if (stop && waiting.empty() && inflight == 0) break;
        if (!progress) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        }
    }
}

extern "C" int fmmhip_server_create(fmmhip_server_t* out, size_t hbm_budget, int slots)
{
    int rc = fmmhip_init();
    if (rc) return rc;
    if (!out || slots < 1 || slots > 64 || hbm_budget < (1u << 20)) return FMMHIP_ERR_INVALID;
    fmmhip_server_t s = new fmmhip_server_s();
    fmm_t two[2] = {STRASSEN_FMM, STRASSEN_FMM};
    for (int t = 0; t < 2 && !rc; t++) {
        rc = fmmhip_alg_create_classical(&s->algs[t][0]);
        if (!rc) rc = fmmhip_alg_create(&s->algs[t][1], &STRASSEN_FMM, 1);
        if (!rc) rc = fmmhip_alg_create(&s->algs[t][2], two, 2);
        // shape 2: 64 x 64 tiles (taken where the 128 x 128 grid pads much)
        for (int c = 0; c < kCands && !rc; c++) rc = fmmhip_alg_set_shape(s->algs[t][c], t ? 2 : -1);
    }
    if (rc) { fmmhip_server_destroy(s); return rc; }
    // no product split: concurrent requests fill the device, and a split
    // would add up to 8 C-sized temporaries to every request's arena share;
    // the classical candidate keeps its k-split, taken only by requests too
    // small to occupy half the CUs (small temporaries; the choice depends
    // on the request alone, so a result is the same alone and in the mix)
    for (int t = 0; t < 2; t++) {
        for (int c = 1; c < kCands; c++) s->algs[t][c]->split = 1;
        for (int c = 0; c < kCands; c++)
            s->list_cap = std::max(s->list_cap, (desc_bound(s->algs[t][c]) + 255) & ~(size_t)255);
    }
    if (hipMalloc(&s->dev, hbm_budget) != hipSuccess) {
        (void)hipGetLastError();
        s->dev = nullptr;
        fmmhip_server_destroy(s);
        return FMMHIP_ERR_NOMEM;
    }
    s->arena.reset((char*)s->dev, hbm_budget);
    s->slots.resize(slots);
    for (Slot& sl : s->slots) {
        sl.pinned_cap = s->list_cap;
        if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc(&sl.pinned, sl.pinned_cap, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            fmmhip_server_destroy(s);
            return FMMHIP_ERR_HIP;
        }
    }
    s->t0 = now_s();
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return FMMHIP_SUCCESS;
}

extern "C" double fmmhip_server_now(fmmhip_server_t s)
{
    return s ? now_s() - s->t0 : 0.0;
}

extern "C" int fmmhip_server_submit(fmmhip_server_t s, const fmmhip_request_t* req, long* id)
{
    if (!s || !req) return FMMHIP_ERR_INVALID;
    const fmmhip_request_t& r = *req;
    if (r.m < 0 || r.n < 0 || r.k < 0 || r.cand < -1 || r.cand >= kCands) return FMMHIP_ERR_INVALID;
    if (r.lda < std::max(r.m, 1L) || r.ldb < std::max(r.k, 1L)) return FMMHIP_ERR_INVALID;
    if (r.C && r.ldc < std::max(r.m, 1L)) return FMMHIP_ERR_INVALID;
    if (!r.C && r.beta != 0.0) return FMMHIP_ERR_INVALID;       // a server-allocated C has no prior value
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return FMMHIP_ERR_INVALID;
    Req q;
    q.r = r;
    const double t = now_s() - s->t0;
    if (q.r.arrival_s < t) q.r.arrival_s = t;
    q.id = s->next_id++;
    // arrival order: insert behind every request that arrives no later
    auto it = s->waiting.end();
    while (it != s->waiting.begin() && s->reqs[*(it - 1)].r.arrival_s > q.r.arrival_s) --it;
    s->reqs.emplace(q.id, q);
    s->waiting.insert(it, q.id);
    s->submitted++;
    if (s->first_arr < 0 || q.r.arrival_s < s->first_arr) s->first_arr = q.r.arrival_s;
    if (id) *id = q.id;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_server_wait(fmmhip_server_t s, long id, int* status)
{
    if (!s) return FMMHIP_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return FMMHIP_ERR_INVALID;
    s->cv_done.wait(lk, [&] { return it->second.done; });
    if (status) *status = it->second.status;
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_server_result(fmmhip_server_t s, long id, void** C, long* ldc, double* latency_ms, int* cand)
{
    if (!s) return FMMHIP_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || !it->second.done) return FMMHIP_ERR_INVALID;
    const Req& q = it->second;
    if (C) *C = q.C;
    if (ldc) *ldc = q.ldc;
    if (latency_ms) *latency_ms = (q.t_done - q.r.arrival_s) * 1e3;
    if (cand) *cand = q.cand;
    return q.status;
}

extern "C" int fmmhip_server_release(fmmhip_server_t s, long id)
{
    if (!s) return FMMHIP_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || !it->second.done) return FMMHIP_ERR_INVALID;
    Req& q = it->second;
    if (q.off1 != kNoBlock) s->arena.release(q.off1, q.bytes1);
    s->reqs.erase(it);
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_server_drain(fmmhip_server_t s)
{
    if (!s) return FMMHIP_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->waiting.empty() && s->inflight == 0; });
    return FMMHIP_SUCCESS;
}

extern "C" int fmmhip_server_stats(fmmhip_server_t s, fmmhip_server_stats_t* st)
{
    if (!s || !st) return FMMHIP_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    std::vector<double> v = s->lat_ms;
    std::sort(v.begin(), v.end());
    if (!v.empty()) {
        auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)(p / 100.0 * (v.size() - 1) + 0.5))]; };
        st->p50_ms = pct(50);
        st->p99_ms = pct(99);
        st->max_ms = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        st->mean_ms = sum / v.size();
    }
    st->span_s = s->first_arr >= 0 ? s->last_done - s->first_arr : 0.0;
    st->arena_bytes = s->arena.cap;
    st->arena_peak = s->arena.peak;
    st->max_inflight = s->max_inflight;
    return FMMHIP_SUCCESS;
}

extern "C" void fmmhip_server_destroy(fmmhip_server_t s)
{
    if (!s) return;
    if (s->th.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->th.join();
    }
    for (Slot& sl : s->slots) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.st) (void)hipStreamDestroy(sl.st);
        if (sl.pinned) (void)hipHostFree(sl.pinned);
    }
    if (s->dev) (void)hipFree(s->dev);
    for (int t = 0; t < 2; t++)
        for (int c = 0; c < kCands; c++) fmmhip_alg_destroy(s->algs[t][c]);
    delete s;
}
