// host.hip -- the join from pageable host memory to sorted host results
// (tj_host_*).
//
// The reference's public runs time this whole path (pageable FP32 source ->
// device -> stages -> sorted canonical ids on the host,
// CEJ/src/run_g5_tensorjoin_public.py:372-547), and most of their wall time
// was spent outside the GPU stages (CEJ/DESIGN_G5_UNIFIED_ARTIFACT.md:33-36).
// Here the host side is a pipeline:
//
//   pageable input --(parallel memcpy)--> pinned staging buffer
//                  --(DMA, copy stream)--> device rows
//   device join (tj_join on the compute stream), chunk by chunk
//   device pairs --(DMA, copy stream)--> pinned staging buffer
//                --(parallel memcpy, + row offset)--> host result
//
// An R x S join runs over contiguous chunks of R rows.  Every chunk is an
// independent R_chunk x S join whose sorted pairs, shifted by the chunk's
// first row, continue the global (i, j) order, so the host result is the
// concatenation.  While the device joins chunk c, the host stages chunk c+1's
// rows and copies chunk c-1's pairs out: transfers and host copies hide
// behind the device work.  A self-join (or the grid method) is one chunk with
// the input upload and the result download pipelined through the staging
// buffers.
//
// Capacities start from an estimate and grow when a chunk reports more pairs
// than its buffers hold; that chunk (and the one already queued behind it)
// then runs again (the retry protocol of CEJ/src/run_g2a_tensorjoin.py:329-378).
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <sys/mman.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <new>
#include <thread>
#include <vector>
#include "tj_internal.h"

using namespace tj;

namespace {

// ------------------------------------------------------------ copy workers
// A fixed pool of host threads for the pageable <-> pinned copies (one
// thread cannot saturate either side of a staging buffer).  Only the calling
// thread makes HIP calls.
class CopyPool {
  public:
    explicit CopyPool(int n) {
        for (int t = 1; t < n; ++t) workers_.emplace_back([this, t] { loop(t); });
        nthreads_ = n;
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto& w : workers_) w.join();
    }
    int threads() const { return nthreads_; }
    // Run f(t, nthreads) on every thread (the caller is t = 0) and wait.
    void run(const std::function<void(int, int)>& f) {
        if (nthreads_ == 1) { f(0, 1); return; }
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &f;
            pending_ = nthreads_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0, nthreads_);
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [this] { return pending_ == 0; });
        job_ = nullptr;
    }

  private:
    void loop(int t) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int, int)>* f;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                f = job_;
            }
            (*f)(t, nthreads_);
            {
                std::lock_guard<std::mutex> g(m_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int, int)>* job_ = nullptr;
    uint64_t gen_ = 0;
    int pending_ = 0;
    int nthreads_ = 1;
    bool stop_ = false;
};

// [lo, hi) of item range n for worker t of nt, split at multiples of `align`
inline void split(int64_t n, int t, int nt, int64_t align, int64_t* lo, int64_t* hi) {
    const int64_t units = (n + align - 1) / align;
    const int64_t a = units * t / nt, b = units * (t + 1) / nt;
    *lo = std::min(n, a * align);
    *hi = std::min(n, b * align);
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int NPIN = 4;                    // staging buffers
constexpr size_t PIN_BYTES = 32ull << 20;  // bytes per staging buffer

}  // namespace

struct tj_host {
    int64_t nR, nS;
    bool self_join;
    int d, method, prec;
    unsigned flags;
    int64_t chunk_rows, nchunks;
    // plans for the full chunk and the last (shorter) chunk
    tj_plan* plan[2] = {nullptr, nullptr};
    int64_t plan_rows[2] = {0, 0};
    int64_t out_cap = 0, amb_cap = 0;
    // device memory
    float* dR = nullptr;
    float* dS = nullptr;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    void* dpairs[2] = {nullptr, nullptr};
    size_t dpairs_bytes = 0;
    uint64_t* dstats = nullptr;  // 2 x TJ_STATS_SLOTS
    void* canon_temp = nullptr;
    size_t canon_temp_bytes = 0;
    uint64_t* dids = nullptr;
    size_t dids_bytes = 0;
    // host side
    uint8_t* pin[NPIN] = {};
    hipEvent_t pin_ev[NPIN] = {};
    bool pin_used[NPIN] = {};
    int pin_next = 0;
    uint64_t* pstats = nullptr;  // pinned, 2 x TJ_STATS_SLOTS
    hipStream_t s_comp = nullptr, s_copy = nullptr;
    hipEvent_t ev_up[2] = {}, ev_done[2] = {};
    CopyPool* pool = nullptr;
    double times[4] = {0, 0, 0, 0};
};

namespace {

int ensure_dev(void** p, size_t* have, size_t need) {
    if (*have >= need && *p) return TJ_OK;
    if (*p) TJ_HIP(hipFree(*p));
    *p = nullptr;
    *have = 0;
    TJ_HIP(hipMalloc(p, need > 0 ? need : 256));
    *have = need;
    return TJ_OK;
}

// (Re)create the chunk plans and their shared workspace for the current
// capacities.  Synchronises the device when buffers are replaced.
int make_plans(tj_host* h) {
    TJ_HIP(hipDeviceSynchronize());
    for (int k = 0; k < 2; ++k) {
        if (h->plan[k]) tj_plan_destroy(h->plan[k]);
        h->plan[k] = nullptr;
    }
    size_t ws = 0;
    for (int k = 0; k < 2; ++k) {
        const int64_t rows = h->plan_rows[k];
        if (rows == 0 && k == 1) continue;
        int rc = tj_plan_create(&h->plan[k], rows, h->self_join ? 0 : h->nS, h->d, h->method,
                                h->method == TJ_METHOD_CERTIFIED ? h->prec : TJ_PREC_FP16, h->out_cap, h->amb_cap,
                                h->flags & TJ_FLAG_AUDIT);
        if (rc) return rc;
        size_t b = 0;
        tj_plan_workspace_bytes(h->plan[k], &b);
        ws = std::max(ws, b);
    }
    int rc = ensure_dev(&h->ws, &h->ws_bytes, ws);
    if (rc) return rc;
    const size_t pb = (size_t)std::max<int64_t>(h->out_cap, 1) * 2 * 8;  // room for int64 pairs
    if (h->dpairs_bytes < pb) {
        for (int k = 0; k < 2; ++k) {
            if (h->dpairs[k]) TJ_HIP(hipFree(h->dpairs[k]));
            h->dpairs[k] = nullptr;
        }
        for (int k = 0; k < 2; ++k) TJ_HIP(hipMalloc(&h->dpairs[k], pb));
        h->dpairs_bytes = pb;
    }
    return TJ_OK;
}

// Pageable src -> device dst through the staging buffers (copy stream).
// Returns once every piece is queued; the last DMA may still run.
int upload(tj_host* h, void* dst, const void* src, size_t bytes) {
    size_t off = 0;
    while (off < bytes) {
        const size_t n = std::min(PIN_BYTES, bytes - off);
        const int b = h->pin_next;
        h->pin_next = (h->pin_next + 1) % NPIN;
        if (h->pin_used[b]) TJ_HIP(hipEventSynchronize(h->pin_ev[b]));
        uint8_t* pin = h->pin[b];
        const uint8_t* s = (const uint8_t*)src + off;
        h->pool->run([&](int t, int nt) {
            int64_t lo, hi;
            split((int64_t)n, t, nt, 4096, &lo, &hi);
            if (hi > lo) memcpy(pin + lo, s + lo, (size_t)(hi - lo));
        });
        TJ_HIP(hipMemcpyAsync((uint8_t*)dst + off, pin, n, hipMemcpyHostToDevice, h->s_copy));
        TJ_HIP(hipEventRecord(h->pin_ev[b], h->s_copy));
        h->pin_used[b] = true;
        off += n;
    }
    return TJ_OK;
}

// Device src -> host dst through the staging buffers, after event `after`.
// Pairs (elem = 4 or 8 bytes, two per row) get `row0` added to their first
// member; other data is copied as is (row0 = 0).  Synchronous.
int download(tj_host* h, void* dst, const void* src, size_t bytes, hipEvent_t after, int elem, int64_t row0) {
    if (bytes == 0) return TJ_OK;
    TJ_HIP(hipStreamWaitEvent(h->s_copy, after, 0));
    const size_t piece = PIN_BYTES / 16 * 16;  // whole pairs of either width
    const size_t np = (bytes + piece - 1) / piece;
    auto issue = [&](size_t k) -> int {
        const int b = (int)(k % NPIN);
        if (h->pin_used[b]) TJ_HIP(hipEventSynchronize(h->pin_ev[b]));
        const size_t off = k * piece, n = std::min(piece, bytes - off);
        TJ_HIP(hipMemcpyAsync(h->pin[b], (const uint8_t*)src + off, n, hipMemcpyDeviceToHost, h->s_copy));
        TJ_HIP(hipEventRecord(h->pin_ev[b], h->s_copy));
        h->pin_used[b] = true;
        return TJ_OK;
    };
    int rc;
    for (size_t k = 0; k < np && k < (size_t)NPIN; ++k)
        if ((rc = issue(k))) return rc;
    for (size_t k = 0; k < np; ++k) {
        const int b = (int)(k % NPIN);
        TJ_HIP(hipEventSynchronize(h->pin_ev[b]));
        const size_t off = k * piece, n = std::min(piece, bytes - off);
        const uint8_t* pin = h->pin[b];
        uint8_t* out = (uint8_t*)dst + off;
        h->pool->run([&](int t, int nt) {
            int64_t lo, hi;
            split((int64_t)n, t, nt, 4096, &lo, &hi);
            if (hi <= lo) return;
            if (row0 == 0) {
                memcpy(out + lo, pin + lo, (size_t)(hi - lo));
            } else if (elem == 4) {
                const int32_t* a = (const int32_t*)(pin + lo);
                int32_t* o = (int32_t*)(out + lo);
                const int64_t m = (hi - lo) / 8;
                const int32_t r0 = (int32_t)row0;
                for (int64_t q = 0; q < m; ++q) {
                    o[2 * q] = a[2 * q] + r0;
                    o[2 * q + 1] = a[2 * q + 1];
                }
            } else {
                const int64_t* a = (const int64_t*)(pin + lo);
                int64_t* o = (int64_t*)(out + lo);
                const int64_t m = (hi - lo) / 16;
                for (int64_t q = 0; q < m; ++q) {
                    o[2 * q] = a[2 * q] + row0;
                    o[2 * q + 1] = a[2 * q + 1];
                }
            }
        });
        if (k + NPIN < np && (rc = issue(k + NPIN))) return rc;
    }
    return TJ_OK;
}

// The host result.  Large buffers are 2 MiB aligned and marked for huge
// pages, so first-touch faults of a fresh result cost one fault per 2 MiB
// rather than per 4 KiB (where the kernel allows transparent huge pages).
struct Out {
    uint8_t* buf = nullptr;
    size_t bytes = 0, cap = 0;
    ~Out() { free(buf); }  // released to the caller on success
    Out() = default;
    Out(const Out&) = delete;
    Out& operator=(const Out&) = delete;
    int reserve(size_t need) {
        if (need <= cap) return TJ_OK;
        const size_t c = std::max(need, cap + cap / 2);
        void* nb = nullptr;
        if (c >= (8u << 20)) {
            const size_t al = (c + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
            if (posix_memalign(&nb, 2u << 20, al) != 0) nb = nullptr;
            if (nb) (void)madvise(nb, al, MADV_HUGEPAGE);
        } else {
            nb = malloc(c);
        }
        TJ_CHECK(nb != nullptr, TJ_ERR_INVALID, "out of host memory for %zu result bytes", c);
        if (bytes) memcpy(nb, buf, bytes);
        free(buf);
        buf = (uint8_t*)nb;
        cap = c;
        return TJ_OK;
    }
};

}  // namespace

extern "C" {

int tj_host_create(tj_host** out, int64_t nR, int64_t nS, int d, int method, int precision, unsigned flags,
                   int64_t chunk_rows, int threads) {
    TJ_CHECK(out != nullptr, TJ_ERR_INVALID, "host pointer is NULL");
    *out = nullptr;
    TJ_CHECK(nR >= 0 && nS >= 0 && d >= 1, TJ_ERR_INVALID, "bad shape nR=%lld nS=%lld d=%d", (long long)nR,
             (long long)nS, d);
    TJ_CHECK(nR < (1ll << 31) && nS < (1ll << 31), TJ_ERR_UNSUPPORTED, "point sets beyond 2^31 rows");
    TJ_CHECK(chunk_rows >= 0 && threads >= 0, TJ_ERR_INVALID, "negative chunk_rows or threads");
    // validate method / precision / shape with a throw-away plan
    tj_plan* probe = nullptr;
    int rc = tj_plan_create(&probe, std::min<int64_t>(nR, 64), nS == 0 ? 0 : std::min<int64_t>(nS, 64), d, method,
                            method == TJ_METHOD_CERTIFIED ? precision : TJ_PREC_FP16, 1, 1, 0);
    if (rc) return rc;
    tj_plan_destroy(probe);

    tj_host* h = new (std::nothrow) tj_host();
    TJ_CHECK(h != nullptr, TJ_ERR_INVALID, "out of host memory");
    h->nR = nR;
    h->self_join = nS == 0;
    h->nS = h->self_join ? nR : nS;
    h->d = d;
    h->method = method;
    h->prec = precision;
    h->flags = flags;
    // R x S: chunks of about 1/8 of R (at least 16384 rows), so that staging
    // and copying out hide behind the device work; one chunk otherwise
    if (h->self_join || method == TJ_METHOD_GRID || nR == 0) {
        h->chunk_rows = std::max<int64_t>(nR, 1);
    } else {
        int64_t c = chunk_rows > 0 ? chunk_rows : std::max<int64_t>(16384, (nR + 7) / 8);
        h->chunk_rows = std::min<int64_t>(c, nR);
    }
    h->nchunks = nR == 0 ? 1 : (nR + h->chunk_rows - 1) / h->chunk_rows;
    h->plan_rows[0] = std::min<int64_t>(h->chunk_rows, nR);
    h->plan_rows[1] = (h->nchunks > 1 && nR % h->chunk_rows) ? nR % h->chunk_rows : 0;
    // initial capacities from the pair count of a chunk; grown on overflow
    const int64_t cp = h->self_join ? nR * (nR - 1) / 2 : h->plan_rows[0] * h->nS;
    h->out_cap = std::min<int64_t>(std::max<int64_t>(cp, 1),
                                   std::max<int64_t>({1ll << 20, std::min<int64_t>(cp / 512, 1ll << 26),
                                                      16 * std::max(h->plan_rows[0], h->nS)}));
    h->amb_cap = std::min<int64_t>(std::max<int64_t>(cp, 1), 1ll << 22);
    const int hw = (int)std::thread::hardware_concurrency();
    const int nt = threads > 0 ? threads : std::max(1, std::min(8, hw > 0 ? hw : 1));
    h->pool = new (std::nothrow) CopyPool(nt);
    auto fail = [&](int code) {
        tj_host_destroy(h);
        return code;
    };
    if (!h->pool) return fail(TJ_ERR_INVALID);
    if (hipStreamCreateWithFlags(&h->s_comp, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&h->s_copy, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(TJ_ERR_HIP);
    }
    for (int k = 0; k < NPIN; ++k) {
        if (hipHostMalloc((void**)&h->pin[k], PIN_BYTES, 0) != hipSuccess ||
            hipEventCreateWithFlags(&h->pin_ev[k], hipEventDisableTiming) != hipSuccess) {
            set_error("pinned staging allocation failed");
            return fail(TJ_ERR_HIP);
        }
    }
    for (int k = 0; k < 2; ++k)
        if (hipEventCreateWithFlags(&h->ev_up[k], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_done[k], hipEventDisableTiming) != hipSuccess) {
            set_error("hipEventCreate failed");
            return fail(TJ_ERR_HIP);
        }
    if (hipHostMalloc((void**)&h->pstats, 2 * TJ_STATS_SLOTS * 8, 0) != hipSuccess ||
        hipMalloc((void**)&h->dstats, 2 * TJ_STATS_SLOTS * 8) != hipSuccess) {
        set_error("stats allocation failed");
        return fail(TJ_ERR_HIP);
    }
    const size_t in_bytes = (size_t)nR * d * 4;
    if (hipMalloc((void**)&h->dR, std::max<size_t>(in_bytes, 256)) != hipSuccess ||
        (!h->self_join && hipMalloc((void**)&h->dS, std::max<size_t>((size_t)nS * d * 4, 256)) != hipSuccess)) {
        set_error("device input allocation failed");
        return fail(TJ_ERR_HIP);
    }
    if ((rc = make_plans(h))) return fail(rc);
    *out = h;
    return TJ_OK;
}

int tj_host_destroy(tj_host* h) {
    if (!h) return TJ_OK;
    if (h->s_comp) (void)hipStreamSynchronize(h->s_comp);
    if (h->s_copy) (void)hipStreamSynchronize(h->s_copy);
    for (int k = 0; k < 2; ++k)
        if (h->plan[k]) tj_plan_destroy(h->plan[k]);
    void* dev[] = {h->dR, h->dS, h->ws, h->dpairs[0], h->dpairs[1], h->dstats, h->canon_temp, h->dids};
    for (void* p : dev)
        if (p) (void)hipFree(p);
    for (int k = 0; k < NPIN; ++k) {
        if (h->pin[k]) (void)hipHostFree(h->pin[k]);
        if (h->pin_ev[k]) (void)hipEventDestroy(h->pin_ev[k]);
    }
    for (int k = 0; k < 2; ++k) {
        if (h->ev_up[k]) (void)hipEventDestroy(h->ev_up[k]);
        if (h->ev_done[k]) (void)hipEventDestroy(h->ev_done[k]);
    }
    if (h->pstats) (void)hipHostFree(h->pstats);
    if (h->s_comp) (void)hipStreamDestroy(h->s_comp);
    if (h->s_copy) (void)hipStreamDestroy(h->s_copy);
    delete h->pool;
    delete h;
    return TJ_OK;
}

void tj_host_free(void* result) { free(result); }

int tj_host_times(const tj_host* h, double* ms, int n) {
    TJ_CHECK(h && ms, TJ_ERR_INVALID, "NULL argument");
    for (int k = 0; k < n && k < 4; ++k) ms[k] = h->times[k];
    return TJ_OK;
}

int tj_host_join(tj_host* h, const float* R, const float* S, double T, int out_format, void** result,
                 int64_t* count, uint64_t* stats) {
    TJ_CHECK(h && result && count, TJ_ERR_INVALID, "NULL argument");
    TJ_CHECK(R != nullptr || h->nR == 0, TJ_ERR_INVALID, "R is NULL");
    TJ_CHECK(h->self_join || S != nullptr || h->nS == 0, TJ_ERR_INVALID, "S is NULL for an R x S join");
    TJ_CHECK(out_format == TJ_OUT_I32 || out_format == TJ_OUT_I64 || out_format == TJ_OUT_U64_CANONICAL,
             TJ_ERR_INVALID, "bad out_format %d", out_format);
    TJ_CHECK(out_format != TJ_OUT_U64_CANONICAL || h->self_join, TJ_ERR_INVALID,
             "canonical directed ids are defined for self-joins");
    TJ_CHECK(T == T && T >= 0.0 && T < INFINITY, TJ_ERR_INVALID, "T must be finite and >= 0");
    *result = nullptr;
    *count = 0;
    const double t0 = now_ms();
    double t_stage = 0, t_wait = 0, t_out = 0;
    const int pair_fmt = out_format == TJ_OUT_I32 ? TJ_OUT_I32 : TJ_OUT_I64;
    const int elem = pair_fmt == TJ_OUT_I32 ? 4 : 8;
    const size_t row_bytes = (size_t)h->d * 4;
    uint64_t acc[TJ_STATS_SLOTS] = {};
    Out out;
    int rc;

    // inputs: S whole (R x S), then R chunk by chunk
    double ts = now_ms();
    if (!h->self_join && (rc = upload(h, h->dS, S, (size_t)h->nS * row_bytes))) return rc;
    auto rows_of = [&](int64_t c) { return std::min(h->chunk_rows, h->nR - c * h->chunk_rows); };
    auto stage = [&](int64_t c) -> int {
        const int64_t r0 = c * h->chunk_rows;
        int e = upload(h, (uint8_t*)h->dR + r0 * row_bytes, (const uint8_t*)R + r0 * row_bytes,
                       (size_t)rows_of(c) * row_bytes);
        if (e) return e;
        TJ_HIP(hipEventRecord(h->ev_up[c & 1], h->s_copy));
        return TJ_OK;
    };
    auto compute = [&](int64_t c) -> int {
        const int64_t r0 = c * h->chunk_rows;
        const int k = rows_of(c) == h->plan_rows[0] ? 0 : 1;
        TJ_HIP(hipStreamWaitEvent(h->s_comp, h->ev_up[c & 1], 0));
        int e = tj_join(h->plan[k], h->dR + r0 * h->d, h->self_join ? nullptr : h->dS, T, h->ws, h->dpairs[c & 1],
                        pair_fmt, h->dstats + (c & 1) * TJ_STATS_SLOTS, h->s_comp);
        if (e) return e;
        TJ_HIP(hipMemcpyAsync(h->pstats + (c & 1) * TJ_STATS_SLOTS, h->dstats + (c & 1) * TJ_STATS_SLOTS,
                              TJ_STATS_SLOTS * 8, hipMemcpyDeviceToHost, h->s_comp));
        TJ_HIP(hipEventRecord(h->ev_done[c & 1], h->s_comp));
        return TJ_OK;
    };
    if (h->nR > 0 && (rc = stage(0))) return rc;
    if (h->nR > 0 && (rc = compute(0))) return rc;
    t_stage += now_ms() - ts;
    int64_t queued = 0;  // chunks [0, queued] are staged and enqueued
    for (int64_t c = 0; h->nR > 0 && c < h->nchunks; ++c) {
        if (c + 1 < h->nchunks && queued < c + 1) {
            ts = now_ms();
            if ((rc = stage(c + 1))) return rc;
            if ((rc = compute(c + 1))) return rc;
            queued = c + 1;
            t_stage += now_ms() - ts;
        }
        // collect chunk c
        double tw = now_ms();
        TJ_HIP(hipEventSynchronize(h->ev_done[c & 1]));
        t_wait += now_ms() - tw;
        const uint64_t* st = h->pstats + (c & 1) * TJ_STATS_SLOTS;
        TJ_CHECK(st[TJ_ST_INVALID] == 0, TJ_ERR_INVALID, "%llu non-finite input coordinates",
                 (unsigned long long)st[TJ_ST_INVALID]);
        const int64_t need_out = (int64_t)st[TJ_ST_OUT];
        const int64_t need_amb = (int64_t)st[TJ_ST_CANDIDATES];
        if (need_out > h->out_cap || (h->method == TJ_METHOD_CERTIFIED && need_amb > h->amb_cap)) {
            // grow, then run chunk c (and the queued c+1) again
            TJ_CHECK(need_out < (1ll << 32) && need_amb < (1ll << 32), TJ_ERR_UNSUPPORTED,
                     "a chunk needs more than 2^32 result slots");
            h->out_cap = std::max(h->out_cap, std::min<int64_t>((1ll << 32) - 1, need_out + need_out / 4 + 1024));
            h->amb_cap = std::max(h->amb_cap, std::min<int64_t>((1ll << 32) - 1, need_amb + need_amb / 4 + 1024));
            if ((rc = make_plans(h))) return rc;
            TJ_HIP(hipEventRecord(h->ev_up[c & 1], h->s_copy));  // rows are resident
            if ((rc = compute(c))) return rc;
            if (queued > c) {
                TJ_HIP(hipEventRecord(h->ev_up[(c + 1) & 1], h->s_copy));
                if ((rc = compute(c + 1))) return rc;
            }
            --c;  // collect chunk c again
            continue;
        }
        for (int k = 0; k < TJ_STATS_SLOTS; ++k) {
            if (k == TJ_ST_MAX_ERR || k == TJ_ST_MAX_RATIO) {
                const double a = __builtin_bit_cast(double, acc[k]), b = __builtin_bit_cast(double, st[k]);
                acc[k] = __builtin_bit_cast(uint64_t, std::max(a, b));
            } else if (k == TJ_ST_SCALE_EXP || (k >= TJ_ST_VIOL_KEY && k <= TJ_ST_VIOL_EST)) {
                if (st[k]) acc[k] = st[k];
            } else {
                acc[k] += st[k];
            }
        }
        double to = now_ms();
        if (out_format == TJ_OUT_U64_CANONICAL) {
            // one chunk: mirror + sort on the device, then copy the ids out
            const int64_t n = h->nR;
            size_t need = 0;
            if ((rc = tj_canonical_u64(nullptr, &need, nullptr, nullptr, need_out, n, nullptr, nullptr))) return rc;
            if ((rc = ensure_dev(&h->canon_temp, &h->canon_temp_bytes, need))) return rc;
            const size_t ids = (size_t)(2 * need_out + n);
            if ((rc = ensure_dev((void**)&h->dids, &h->dids_bytes, std::max<size_t>(ids, 1) * 8))) return rc;
            if ((rc = tj_canonical_u64(h->canon_temp, &h->canon_temp_bytes, (const int64_t*)h->dpairs[0],
                                       h->dstats + TJ_ST_OUT, need_out, n, h->dids, h->s_comp)))
                return rc;
            TJ_HIP(hipEventRecord(h->ev_done[0], h->s_comp));
            if ((rc = out.reserve(ids * 8))) return rc;
            if ((rc = download(h, out.buf, h->dids, ids * 8, h->ev_done[0], 8, 0))) return rc;
            out.bytes = ids * 8;
            *count = (int64_t)ids;
        } else {
            const size_t bytes = (size_t)need_out * 2 * elem;
            // the first chunk sizes the whole result (chunks are alike), later
            // ones grow it only if they are denser
            const size_t want = c == 0 ? bytes * (size_t)h->nchunks + bytes / 8 : out.bytes + bytes;
            if ((rc = out.reserve(std::max(want, out.bytes + bytes)))) return rc;
            if ((rc = download(h, out.buf + out.bytes, h->dpairs[c & 1], bytes, h->ev_done[c & 1], elem,
                               c * h->chunk_rows)))
                return rc;
            out.bytes += bytes;
            *count += need_out;
        }
        t_out += now_ms() - to;
    }
    if (h->nR == 0 && out_format == TJ_OUT_U64_CANONICAL) *count = 0;
    TJ_HIP(hipStreamSynchronize(h->s_copy));
    if (stats) memcpy(stats, acc, sizeof(acc));
    *result = out.buf;
    out.buf = nullptr;
    h->times[0] = now_ms() - t0;
    h->times[1] = t_stage;
    h->times[2] = t_wait;
    h->times[3] = t_out;
    return TJ_OK;
}

}  // extern "C"
