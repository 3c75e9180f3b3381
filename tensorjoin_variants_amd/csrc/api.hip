// api.hip -- the C ABI of libtjoin.so: plans, the join pipeline, and the
// exported stages.  See include/tjoin.h for the contract of each entry point.
//
// The pipeline of one certified join (all on the caller's stream, no host
// synchronisation between stages, so it can be captured in a hipGraph):
//   fill stats -> scale scan (R, S) -> parameter block -> row metadata (R, S)
//   -> persistent MFMA screen -> FP64 refinement of the ambiguous queue
//   -> radix sort of the output keys -> unpack to (i, j) rows.
// It replaces the host batch loop of Matrix.run (TG/src/g2_operator_r1.py:42-61),
// which synchronised four times per 4,096-tile batch.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <new>
#include "plan_impl.h"

namespace tj {
int screen_fp16(int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int query, int rb);
int screen_fp16x3(int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int query, int rb);
int screen_bf16(int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int query, int rb);
int screen_bf16x3(int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int query, int rb);
int screen_fp8(int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int query, int rb);
int screen_fp8x3(int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int query, int rb);
int launch_mirror(const int64_t* pairs, const uint64_t* count, int64_t cap, int64_t n, uint64_t* ids, uint64_t* total,
                  hipStream_t s);

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
int hip_fail(hipError_t e, const char* what) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return TJ_ERR_HIP;
}
int num_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

static int screen_call(int prec, int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int query,
                       int rb) {
    switch (prec) {
        case P_FP16: return screen_fp16(dp, self_join, a, grid, s, query, rb);
        case P_FP16X3: return screen_fp16x3(dp, self_join, a, grid, s, query, rb);
        case P_BF16: return screen_bf16(dp, self_join, a, grid, s, query, rb);
        case P_BF16X3: return screen_bf16x3(dp, self_join, a, grid, s, query, rb);
        case P_FP8: return screen_fp8(dp, self_join, a, grid, s, query, rb);
        case P_FP8X3: return screen_fp8x3(dp, self_join, a, grid, s, query, rb);
        default: set_error("unknown precision %d", prec); return TJ_ERR_INVALID;
    }
}
int launch_screen(int prec, int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int rb) {
    return screen_call(prec, dp, self_join, a, grid, s, 0, rb);
}
int screen_rows_per_wg(int prec, int dp, bool self_join, int rb) {
    ScreenArgs dummy{};
    return screen_call(prec, dp, self_join, dummy, 0, nullptr, 2, rb);
}
int screen_max_grid(int prec, int dp, bool self_join, int rb) {
    ScreenArgs dummy{};
    int occ = screen_call(prec, dp, self_join, dummy, 0, nullptr, 1, rb);
    if (occ < 1) occ = 1;
    return occ * num_cus();
}

int screen_layout(int64_t nA, int64_t nB, bool self_join, int dp, int prec, ScreenLayout* L, int rb,
                  int64_t split_cols) {
    const int es = prec_esize(prec);
    const bool x3 = prec_terms(prec) == 3;
    L->bm = screen_rows_per_wg(prec, dp, self_join, rb);
    if (L->bm <= 0) return L->bm < 0 ? L->bm : TJ_ERR_UNSUPPORTED;
    L->rb = rb;
    L->panels = (int)((nA + L->bm - 1) / L->bm);
    const int maxg = screen_max_grid(prec, dp, self_join, rb);
    const int64_t panels = std::max(L->panels, 1);
    int64_t cs;
    if (self_join) {
        // Column splits of the triangle's rows (panel p sweeps columns from
        // its first row on, split by split; splits past nB are empty items):
        // enough items to balance the persistent grid when a small join has
        // few panels, none narrower than 1024 columns.
        const int64_t want = (8ll * maxg + panels - 1) / panels;
        const int64_t by_width = (nB + 1023) / 1024;
        cs = std::max<int64_t>(1, std::min(want, by_width));
    } else {
        // Column splits of about 4 MiB of quantized S (one XCD's L2), items
        // ordered split-major: the workgroups running at any time sweep the
        // same chunk of S, which then stays in L2, instead of each panel
        // streaming all of S from memory.  Also enough items to balance the
        // persistent grid when there are few panels.
        const int64_t row_bytes = (int64_t)dp * es * (x3 ? 2 : 1);
        const int64_t by_l2 = std::max<int64_t>(1, (nB * row_bytes + (4ll << 20) - 1) / (4ll << 20));
        const int64_t want = (4ll * maxg + panels - 1) / panels;
        const int64_t by_width = (nB + 1023) / 1024;  // splits of at least 1024 columns
        cs = std::max<int64_t>(1, std::min(std::max(by_l2, want), by_width));
    }
    int64_t per = ((nB + cs - 1) / cs + 31) / 32 * 32;
    if (split_cols > 0) per = (split_cols + 31) / 32 * 32;  // the caller's column split width
    if (per > nB) per = (nB + 31) / 32 * 32;
    L->cols_per_split = (int)std::max<int64_t>(per, 32);
    L->col_splits = (int)((nB + L->cols_per_split - 1) / L->cols_per_split);
    const int64_t items = (int64_t)L->panels * L->col_splits;
    L->grid = (int)std::max<int64_t>(1, std::min<int64_t>(items, maxg));
    return TJ_OK;
}

// Record what fn enqueues into an instantiated graph.  The capture runs on a
// private stream (default priority) so that the caller's stream, which may
// be the legacy default stream, is never put in capture mode.
int capture_graph(hipGraphExec_t* exec, int (*fn)(void*, hipStream_t), void* ctx) {
    hipStream_t cs;
    TJ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return hip_fail(e, "hipStreamBeginCapture");
    }
    const int rc = fn(ctx, cs);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cs, &graph);
    (void)hipStreamDestroy(cs);
    if (rc != TJ_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    return TJ_OK;
}
}  // namespace tj

using namespace tj;


static size_t carve(size_t& cur, size_t bytes) {
    size_t at = (cur + 255) / 256 * 256;
    cur = at + bytes;
    return at;
}

extern "C" {

int tj_version(void) { return TJ_VERSION; }

const char* tj_last_error_string(void) { return g_err; }

int tj_device_info(int* cu_count, int* arch_major, int* arch_minor) {
    int dev = 0;
    hipDeviceProp_t prop;
    TJ_HIP(hipGetDevice(&dev));
    TJ_HIP(hipGetDeviceProperties(&prop, dev));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (arch_major) *arch_major = prop.major;
    if (arch_minor) *arch_minor = prop.minor;
    return TJ_OK;
}

}  // extern "C"

int tj::plan_init(tj_plan* p, int64_t nR, int64_t nS, int d, int method, int precision, int64_t out_capacity,
                  int64_t amb_capacity, unsigned flags) {
    TJ_CHECK(nR >= 0 && nS >= 0 && d >= 1, TJ_ERR_INVALID, "bad shape nR=%lld nS=%lld d=%d", (long long)nR,
             (long long)nS, d);
    TJ_CHECK(method == TJ_METHOD_CERTIFIED || method == TJ_METHOD_PLAIN_FP64 || method == TJ_METHOD_GRID ||
                 method == TJ_METHOD_PLAIN_FP32,
             TJ_ERR_INVALID, "method %d is not served by tj_join", method);
    TJ_CHECK(method != TJ_METHOD_GRID || d <= 3, TJ_ERR_UNSUPPORTED, "the grid method serves d <= 3 (d=%d)", d);
    TJ_CHECK(out_capacity >= 0 && out_capacity < (1ll << 32), TJ_ERR_INVALID, "out_capacity must be in [0, 2^32)");
    TJ_CHECK(amb_capacity >= 0 && amb_capacity < (1ll << 32), TJ_ERR_INVALID, "amb_capacity must be in [0, 2^32)");
    TJ_CHECK(nR < (1ll << 31) && nS < (1ll << 31), TJ_ERR_UNSUPPORTED, "point sets beyond 2^31 rows");
    p->nR = nR;
    p->self_join = (nS == 0);
    p->nS = p->self_join ? nR : nS;
    p->d = d;
    p->method = method;
    p->prec = method == TJ_METHOD_CERTIFIED ? precision : P_NONE;
    p->out_cap = out_capacity;
    p->amb_cap = method == TJ_METHOD_CERTIFIED ? amb_capacity : 0;
    p->flags = flags;
    p->jbits = bits_for(p->nS > 1 ? p->nS : 2);
    p->keybits = bits_for(p->nR > 1 ? p->nR : 2) + p->jbits;
    TJ_CHECK(p->keybits <= 63, TJ_ERR_UNSUPPORTED, "pair keys need %d bits", p->keybits);
    size_t cur = 0;
    p->off_maxabs = carve(cur, 256);
    p->off_params = carve(cur, sizeof(JoinParams));
    if (method == TJ_METHOD_CERTIFIED) {
        TJ_CHECK(p->prec >= P_FP16 && p->prec <= P_FP8X3, TJ_ERR_INVALID, "unknown precision %d", precision);
        p->dp = padded_dim(d, p->prec);
        const int es = prec_esize(p->prec);
        TJ_CHECK(p->dp > 0 && !(prec_terms(p->prec) == 3 && !prec_is_fp8(p->prec) && p->dp > 256), TJ_ERR_UNSUPPORTED,
                 "the screen serves d <= 512 (d <= 256 for a 16-bit hi/lo split); d=%d", d);
        const bool x3 = prec_terms(p->prec) == 3;
        p->off_metaR = carve(cur, sizeof(RowMeta) * (size_t)nR);
        p->off_qRh = carve(cur, (size_t)nR * p->dp * es);
        p->off_qRl = x3 ? carve(cur, (size_t)nR * p->dp * es) : p->off_qRh;
        if (!p->self_join) {
            p->off_metaS = carve(cur, sizeof(RowMeta) * (size_t)nS);
            p->off_qSh = carve(cur, (size_t)nS * p->dp * es);
            p->off_qSl = x3 ? carve(cur, (size_t)nS * p->dp * es) : p->off_qSh;
        } else {
            p->off_metaS = p->off_metaR;
            p->off_qSh = p->off_qRh;
            p->off_qSl = p->off_qRl;
        }
        p->off_amb = carve(cur, 8 * (size_t)(p->amb_cap > 0 ? p->amb_cap : 1));
        p->off_aux = carve(cur, 4 * (size_t)(p->amb_cap > 0 ? p->amb_cap : 1));
        ScreenLayout sl;
        int rc;
        if ((rc = screen_layout(nR, p->nS, p->self_join, p->dp, p->prec, &sl, 0, 0))) return rc;
        p->bm = sl.bm;
        p->panels = sl.panels;
        p->col_splits = sl.col_splits;
        p->cols_per_split = sl.cols_per_split;
        p->grid = sl.grid;
    }
    if (method == TJ_METHOD_PLAIN_FP32) {
        p->off_normR = carve(cur, plain32_norm_bytes(nR));
        p->off_normS = p->self_join ? p->off_normR : carve(cur, plain32_norm_bytes(nS));
    }
    if (method == TJ_METHOD_GRID) {
        p->glay = grid_layout(p->nR, p->nS, p->out_cap);
        p->off_grid = carve(cur, p->glay.total);
        p->off_out = p->off_sort = p->off_grid;
    } else {
        p->off_out = carve(cur, 8 * (size_t)(p->out_cap > 0 ? p->out_cap : 1));
        p->off_sort = carve(cur, sort_temp_bytes(p->out_cap));
    }
    p->total = (cur + 255) / 256 * 256;
    p->timing = false;
    p->has_graph = false;
    return TJ_OK;
}

extern "C" {

int tj_plan_create(tj_plan** out, int64_t nR, int64_t nS, int d, int method, int precision, int64_t out_capacity,
                   int64_t amb_capacity, unsigned flags) {
    TJ_CHECK(out != nullptr, TJ_ERR_INVALID, "plan pointer is NULL");
    *out = nullptr;
    tj_plan* p = new (std::nothrow) tj_plan();
    TJ_CHECK(p != nullptr, TJ_ERR_INVALID, "out of host memory");
    const int rc = plan_init(p, nR, nS, d, method, precision, out_capacity, amb_capacity, flags);
    if (rc != TJ_OK) {
        delete p;
        return rc;
    }
    if (flags & TJ_FLAG_TIMING) {
        for (int k = 0; k < 5; ++k) {
            if (hipEventCreate(&p->ev[k]) != hipSuccess) {
                for (int q = 0; q < k; ++q) (void)hipEventDestroy(p->ev[q]);
                delete p;
                set_error("hipEventCreate failed");
                return TJ_ERR_HIP;
            }
        }
        p->timing = true;
    }
    *out = p;
    return TJ_OK;
}

int tj_plan_destroy(tj_plan* plan) {
    if (plan && plan->has_graph) (void)hipGraphExecDestroy(plan->graph);
    if (plan && plan->timing)
        for (int k = 0; k < 5; ++k) (void)hipEventDestroy(plan->ev[k]);
    delete plan;
    return TJ_OK;
}

int tj_plan_set_tile(tj_plan* p, int row_blocks, int64_t cols_per_split) {
    TJ_CHECK(p != nullptr, TJ_ERR_INVALID, "plan is NULL");
    TJ_CHECK(p->method == TJ_METHOD_CERTIFIED, TJ_ERR_INVALID, "only a certified plan has a screen tile");
    TJ_CHECK(row_blocks >= 0 && cols_per_split >= 0, TJ_ERR_INVALID, "negative tile parameter");
    TJ_CHECK(cols_per_split <= ((int64_t)1 << 30), TJ_ERR_INVALID, "column split of %lld columns",
             (long long)cols_per_split);
    ScreenLayout sl;
    int rc;
    if ((rc = screen_layout(p->nR, p->nS, p->self_join, p->dp, p->prec, &sl, row_blocks, cols_per_split))) return rc;
    if (p->has_graph) {  // a captured graph holds the old launch shape
        (void)hipGraphExecDestroy(p->graph);
        p->has_graph = false;
    }
    p->tile_rb = row_blocks;
    p->tile_split = cols_per_split;
    p->bm = sl.bm;
    p->panels = sl.panels;
    p->col_splits = sl.col_splits;
    p->cols_per_split = sl.cols_per_split;
    p->grid = sl.grid;
    return TJ_OK;
}

int tj_plan_workspace_bytes(const tj_plan* plan, size_t* bytes) {
    TJ_CHECK(plan && bytes, TJ_ERR_INVALID, "NULL argument");
    *bytes = plan->total;
    return TJ_OK;
}

int tj_plan_describe(const tj_plan* p, char* buf, size_t len) {
    TJ_CHECK(p && buf, TJ_ERR_INVALID, "NULL argument");
    snprintf(buf, len,
             "{\"nR\":%lld,\"nS\":%lld,\"self\":%d,\"d\":%d,\"dp\":%d,\"method\":%d,\"precision\":%d,"
             "\"jbits\":%d,\"keybits\":%d,\"rows_per_wg\":%d,\"panels\":%d,\"col_splits\":%d,"
             "\"cols_per_split\":%d,\"grid\":%d,\"row_blocks\":%d,\"workspace\":%zu}",
             (long long)p->nR, (long long)p->nS, (int)p->self_join, p->d, p->dp, p->method, p->prec, p->jbits,
             p->keybits, p->bm, p->panels, p->col_splits, p->cols_per_split, p->grid, p->bm / 128, p->total);
    return TJ_OK;
}

int tj_plan_stage_times(tj_plan* p, float* ms, int n) {
    TJ_CHECK(p && ms, TJ_ERR_INVALID, "NULL argument");
    TJ_CHECK(p->timing, TJ_ERR_INVALID, "plan was created without TJ_FLAG_TIMING");
    TJ_HIP(hipEventSynchronize(p->ev[4]));
    for (int k = 0; k < n && k < 4; ++k) TJ_HIP(hipEventElapsedTime(&ms[k], p->ev[k], p->ev[k + 1]));
    return TJ_OK;
}

}  // extern "C"

int tj::join_check(const tj_plan* p, const float* R, const float* S, double T, void* workspace, void* pairs,
                   int out_format, uint64_t* stats) {
    TJ_CHECK(p != nullptr, TJ_ERR_INVALID, "plan is NULL");
    TJ_CHECK(workspace != nullptr && stats != nullptr, TJ_ERR_INVALID, "workspace and stats are required");
    TJ_CHECK(R != nullptr || p->nR == 0, TJ_ERR_INVALID, "R is NULL");
    TJ_CHECK(p->self_join || S != nullptr || p->nS == 0, TJ_ERR_INVALID, "S is NULL for an R x S plan");
    TJ_CHECK(pairs != nullptr || p->out_cap == 0, TJ_ERR_INVALID, "pairs buffer is NULL");
    TJ_CHECK(out_format == TJ_OUT_I32 || out_format == TJ_OUT_I64, TJ_ERR_INVALID, "bad out_format %d", out_format);
    TJ_CHECK(out_format == TJ_OUT_I64 || (p->nR < (1ll << 31) && p->nS < (1ll << 31)), TJ_ERR_INVALID,
             "int32 output needs fewer than 2^31 rows");
    TJ_CHECK(T == T && T >= 0.0 && T < INFINITY, TJ_ERR_INVALID, "T must be finite and >= 0");
    return TJ_OK;
}


int tj::join_enqueue(const JoinCall& c, hipStream_t s) {
    tj_plan* p = c.p;
    const float* R = c.R;
    const float* S = c.S;
    const double T = c.T;
    void* pairs = c.pairs;
    const int out_format = c.out_format;
    uint64_t* stats = c.stats;
    uint8_t* ws = (uint8_t*)c.workspace;
    if (p->self_join) S = R;
    uint32_t* maxabs = (uint32_t*)(ws + p->off_maxabs);
    JoinParams* params = (JoinParams*)(ws + p->off_params);
    uint64_t* out_keys = (uint64_t*)(ws + p->off_out);
    int rc;
    if ((rc = launch_fill_u64(stats, TJ_STATS_SLOTS, 0, s))) return rc;
    if ((rc = launch_fill_u64((uint64_t*)maxabs, 1, 0, s))) return rc;
    const int64_t pairs_total = p->self_join ? p->nR * (p->nR - 1) / 2 : p->nR * p->nS;
    if ((rc = launch_fill_u64(stats + TJ_ST_PAIRS, 1, (uint64_t)pairs_total, s))) return rc;
#define TJ_MARK(k) \
    if (c.marks && p->timing) TJ_HIP(hipEventRecord(p->ev[k], s))
    TJ_MARK(0);

    if (p->method == TJ_METHOD_CERTIFIED) {
        RowMeta* mR = (RowMeta*)(ws + p->off_metaR);
        RowMeta* mS = (RowMeta*)(ws + p->off_metaS);
        if ((rc = launch_scale_scan(R, p->nR, p->d, maxabs, stats, s))) return rc;
        if (!p->self_join && (rc = launch_scale_scan(S, p->nS, p->d, maxabs, stats, s))) return rc;
        if ((rc = launch_params(maxabs, T, p->d, p->dp, p->prec, params, stats, s))) return rc;
        if ((rc = launch_rowmeta(R, p->nR, p->d, p->dp, p->prec, params, ws + p->off_qRh, ws + p->off_qRl, mR, stats, s)))
            return rc;
        if (!p->self_join &&
            (rc = launch_rowmeta(S, p->nS, p->d, p->dp, p->prec, params, ws + p->off_qSh, ws + p->off_qSl, mS, stats, s)))
            return rc;
        TJ_MARK(1);
        if (p->nR > 0 && p->nS > 0) {
            ScreenArgs a{};
            a.qa_hi = ws + p->off_qRh; a.qa_lo = ws + p->off_qRl; a.ma = mR; a.na = p->nR;
            a.qb_hi = ws + p->off_qSh; a.qb_lo = ws + p->off_qSl; a.mb = mS; a.nb = p->nS;
            a.params = params;
            a.out_keys = out_keys; a.out_cap = p->out_cap;
            a.amb_keys = (uint64_t*)(ws + p->off_amb); a.amb_aux = (float*)(ws + p->off_aux); a.amb_cap = p->amb_cap;
            a.stats = stats;
            a.jbits = p->jbits;
            a.panels = p->panels; a.col_splits = p->col_splits; a.cols_per_split = p->cols_per_split;
            a.audit = (p->flags & TJ_FLAG_AUDIT) ? 1 : 0;
            if ((rc = launch_screen(p->prec, p->dp, p->self_join, a, p->grid, s, p->tile_rb))) return rc;
            TJ_MARK(2);
            if ((rc = launch_certify(R, S, p->d, p->jbits, a.amb_keys, a.amb_aux, p->amb_cap, mR, mS, params, out_keys,
                                     p->out_cap, stats, a.audit, s)))
                return rc;
        } else {
            TJ_MARK(2);
        }
        TJ_MARK(3);
    } else if (p->method == TJ_METHOD_GRID) {
        // (the grid's bounds pass counts the non-finite coordinates)
        TJ_MARK(1);
        if ((rc = launch_grid_join(R, p->nR, S, p->nS, p->d, T, p->self_join, p->self_join, ws + p->off_grid, p->glay,
                                   p->out_cap, out_format, pairs, nullptr, stats, s)))
            return rc;
        TJ_MARK(2);
        TJ_MARK(3);
        TJ_MARK(4);
        return TJ_OK;
    } else if (p->method == TJ_METHOD_PLAIN_FP32) {
        TJ_MARK(1);
        if ((rc = launch_plain32(R, p->nR, S, p->nS, p->d, p->jbits, T, p->self_join, ws + p->off_normR,
                                 ws + p->off_normS, out_keys, p->out_cap, stats, s)))
            return rc;
        TJ_MARK(2);
        TJ_MARK(3);
    } else {
        TJ_MARK(1);
        if ((rc = launch_plain(R, p->nR, S, p->nS, p->d, p->jbits, T, p->self_join, out_keys, p->out_cap, stats, s)))
            return rc;
        TJ_MARK(2);
        TJ_MARK(3);
    }
    uint64_t* sorted = out_keys;
    if ((rc = launch_sort(ws + p->off_sort, out_keys, stats + TJ_ST_OUT, p->out_cap, p->keybits, s, &sorted))) return rc;
    if ((rc = launch_unpack(sorted, stats + TJ_ST_OUT, p->out_cap, p->jbits, pairs, out_format, s))) return rc;
    TJ_MARK(4);
#undef TJ_MARK
    return TJ_OK;
}

extern "C" {

int tj_join(tj_plan* p, const float* R, const float* S, double T, void* workspace, void* pairs, int out_format,
            uint64_t* stats, void* stream) {
    int rc = join_check(p, R, S, T, workspace, pairs, out_format, stats);
    if (rc) return rc;
    const JoinCall c{p, R, S, T, workspace, pairs, out_format, stats, true};
    return join_enqueue(c, (hipStream_t)stream);
}

static int join_capture_fn(void* ctx, hipStream_t s) { return join_enqueue(*(const JoinCall*)ctx, s); }

int tj_plan_capture(tj_plan* p, const float* R, const float* S, double T, void* workspace, void* pairs,
                    int out_format, uint64_t* stats) {
    int rc = join_check(p, R, S, T, workspace, pairs, out_format, stats);
    if (rc) return rc;
    JoinCall c{p, R, S, T, workspace, pairs, out_format, stats, false};
    hipGraphExec_t exec;
    if ((rc = capture_graph(&exec, join_capture_fn, &c))) return rc;
    if (p->has_graph) (void)hipGraphExecDestroy(p->graph);
    p->graph = exec;
    p->has_graph = true;
    return TJ_OK;
}

int tj_plan_replay(tj_plan* p, void* stream) {
    TJ_CHECK(p != nullptr, TJ_ERR_INVALID, "plan is NULL");
    TJ_CHECK(p->has_graph, TJ_ERR_INVALID, "no graph captured for this plan");
    TJ_HIP(hipGraphLaunch(p->graph, (hipStream_t)stream));
    return TJ_OK;
}

int tj_scale_scan(const float* x, int64_t n, int d, uint32_t* maxabs_bits, uint64_t* stats, void* stream) {
    TJ_CHECK(maxabs_bits && stats && (x || n == 0), TJ_ERR_INVALID, "NULL argument");
    return launch_scale_scan(x, n, d, maxabs_bits, stats, (hipStream_t)stream);
}

int tj_params(const uint32_t* maxabs_bits, double T, int d, int precision, void* params, uint64_t* stats, void* stream) {
    TJ_CHECK(maxabs_bits && params && stats, TJ_ERR_INVALID, "NULL argument");
    int dp = precision == P_NONE ? d : padded_dim(d, precision);
    TJ_CHECK(dp > 0, TJ_ERR_UNSUPPORTED, "d=%d not served at precision %d", d, precision);
    return launch_params(maxabs_bits, T, d, dp, precision, (JoinParams*)params, stats, (hipStream_t)stream);
}

int tj_rowmeta(const float* x, int64_t n, int d, int precision, const void* params, void* q_hi, void* q_lo, float* meta,
               uint64_t* stats, void* stream) {
    TJ_CHECK(params && q_hi && meta && (x || n == 0), TJ_ERR_INVALID, "NULL argument");
    int dp = padded_dim(d, precision);
    TJ_CHECK(dp > 0, TJ_ERR_UNSUPPORTED, "d=%d not served at precision %d", d, precision);
    TJ_CHECK(prec_terms(precision) == 1 || q_lo, TJ_ERR_INVALID, "hi/lo precision needs q_lo");
    return launch_rowmeta(x, n, d, dp, precision, (const JoinParams*)params, q_hi, q_lo ? q_lo : q_hi, (RowMeta*)meta,
                          stats, (hipStream_t)stream);
}

int tj_rowmeta_sizes(int64_t n, int d, int precision, size_t* q_bytes, size_t* meta_bytes) {
    TJ_CHECK(q_bytes && meta_bytes, TJ_ERR_INVALID, "NULL argument");
    TJ_CHECK(n >= 0 && d >= 1, TJ_ERR_INVALID, "bad shape n=%lld d=%d", (long long)n, d);
    TJ_CHECK(precision >= P_FP16 && precision <= P_FP8X3, TJ_ERR_INVALID, "unknown precision %d", precision);
    const int dp = padded_dim(d, precision);
    TJ_CHECK(dp > 0, TJ_ERR_UNSUPPORTED, "d=%d not served at precision %d", d, precision);
    *q_bytes = (size_t)n * dp * prec_esize(precision);
    *meta_bytes = sizeof(RowMeta) * (size_t)n;
    return TJ_OK;
}

int tj_screen(int precision, int d, const void* qa_hi, const void* qa_lo, const float* meta_a, int64_t na,
              const void* qb_hi, const void* qb_lo, const float* meta_b, int64_t nb, const void* params,
              uint64_t* out_keys, int64_t out_capacity, uint64_t* cand_keys, float* cand_s, int64_t cand_capacity,
              unsigned flags, uint64_t* stats, void* stream) {
    const bool self_join = (nb == 0);
    TJ_CHECK(precision >= P_FP16 && precision <= P_FP8X3, TJ_ERR_INVALID, "unknown precision %d", precision);
    TJ_CHECK(na >= 0 && nb >= 0 && na < (1ll << 31) && nb < (1ll << 31), TJ_ERR_INVALID, "bad sizes");
    TJ_CHECK(out_capacity >= 0 && out_capacity < (1ll << 32) && cand_capacity >= 0 && cand_capacity < (1ll << 32),
             TJ_ERR_INVALID, "capacities must be in [0, 2^32)");
    TJ_CHECK(params && stats && (out_keys || out_capacity == 0) && ((cand_keys && cand_s) || cand_capacity == 0),
             TJ_ERR_INVALID, "NULL argument");
    TJ_CHECK(na == 0 || (qa_hi && meta_a), TJ_ERR_INVALID, "rows of A are NULL");
    TJ_CHECK(self_join || (qb_hi && meta_b), TJ_ERR_INVALID, "rows of B are NULL");
    const int x3 = prec_terms(precision) == 3;
    TJ_CHECK(!x3 || na == 0 || (qa_lo && (self_join || qb_lo)), TJ_ERR_INVALID, "hi/lo precision needs the lo codes");
    const int dp = padded_dim(d, precision);
    TJ_CHECK(dp > 0 && (!x3 || prec_is_fp8(precision) || dp <= 256), TJ_ERR_UNSUPPORTED,
             "d=%d not served at precision %d", d, precision);
    const int64_t nB = self_join ? na : nb;
    if (na == 0 || nB == 0) return TJ_OK;
    ScreenLayout sl;
    int rc;
    if ((rc = screen_layout(na, nB, self_join, dp, precision, &sl, 0, 0))) return rc;
    hipStream_t s = (hipStream_t)stream;
    ScreenArgs a{};
    a.qa_hi = qa_hi; a.qa_lo = x3 ? qa_lo : qa_hi; a.ma = (const RowMeta*)meta_a; a.na = na;
    if (self_join) {
        a.qb_hi = a.qa_hi; a.qb_lo = a.qa_lo; a.mb = a.ma;
    } else {
        a.qb_hi = qb_hi; a.qb_lo = x3 ? qb_lo : qb_hi; a.mb = (const RowMeta*)meta_b;
    }
    a.nb = nB;
    a.params = (const JoinParams*)params;
    a.out_keys = out_keys; a.out_cap = out_capacity;
    a.amb_keys = cand_keys; a.amb_aux = cand_s; a.amb_cap = cand_capacity;
    a.stats = stats;
    a.jbits = bits_for(nB > 1 ? nB : 2);
    a.panels = sl.panels; a.col_splits = sl.col_splits; a.cols_per_split = sl.cols_per_split;
    a.audit = (flags & TJ_FLAG_AUDIT) ? 1 : 0;
    // the persistent grid takes its items from a cursor that starts at zero
    if ((rc = launch_fill_u64(stats + TJ_ST_WORK, 1, 0, s))) return rc;
    return launch_screen(precision, dp, self_join, a, sl.grid, s, 0);
}

int tj_certify(const float* R, const float* S, int d, int64_t nS, const uint64_t* cand_keys, const float* cand_s,
               int64_t cand_capacity, const float* meta_R, const float* meta_S, const void* params, uint64_t* out_keys,
               int64_t out_capacity, unsigned flags, uint64_t* stats, void* stream) {
    TJ_CHECK(R && meta_R && params && stats && (out_keys || out_capacity == 0) &&
                 ((cand_keys && cand_s) || cand_capacity == 0),
             TJ_ERR_INVALID, "NULL argument");
    TJ_CHECK(nS >= 0 && nS < (1ll << 31) && d >= 1, TJ_ERR_INVALID, "bad shape");
    TJ_CHECK(cand_capacity >= 0 && cand_capacity < (1ll << 32) && out_capacity >= 0 && out_capacity < (1ll << 32),
             TJ_ERR_INVALID, "capacities must be in [0, 2^32)");
    const bool self_join = (S == nullptr);
    TJ_CHECK(self_join || meta_S, TJ_ERR_INVALID, "meta_S is NULL for an R x S join");
    return launch_certify(R, self_join ? R : S, d, bits_for(nS > 1 ? nS : 2), cand_keys, cand_s, cand_capacity,
                          (const RowMeta*)meta_R, (const RowMeta*)(self_join ? meta_R : meta_S),
                          (const JoinParams*)params, out_keys, out_capacity, stats, (flags & TJ_FLAG_AUDIT) ? 1 : 0,
                          (hipStream_t)stream);
}

int tj_refine_fp64(const float* R, const float* S, int d, int jbits, double T, const uint64_t* keys,
                   const uint64_t* count, int64_t capacity, uint64_t* out_keys, int64_t out_capacity, uint64_t* stats,
                   void* stream) {
    TJ_CHECK(R && keys && count && stats && (out_keys || out_capacity == 0), TJ_ERR_INVALID, "NULL argument");
    TJ_CHECK(d >= 1 && jbits >= 1 && jbits <= 32, TJ_ERR_INVALID, "bad d=%d or jbits=%d", d, jbits);
    TJ_CHECK(T == T && T >= 0.0, TJ_ERR_INVALID, "T must be >= 0");
    TJ_CHECK(capacity >= 0 && capacity < (1ll << 32) && out_capacity >= 0 && out_capacity < (1ll << 32),
             TJ_ERR_INVALID, "capacities must be in [0, 2^32)");
    return launch_refine(R, S ? S : R, d, jbits, T, keys, count, capacity, out_keys, out_capacity, stats,
                         (hipStream_t)stream);
}

int tj_plain_fp64(const float* R, int64_t nR, const float* S, int64_t nS, int d, int jbits, double T, uint64_t* out_keys,
                  int64_t out_capacity, uint64_t* stats, void* stream) {
    TJ_CHECK(R && out_keys && stats, TJ_ERR_INVALID, "NULL argument");
    const bool self_join = (S == nullptr);
    return launch_plain(R, nR, self_join ? R : S, self_join ? nR : nS, d, jbits, T, self_join, out_keys, out_capacity,
                        stats, (hipStream_t)stream);
}

int tj_grid_workspace_bytes(int64_t nR, int64_t nS, int64_t capacity, size_t* bytes) {
    TJ_CHECK(bytes != nullptr, TJ_ERR_INVALID, "bytes is NULL");
    TJ_CHECK(nR >= 0 && nS >= 0 && nR < (1ll << 31) && nS < (1ll << 31), TJ_ERR_INVALID, "bad sizes");
    TJ_CHECK(capacity >= 0 && capacity < (1ll << 32), TJ_ERR_INVALID, "capacity must be in [0, 2^32)");
    *bytes = grid_layout(nR, nS == 0 ? nR : nS, capacity).total;
    return TJ_OK;
}

static int grid_stage(const float* R, int64_t nR, const float* S, int64_t nS, int d, double T, bool upper,
                      void* workspace, size_t workspace_bytes, int64_t capacity, int emit, void* out, uint32_t* counts,
                      uint64_t* stats, hipStream_t s) {
    TJ_CHECK(d >= 1 && d <= 3, TJ_ERR_UNSUPPORTED, "the grid serves d <= 3 (d=%d)", d);
    TJ_CHECK(nR >= 0 && nS >= 0 && nR < (1ll << 31) && nS < (1ll << 31), TJ_ERR_INVALID, "bad sizes");
    TJ_CHECK(capacity >= 0 && capacity < (1ll << 32), TJ_ERR_INVALID, "capacity must be in [0, 2^32)");
    TJ_CHECK(T == T && T >= 0.0 && T < INFINITY, TJ_ERR_INVALID, "T must be finite and >= 0");
    TJ_CHECK(stats && workspace && (R || nR == 0) && (out || capacity == 0), TJ_ERR_INVALID, "NULL argument");
    const bool self_join = (nS == 0);
    TJ_CHECK(self_join || S, TJ_ERR_INVALID, "S is NULL for an R x S join");
    const GridLayout L = grid_layout(nR, self_join ? nR : nS, capacity);
    TJ_CHECK(workspace_bytes >= L.total, TJ_ERR_WORKSPACE, "grid workspace needs %zu bytes", L.total);
    int rc;
    if ((rc = launch_fill_u64(stats, TJ_STATS_SLOTS, 0, s))) return rc;
    const int64_t nB = self_join ? nR : nS;
    const int64_t pairs_total = (self_join && upper) ? nR * (nR - 1) / 2 : nR * nB;
    if ((rc = launch_fill_u64(stats + TJ_ST_PAIRS, 1, (uint64_t)pairs_total, s))) return rc;
    return launch_grid_join(R, nR, self_join ? R : S, nB, d, T, self_join, upper, (uint8_t*)workspace, L, capacity,
                            emit, out, counts, stats, s);
}

int tj_grid_join(const float* R, int64_t nR, const float* S, int64_t nS, int d, double T, void* workspace,
                 size_t workspace_bytes, int64_t capacity, void* pairs, int out_format, uint64_t* stats, void* stream) {
    TJ_CHECK(out_format == TJ_OUT_I32 || out_format == TJ_OUT_I64, TJ_ERR_INVALID, "bad out_format %d", out_format);
    return grid_stage(R, nR, S, nS, d, T, nS == 0, workspace, workspace_bytes, capacity, out_format, pairs, nullptr,
                      stats, (hipStream_t)stream);
}

int tj_grid_neighbors(const float* X, int64_t n, int d, double T, void* workspace, size_t workspace_bytes,
                      int64_t capacity, uint32_t* counts, uint32_t* table, uint64_t* stats, void* stream) {
    TJ_CHECK(counts || n == 0, TJ_ERR_INVALID, "counts is NULL");
    return grid_stage(X, n, nullptr, 0, d, T, false, workspace, workspace_bytes, capacity, GRID_EMIT_TABLE, table,
                      counts, stats, (hipStream_t)stream);
}

int tj_sort_keys(void* temp, size_t* temp_bytes, uint64_t* keys, const uint64_t* count, int64_t capacity, int bits,
                 void* stream) {
    TJ_CHECK(temp_bytes != nullptr, TJ_ERR_INVALID, "temp_bytes is NULL");
    TJ_CHECK(capacity >= 0 && capacity < (1ll << 32), TJ_ERR_INVALID, "capacity must be in [0, 2^32)");
    TJ_CHECK(bits >= 0 && bits <= 64, TJ_ERR_INVALID, "bits must be in [0, 64]");
    if (temp == nullptr) { *temp_bytes = sort_temp_bytes(capacity); return TJ_OK; }
    TJ_CHECK(*temp_bytes >= sort_temp_bytes(capacity), TJ_ERR_WORKSPACE, "sort temp too small");
    TJ_CHECK(keys && count, TJ_ERR_INVALID, "NULL argument");
    return launch_sort(temp, keys, count, capacity, bits, (hipStream_t)stream);
}

int tj_unpack_pairs(const uint64_t* keys, const uint64_t* count, int64_t capacity, int jbits, void* pairs, int out_format,
                    void* stream) {
    TJ_CHECK(keys && count && pairs, TJ_ERR_INVALID, "NULL argument");
    TJ_CHECK(out_format == TJ_OUT_I32 || out_format == TJ_OUT_I64, TJ_ERR_INVALID, "bad out_format");
    return launch_unpack(keys, count, capacity, jbits, pairs, out_format, (hipStream_t)stream);
}

int tj_canonical_u64(void* temp, size_t* temp_bytes, const int64_t* pairs_i64, const uint64_t* count, int64_t capacity,
                     int64_t n, uint64_t* ids, void* stream) {
    TJ_CHECK(temp_bytes != nullptr, TJ_ERR_INVALID, "temp_bytes is NULL");
    TJ_CHECK(n >= 0 && n < (1ll << 31), TJ_ERR_INVALID, "bad n");
    const int64_t idcap = 2 * capacity + n;
    TJ_CHECK(idcap < (1ll << 32), TJ_ERR_INVALID, "capacity too large");
    const size_t need = sort_temp_bytes(idcap) + 256;
    if (temp == nullptr) { *temp_bytes = need; return TJ_OK; }
    TJ_CHECK(*temp_bytes >= need, TJ_ERR_WORKSPACE, "canonical temp too small");
    TJ_CHECK(pairs_i64 && count && ids, TJ_ERR_INVALID, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    uint64_t* total = (uint64_t*)temp;
    void* sort_temp = (uint8_t*)temp + 256;
    int rc;
    if ((rc = launch_mirror(pairs_i64, count, capacity, n, ids, total, s))) return rc;
    const int bits = bits_for(n > 1 ? n * n : 2);
    return launch_sort(sort_temp, ids, total, idcap, bits, s);
}

}  // extern "C"
