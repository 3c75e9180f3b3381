// plan_impl.h -- the plan of one join (shapes, decomposition, workspace
// layout) and the pipeline that enqueues it, shared by the plan entry points
// and the allocation-free join.
#pragma once
#include "tj_internal.h"

struct tj_plan {
    int64_t nR, nS;
    bool self_join;
    int d, dp, method, prec;
    int64_t out_cap, amb_cap;
    unsigned flags;
    int jbits, keybits;
    int bm, panels, col_splits, cols_per_split, grid;
    int tile_rb;           // 32-row blocks per wave of the screen (0 = the body's default)
    int64_t tile_split;    // column split width of an R x S screen (0 = the default rule)
    size_t off_grid;
    tj::GridLayout glay;
    size_t off_maxabs, off_params, off_metaR, off_metaS, off_qRh, off_qRl, off_qSh, off_qSl, off_out, off_amb,
        off_aux, off_sort, off_normR, off_normS, total;
    hipEvent_t ev[5];
    bool timing;
    hipGraphExec_t graph;
    bool has_graph;
};

namespace tj {
// One tj_join's arguments.  marks: record the plan's stage events.
struct JoinCall {
    tj_plan* p;
    const float* R;
    const float* S;
    double T;
    void* workspace;
    void* pairs;
    int out_format;
    uint64_t* stats;
    bool marks;
};

// Shapes, method, precision, capacities -> the plan's fields and workspace
// layout, in a caller-provided struct (no allocation of any kind).
int plan_init(tj_plan* p, int64_t nR, int64_t nS, int d, int method, int precision, int64_t out_capacity,
              int64_t amb_capacity, unsigned flags);
int join_check(const tj_plan* p, const float* R, const float* S, double T, void* workspace, void* pairs,
               int out_format, uint64_t* stats);
// Enqueue the whole join on s (no host synchronisation).
int join_enqueue(const JoinCall& c, hipStream_t s);
}  // namespace tj
