// batch.hip -- many independent self-joins of uneven sizes and widths in one
// pipeline: one scale scan, one parameter pass, one row-metadata pass, one
// persistent screen launch over every join's panels (one per precision
// family present: FP16, FP8), one certificate pass, and the output grouped
// by row (per-row counts, offsets, scatter, per-row sort) with per-join
// offsets.
//
// The reference runs one join at a time and slices its triangular tile list
// into host batches of 4,096 tiles with a blocking count read after every
// stage (CEJ/src/run_g5_tensorjoin_public.py:377-381,410-535;
// TG/src/g2_operator_r1.py:45-57).  Here the unit of work is a panel of one
// join; the work list of all joins is ordered largest first (LPT) at plan
// creation and drained by a persistent grid through an atomic cursor, so
// joins of 64 rows and of 16,384 rows, of width 16 and of width 256, share
// the machine without a host round trip.  Every count stays on the device:
// the whole sequence is capturable in a hipGraph (tj_batch_capture).
//
// Row spaces.  Joins are numbered in caller order; their rows are numbered
// consecutively in that order (the batch row space, gbase = first row of a
// join).  Joins of one (precision, padded width) form a group whose quantized
// rows and metadata are stored contiguously (grow = first row of a join in
// its group).  A pair key is (batch row of i) << jbits | (local j); rows in
// batch order with their ids sorted are the output ordered by join, then
// (i, j).
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <new>
#include <vector>
#include "certify_impl.h"
#include "meta_impl.h"

namespace tjbatch {

struct BatchJoin {
    int64_t gbase;     // first row in the batch row space
    int64_t grow;      // first row in the group's arrays
    int64_t ball_off;  // first float of the join's panel balls (tile pruning)
    int32_t n, d, dp, prec;
    int32_t group, bm;
};
struct BatchGroup {
    int64_t q_off, meta_off;  // workspace offsets of the quantized rows and metadata
    int32_t combo, prec, dp, pad;
};

// (precision, padded width) instantiations of the batched screen body, in
// four families, one persistent kernel each: a kernel holding more bodies
// exceeds 256 VGPRs and spills.  A batch whose joins all fall in one family
// (widths 16/64/256 at FP16, say) is a single launch.
#define TJ_BATCH_F16A(X) X(0, P_FP16, 16) X(2, P_FP16, 64) X(4, P_FP16, 256)
#define TJ_BATCH_F16B(X) X(1, P_FP16, 32) X(3, P_FP16, 128)
#define TJ_BATCH_F8A(X) X(5, P_FP8, 64) X(7, P_FP8, 256)
#define TJ_BATCH_F8B(X) X(6, P_FP8, 128)
#define TJ_BATCH_COMBOS(X) TJ_BATCH_F16A(X) TJ_BATCH_F16B(X) TJ_BATCH_F8A(X) TJ_BATCH_F8B(X)
constexpr int N_FAMILIES = 4;
inline int family_of(int prec, int dp) {
    if (prec == P_FP8) return dp == 128 ? 3 : 2;
    return (dp == 32 || dp == 128) ? 1 : 0;
}

constexpr int cmax(int a, int b) { return a > b ? a : b; }
#define TJ_X_RING(ID, PR, DPV) cmax(tjscreen::ring_bytes<PR, DPV>(),
#define TJ_X_CLOSE(ID, PR, DPV) )
constexpr int RING_MAX = TJ_BATCH_COMBOS(TJ_X_RING) 0 TJ_BATCH_COMBOS(TJ_X_CLOSE);
#define TJ_X_NST(ID, PR, DPV) cmax(tjscreen::Geo<PR, DPV>::NSTAGE,
constexpr int NSTAGE_MAX = TJ_BATCH_COMBOS(TJ_X_NST) 0 TJ_BATCH_COMBOS(TJ_X_CLOSE);
#undef TJ_X_RING
#undef TJ_X_NST
#undef TJ_X_CLOSE

int combo_of(int prec, int dp) {
#define X(ID, PR, DPV) \
    if (prec == PR && dp == DPV) return ID;
    TJ_BATCH_COMBOS(X)
#undef X
    return -1;
}
int rows_per_wg(int combo) {
    switch (combo) {
#define X(ID, PR, DPV) \
    case ID: return 4 * tjscreen::rb_batch<PR, DPV>() * 32;
        TJ_BATCH_COMBOS(X)
#undef X
        default: return 0;
    }
}

// Work units of the per-row passes.  A group of GW = group_width(dp) lanes
// handles one row, so a wave holds 64 / GW rows of one join at a time; the
// scale scan takes SCAN_SUB such row sets per unit (one atomic per unit),
// the row-metadata pass 64 rows per unit (lane L finishes row L), the
// spatial order ORDER_ROWS rows, one per lane.  Chunk prefix tables: units
// of joins 0..t-1 (device int64 [joins + 1]).
constexpr int SCAN_SUB = 8;

// Rows of one row-metadata unit (one wave): 16 steps of 64 / width rows.
__host__ __device__ inline int64_t meta_unit_rows(int dp) { return 16 * (64 / tjmeta::meta_width(dp)); }

__device__ inline int find_join(const int64_t* __restrict__ cs, int nj, int64_t c) {
    int lo = 0, hi = nj;  // cs[lo] <= c < cs[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cs[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------ scale scan
// max |x| per join (binary32 bits) and the count of non-finite coordinates.
// Each wave takes a contiguous range of scan units, reads each join's part
// of it as one flat run of 16-byte loads, and flushes the join's maximum
// with one atomic when it moves to the next join (units of one join are
// consecutive; a strided unit order cost one atomic per unit).
__global__ void __launch_bounds__(256) batch_scan_kernel(const float* const* __restrict__ xs,
                                                         const BatchJoin* __restrict__ joins,
                                                         const int64_t* __restrict__ cstart, int nj,
                                                         uint32_t* __restrict__ maxabs, uint64_t* __restrict__ stats) {
    const int lane = threadIdx.x & 63;
    const int64_t wv = (int64_t)blockIdx.x * 4 + tjscreen::uni((int)(threadIdx.x >> 6));
    const int64_t nw = (int64_t)gridDim.x * 4;
    const int64_t units = cstart[nj];
    const int64_t per = (units + nw - 1) / nw;
    const int64_t u0 = wv * per;
    const int64_t u1 = u0 + per < units ? u0 + per : units;
    uint32_t bad = 0;
    for (int64_t c = u0; c < u1;) {
        const int t = tjscreen::uni(find_join(cstart, nj, c));
        const BatchJoin J = joins[t];
        const int64_t rows_per_unit = (int64_t)SCAN_SUB * (64 / tjmeta::group_width(J.dp));
        const int64_t ce = u1 < cstart[t + 1] ? u1 : cstart[t + 1];
        const int64_t r0 = (c - cstart[t]) * rows_per_unit;
        int64_t r1 = (ce - cstart[t]) * rows_per_unit;
        if (r1 > J.n) r1 = J.n;
        const float* base = xs[t] + r0 * (int64_t)J.d;
        const int64_t cnt = r1 > r0 ? (r1 - r0) * (int64_t)J.d : 0;
        uint32_t m = 0;
        auto take = [&](float f) {
            const uint32_t u = __builtin_bit_cast(uint32_t, f) & 0x7FFFFFFFu;
            const bool nf = u >= 0x7F800000u;  // inf or NaN: counted, not a maximum
            bad += nf ? 1u : 0u;
            m = (!nf && u > m) ? u : m;
        };
        int64_t head = (int64_t)(((16 - ((uintptr_t)base & 15)) & 15) / 4);
        if (head > cnt) head = cnt;
        const int64_t nv = (cnt - head) / 4;
        const float4* xv = (const float4*)(base + head);
        // four 16-byte loads per lane in flight, then the remainder
        int64_t k = lane;
        for (; k + 192 < nv; k += 256) {
            const float4 v0 = xv[k], v1 = xv[k + 64], v2 = xv[k + 128], v3 = xv[k + 192];
            take(v0.x); take(v0.y); take(v0.z); take(v0.w);
            take(v1.x); take(v1.y); take(v1.z); take(v1.w);
            take(v2.x); take(v2.y); take(v2.z); take(v2.w);
            take(v3.x); take(v3.y); take(v3.z); take(v3.w);
        }
        for (; k < nv; k += 64) {
            const float4 v = xv[k];
            take(v.x); take(v.y); take(v.z); take(v.w);
        }
        for (int64_t k = lane; k < head; k += 64) take(base[k]);
        for (int64_t k = head + 4 * nv + lane; k < cnt; k += 64) take(base[k]);
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t om = __shfl_xor(m, o);
            m = om > m ? om : m;
        }
        if (lane == 0 && m) atomicMax(&maxabs[t], m);
        c = ce;
    }
    for (int o = 32; o > 0; o >>= 1) bad += __shfl_xor(bad, o);
    if (lane == 0 && bad) atomicAdd((unsigned long long*)&stats[TJ_ST_INVALID], (unsigned long long)bad);
}

// ------------------------------------------------------------- parameters
// One thread per join.  A threshold that is not finite and >= 0 counts as
// an invalid input (the host layer rejects such calls before they start).
__global__ void batch_params_kernel(const BatchJoin* __restrict__ joins, int nj, const double* __restrict__ T,
                                    const uint32_t* __restrict__ maxabs, JoinParams* __restrict__ params,
                                    uint64_t* __restrict__ stats) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nj) return;
    const BatchJoin J = joins[t];
    const double Tt = T[t];
    if (!(Tt >= 0.0 && Tt < INFINITY)) atomicAdd((unsigned long long*)&stats[TJ_ST_INVALID], 1ull);
    params[t] = tjmeta::make_params(maxabs[t], Tt, J.d, J.dp, J.prec);
}

// ------------------------------------------- spatial order and tile bounds
// Certified tile pruning.  The rows of each join are grouped by their nearest
// of NPIV pivots (farthest-point sampling on a sample of the join's rows,
// pivots chained nearest-first), so that a panel of consecutive rows is
// spatially compact; every panel gets a bounding ball (centre c, radius
// r >= max |x - c|, rounded up); and panel p sweeps only the panels q whose
// ball bound
//   |x_i - x_j| >= |c_p - c_q| - r_p - r_q   (triangle inequality)
// does not exceed sqrt(T (1 + 2g)) -- pairs of the others are certified
// outside for the binary64 predicate as well (d2_seq >= d2 (1 - g)).  The order is a
// permutation of each join's rows; output keys are mapped back to the
// caller's row numbers before the sort, so results do not depend on it.
// No counterpart in the reference's engine; the reference studied tile
// pruning from projected bounding boxes separately
// (tensorjoin_20260908_block_feasibility/src/executor.py:76-84).
constexpr int NPIV = 8;
constexpr int PIV_BITS = 3;

// perm[k] (k in the batch row space) = the caller's local row at permuted
// position k; stored as the low jbits of the sorted order keys.
__device__ inline int64_t perm_of(const uint64_t* perm, int64_t k, uint64_t jmask) {
    return perm ? (int64_t)(perm[k] & jmask) : -1;
}

// One workgroup per join: farthest-point sampling over m = min(n, 64,
// 8192 / d) sample rows held in LDS (binary32 distances: only the order's
// quality depends on them).  Sample rows sit d + 1 floats apart (with a
// stride of d, the 64 threads of a distance pass hit one LDS bank when d is
// a multiple of 64); the farthest sample is a wave reduction; the pivot
// chain's 8 x 8 distances are formed by 64 threads at once.
__global__ void __launch_bounds__(256) batch_pivot_kernel(const float* const* __restrict__ xs,
                                                          const BatchJoin* __restrict__ joins,
                                                          int32_t* __restrict__ pivots) {
    __shared__ float sx[8192 + 64];  // m (d + 1) <= 8192 + m
    __shared__ float smin[64];
    __shared__ float sdm[NPIV][NPIV];
    __shared__ int spick;
    __shared__ int spiv[NPIV];
    const int t = blockIdx.x;
    const BatchJoin J = joins[t];
    const int tid = threadIdx.x;
    const int n = J.n, d = J.d, ds = d + 1;
    int m = n < 64 ? n : 64;
    if (m > 8192 / d) m = 8192 / d;
    if (m < 1) m = 1;
    if (n == 0) {
        if (tid < NPIV) pivots[t * NPIV + tid] = 0;
        return;
    }
    // (m >= 1 samples; every pivot below is a sample index < m)
    const float* x = xs[t];
    for (int e = tid; e < m * d; e += 256) {
        const int s2 = e / d, k = e - s2 * d;
        const int64_t row = (int64_t)s2 * n / m;
        sx[s2 * ds + k] = x[row * d + k];
    }
    if (tid < 64) smin[tid] = INFINITY;
    if (tid == 0) spick = 0;
    __syncthreads();
    for (int k = 0; k < NPIV; ++k) {
        const int pk = spick;
        if (tid == 0) spiv[k] = pk;
        if (tid < m) {
            float acc = 0.0f;
            for (int c = 0; c < d; ++c) {
                const float u = sx[tid * ds + c] - sx[pk * ds + c];
                acc += u * u;
            }
            smin[tid] = fminf(smin[tid], acc);
        }
        __syncthreads();
        if (tid < 64) {
            // the first sample of largest distance (NaN distances never win;
            // none valid: sample 0)
            const float v = tid < m ? smin[tid] : -1.0f;
            float mx = v;
            for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
            const uint64_t hit = __ballot(v == mx && mx > -1.0f);
            if (tid == 0) spick = hit ? __builtin_ctzll(hit) : 0;
        }
        __syncthreads();
    }
    // Chain the pivots greedily (each next = the nearest one not yet taken),
    // so that pivots of one cluster get adjacent buckets and a cluster's rows
    // stay one contiguous run of the order.
    if (tid < NPIV * NPIV) {
        const int a2 = tid / NPIV, b2 = tid % NPIV;
        float acc = 0.0f;
        for (int c = 0; c < d; ++c) {
            const float u = sx[spiv[a2] * ds + c] - sx[spiv[b2] * ds + c];
            acc += u * u;
        }
        sdm[a2][b2] = acc;
    }
    __syncthreads();
    if (tid == 0) {
        bool used[NPIV] = {};
        int cur = 0;
        used[0] = true;
        pivots[t * NPIV] = (int32_t)((int64_t)spiv[0] * n / m);
        for (int pos = 1; pos < NPIV; ++pos) {
            int best = -1;
            float bv = INFINITY;
            for (int k = 0; k < NPIV; ++k) {
                if (used[k]) continue;
                const float acc = sdm[cur][k];
                if (best < 0 || acc < bv) { bv = acc; best = k; }
            }
            used[best] = true;
            cur = best;
            pivots[t * NPIV + pos] = (int32_t)((int64_t)spiv[best] * n / m);
        }
    }
}

// Order keys, one per row: (join << (PIV_BITS + jbits)) | (nearest pivot <<
// jbits) | local row.  Sorted, position k of the batch row space holds the
// key of the row placed there (the joins stay contiguous and in order).
// The nearest pivot is taken on the first ORDER_DIMS coordinates only (the
// order only steers how much the tile bounds can skip; any order is
// correct), so a row costs one 64-byte read however wide it is.  One wave
// per 64 consecutive rows of a join, one row per lane: the lane holds its
// row's leading coordinates and forms the NPIV distances itself, the pivot
// coordinates being wave-uniform (scalar) loads -- no cross-lane step (16
// lanes per row with a butterfly sum per pivot took 4.6 times the VALU work).
constexpr int ORDER_DIMS = 16;
constexpr int ORDER_ROWS = 64;  // rows per order unit (one wave)

__global__ void __launch_bounds__(256) batch_order_kernel(const float* const* __restrict__ xs,
                                                          const BatchJoin* __restrict__ joins,
                                                          const int64_t* __restrict__ cstart, int nj,
                                                          const int32_t* __restrict__ pivots, int jbits,
                                                          uint64_t* __restrict__ keys) {
    const int64_t c = (int64_t)blockIdx.x * 4 + tjscreen::uni((int)(threadIdx.x >> 6));
    if (c >= cstart[nj]) return;
    const int lane = threadIdx.x & 63;
    const int t = tjscreen::uni(find_join(cstart, nj, c));
    const BatchJoin J = joins[t];
    const int dd = J.d < ORDER_DIMS ? J.d : ORDER_DIMS;
    const float* x = xs[t];
    const int64_t local = (c - cstart[t]) * ORDER_ROWS + lane;
    const bool act = local < J.n;
    const int64_t lc = act ? local : J.n - 1;
    // (the pivot rows below are read through a global-space pointer: their
    // wave-uniform loads then become scalar loads)
    typedef const __attribute__((address_space(1))) float gfloat;
    const float* xr = x + lc * (int64_t)J.d;
    float xv[ORDER_DIMS];
    if ((J.d & 3) == 0 && ((uintptr_t)x & 15) == 0) {
        // dd is a multiple of 4 here (16, or d < 16)
#pragma unroll
        for (int q = 0; q < ORDER_DIMS / 4; ++q) {
            const float4 v = 4 * q < dd ? *(const float4*)(xr + 4 * q) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            xv[4 * q] = v.x; xv[4 * q + 1] = v.y; xv[4 * q + 2] = v.z; xv[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < ORDER_DIMS; ++k) xv[k] = k < dd ? xr[k < dd ? k : 0] : 0.0f;
    }
    const int32_t* piv = pivots + t * NPIV;
    float best = INFINITY;
    int bk = 0;
#pragma unroll
    for (int p = 0; p < NPIV; ++p) {
        const gfloat* pr = (const gfloat*)(x + (int64_t)piv[p] * J.d);
        float acc = 0.0f;
        if (dd == ORDER_DIMS) {
#pragma unroll
            for (int k = 0; k < ORDER_DIMS; ++k) {
                const float u = xv[k] - pr[k];
                acc = fmaf(u, u, acc);
            }
        } else {
            // narrow rows (d < ORDER_DIMS): coordinates past d are 0 in xv
            // and are not read from the pivot
#pragma unroll
            for (int k = 0; k < ORDER_DIMS; ++k) {
                const float u = xv[k] - (k < dd ? pr[k < dd ? k : 0] : 0.0f);
                acc = fmaf(u, u, acc);
            }
        }
        if (acc < best) { best = acc; bk = p; }  // NaN rows stay in bucket 0
    }
    if (act) keys[J.gbase + lc] = ((uint64_t)t << (PIV_BITS + jbits)) | ((uint64_t)bk << jbits) | (uint64_t)lc;
}

// One workgroup per work item (join t, panel p): the panel's bounding ball
// over its rows in the permuted order.  Centre: binary32 mean (any value is
// valid); radius: binary64 sum of squares, rounded up (each difference,
// square and addition loses at most 2^-53 relative: factor 1 + (d+4) 2^-52,
// then the square root and the binary32 rounding upward).
__global__ void __launch_bounds__(256) batch_ball_kernel(const float* const* __restrict__ xs,
                                                         const BatchJoin* __restrict__ joins,
                                                         const int2* __restrict__ items,
                                                         const uint64_t* __restrict__ perm, int jbits,
                                                         float* __restrict__ balls) {
    __shared__ float spart[4][256];
    __shared__ __attribute__((aligned(16))) float sc[256];
    __shared__ double sr[4];
    const int2 it = items[blockIdx.x];
    const BatchJoin J = joins[it.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t jmask = (1ull << jbits) - 1;
    const int64_t r0 = (int64_t)it.y * J.bm;
    const int64_t r1 = r0 + J.bm < J.n ? r0 + J.bm : J.n;
    const float* x = xs[it.x];
    const int d = J.d;  // <= 256: at most four coordinates per lane
    // centre: the mean of up to 32 rows spread over the panel (any centre is
    // valid; the radius below is taken over every row), 8 per wave.  The
    // rows' order positions are read first, then their coordinates, so a
    // wave has its loads in flight together (a loop of dependent pairs of
    // loads, one row after another, was most of this kernel's time).
    const int64_t nr = r1 - r0;
    const int ns = nr < 32 ? (int)nr : 32;
    const int wv = tjscreen::uni(wave);
    int64_t ro[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int sidx = wv + 4 * u < ns ? wv + 4 * u : ns - 1;
        ro[u] = perm_of(perm, J.gbase + r0 + (int64_t)sidx * nr / ns, jmask) * d;
    }
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (64 * m >= d) break;
        // lanes past d read coordinate d - 1 and store nothing below
        const int k = lane + 64 * m < d ? lane + 64 * m : d - 1;
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = x[ro[u] + k];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (wv + 4 * u < ns) acc[m] += v[u];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
        if (lane + 64 * m < d) spart[wave][lane + 64 * m] = acc[m];
    __syncthreads();
    if (tid < d) sc[tid] = ((spart[0][tid] + spart[1][tid]) + (spart[2][tid] + spart[3][tid])) / (float)ns;
    __syncthreads();
    // radius: binary32 sums, LPR lanes per row (64 / LPR rows per wave at a
    // time).  Every difference and square of binary32 values is rounded once
    // and all terms are non-negative, so the exact |x - c|^2 <= (S + (d+2)
    // 2^-125) / (1 - (d+2) 2^-24) whatever the summation order.  A NaN or
    // infinite sum (centre overflow) gives an infinite radius: no pruning.
    // 16-byte loads when every row starts on a 16-byte boundary; then LPR is
    // the row's 16-byte count rounded up to a power of two (at most 16), so
    // narrow rows keep every lane busy (d = 16: 16 rows per wave, not 4)
    const bool vec = (d & 3) == 0 && ((uintptr_t)x & 15) == 0;
    int lpr = 16;
    if (vec) {
        lpr = 1;
        while (lpr < (d >> 2) && lpr < 16) lpr <<= 1;
    }
    const int lg = __builtin_ctz(lpr);
    const int g16 = lane >> lg, l16 = lane & (lpr - 1);
    const int rpw = 64 >> lg;
    float rmax = 0.0f;
    bool bad = false;
    // (the next step's order positions are read one step ahead)
    int64_t rb = r0 + wv * rpw;
    int64_t pnext = 0;
    if (rb < r1) pnext = perm_of(perm, J.gbase + (rb + g16 < r1 ? rb + g16 : r0), jmask);
    for (; rb < r1; rb += 4 * rpw) {
        const int64_t r = rb + g16;
        const bool act = r < r1;
        const int64_t pcur = pnext;
        const int64_t rn = rb + 4 * rpw + g16;
        if (rb + 4 * rpw < r1) pnext = perm_of(perm, J.gbase + (rn < r1 ? rn : r0), jmask);
        const float* xr = x + pcur * d;
        float a = 0.0f;
        if (vec) {
            for (int k = 4 * l16; k < d; k += 4 * lpr) {
                const float4 v = *(const float4*)(xr + k);
                const float4 cc = *(const float4*)(sc + k);
                const float u0 = v.x - cc.x, u1 = v.y - cc.y, u2 = v.z - cc.z, u3 = v.w - cc.w;
                a += u0 * u0;
                a += u1 * u1;
                a += u2 * u2;
                a += u3 * u3;
            }
        } else {
            for (int k = l16; k < d; k += 16) {
                const float u = xr[k] - sc[k];
                a += u * u;
            }
        }
        for (int o = lpr >> 1; o > 0; o >>= 1)
            a += __builtin_bit_cast(float, tj::xor_lane(__builtin_bit_cast(uint32_t, a), o, lane));
        if (act) {
            bad |= !(a < INFINITY);
            rmax = fmaxf(rmax, a);
        }
    }
    bad = __any(bad);
    for (int o = 32; o > 0; o >>= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, o));
    if (lane == 0) sr[wave] = bad ? (double)INFINITY : (double)rmax;
    __syncthreads();
    float* ball = balls + J.ball_off + (int64_t)it.y * (d + 1);
    if (tid < d) ball[tid] = sc[tid];
    if (tid == 0) {
        const double S = fmax(fmax(sr[0], sr[1]), fmax(sr[2], sr[3]));  // no NaN here
        const double n = (double)d + 2.0;
        double r2 = (S + n * 0x1p-125) / (1.0 - n * 0x1p-24);
        r2 += r2 * 0x1p-50;
        ball[d] = f32_up(sqrt(r2) * (1.0 + 0x1p-50));
    }
}

// One wave per work item (join t, panel p): the column ranges the panel
// still has to sweep -- runs of consecutive panels q >= p whose ball bound
// is not certified outside, at most MAXR runs (past that, a new run is
// merged into the last one: screening more is always sound) -- and the
// pairs it skips (TJ_ST_PRUNED, TJ_ST_PRUNED_WORK).  Ranges are local
// columns at panel boundaries (multiples of 32), the last one may end at n.
constexpr int MAXR = 4;
constexpr int RSTRIDE = 1 + 2 * MAXR;  // int32 per item: count, then (begin, end) pairs

__global__ void __launch_bounds__(256) batch_prune_kernel(const BatchJoin* __restrict__ joins,
                                                          const int2* __restrict__ items, int64_t n_items,
                                                          const JoinParams* __restrict__ params,
                                                          const float* __restrict__ balls,
                                                          int32_t* __restrict__ ranges, uint64_t* __restrict__ stats) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the skipped pairs (and their work) of this wave's items, added to the
    // stats by one atomic pair per workgroup: an atomic per item on the
    // same two counters serialised and cost as much as the pruning
    uint64_t skip_pairs = 0, skip_work = 0;
    for (int64_t w = (int64_t)blockIdx.x * 4 + tjscreen::uni(wave); w < n_items; w += (int64_t)gridDim.x * 4) {
        const int2 it = items[w];
        const BatchJoin J = joins[it.x];
        const JoinParams& P = params[it.x];
        const int d = J.d;
        const int panels = (J.n + J.bm - 1) / J.bm;
        const float* bp = balls + J.ball_off + (int64_t)it.y * (d + 1);
        const double rp = (double)bp[d];
        const double lim = P.T * (1.0 + 2.0 * P.g64) * (1.0 + 0x1p-50);
        int nr = 0;
        int rb[MAXR], re[MAXR];  // runs in panel units [rb, re)
        // 64 / GW panels q at a time, a group of GW lanes per panel (GW grows with
        // d so that a lane sums at most 16 of the squared differences, in
        // binary64: the bound below holds for any summation order), then the
        // kept panels in order from the wave's mask
        const int GW = d <= 16 ? 1 : d <= 64 ? 4 : d <= 256 ? 16 : d <= 512 ? 32 : 64;
        const int QPS = 64 / GW;
        const int sub = lane % GW;
        for (int q0 = it.y; q0 < panels; q0 += QPS) {
            const int q = q0 + lane / GW;
            bool keep = q == it.y;
            double a = 0.0;
            const float* bq = balls + J.ball_off + (int64_t)(q < panels ? q : it.y) * (d + 1);
            for (int k = sub; k < d; k += GW) {
                const double u = (double)bp[k] - (double)bq[k];
                a += u * u;
            }
            for (int o = GW >> 1; o > 0; o >>= 1) a += __shfl_xor(a, o);
            if (q < panels && !keep) {
                // |c_p - c_q| from below, then the gap from below
                const double dist = sqrt(a * (1.0 - (double)(d + 4) * 0x1p-52)) * (1.0 - 0x1p-50);
                const double rs = rp + (double)bq[d];
                double gap = dist - rs;
                gap -= (dist + rs) * 0x1p-50;
                keep = !(gap > 0.0 && gap * gap * (1.0 - 0x1p-50) > lim);
            }
            const uint64_t kb = __ballot(q < panels && keep && sub == 0);  // one bit per group: lane GW g
            for (int g = 0; g < QPS; ++g) {
                if (((kb >> (g * GW)) & 1ull) == 0) continue;
                const int qk = q0 + g;
                if (nr > 0 && re[nr - 1] == qk) re[nr - 1] = qk + 1;   // extends the current run
                else if (nr < MAXR) { rb[nr] = qk; re[nr] = qk + 1; ++nr; }
                else re[nr - 1] = qk + 1;                              // merged into the last run
            }
        }
        if (lane == 0) {
            int32_t* out = ranges + w * RSTRIDE;
            out[0] = nr;
            int64_t swept = 0;
            for (int r = 0; r < nr; ++r) {
                const int64_t c0 = (int64_t)rb[r] * J.bm;
                const int64_t c1 = (int64_t)re[r] * J.bm < J.n ? (int64_t)re[r] * J.bm : J.n;
                out[1 + 2 * r] = (int32_t)c0;
                out[2 + 2 * r] = (int32_t)c1;
                swept += c1 - c0;
            }
            const int64_t r0 = (int64_t)it.y * J.bm;
            const int64_t rows = r0 + J.bm < J.n ? (int64_t)J.bm : J.n - r0;
            const int64_t skipped = rows * (J.n - r0 - swept);
            if (skipped > 0) {
                skip_pairs += (uint64_t)skipped;
                skip_work += (uint64_t)skipped * (uint64_t)d;
            }
        }
    }
    __shared__ uint64_t sp[4], sw[4];
    if (lane == 0) { sp[wave] = skip_pairs; sw[wave] = skip_work; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t a = sp[0] + sp[1] + sp[2] + sp[3], b = sw[0] + sw[1] + sw[2] + sw[3];
        if (a) {
            atomicAdd((unsigned long long*)&stats[TJ_ST_PRUNED], (unsigned long long)a);
            atomicAdd((unsigned long long*)&stats[TJ_ST_PRUNED_WORK], (unsigned long long)b);
        }
    }
}

// ------------------------------------------------------ row-grouped output
// The output keys (batch row of i, local j) are ordered without a global
// sort, in three steps:
//  1. by join: a count pass adds each key to its join's count, an exclusive
//     scan over the joins gives each join its output range, a scatter pass
//     moves each key, as its local (i, j) (with tile pruning mapped back
//     from the permuted rows to the caller's, i < j), into its join's range.
//     A flush of the screen stages the survivors of one work item together,
//     so the keys of a join arrive in long runs and a wave adds one count
//     per run of equal joins among its lanes;
//  2. by row within a join, one workgroup per join: row counts (LDS
//     counters for joins of up to ROWS_LDS rows, global ones beyond),
//     their scan (the row offsets), and a scatter of each j into its row's
//     range;
//  3. the per-row sort (rowsort.hip) orders every row and emits the local
//     (i, j) rows.
__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

constexpr int ROWS_LDS = 16384;  // rows of a join whose counters live in LDS (64 KiB)
constexpr int JR_THREADS = 1024;

template <bool SCATTER>
__global__ void __launch_bounds__(256) batch_bucket_kernel(const uint64_t* __restrict__ keys,
                                                           const uint64_t* __restrict__ count, int64_t cap, int jbits,
                                                           const int32_t* __restrict__ row_join,
                                                           const BatchJoin* __restrict__ joins,
                                                           const uint64_t* __restrict__ perm,
                                                           uint32_t* __restrict__ jcnt, const uint64_t* __restrict__ joff,
                                                           uint64_t* __restrict__ bucket) {
    int64_t cnt = (int64_t)*count;
    if (cnt > cap) cnt = cap;
    const uint64_t jmask = (1ull << jbits) - 1;
    const int lane = threadIdx.x & 63;
    // UN groups of 64 consecutive keys per wave step, their loads and table
    // lookups issued before any is used (overlapping latencies)
    constexpr int UN = 4;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * UN;
    // k0 is wave-uniform: every lane of a wave runs the same iterations
    for (int64_t k0 = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) * UN; k0 < cnt; k0 += stride) {
        int t[UN];
        uint64_t local[UN];
        uint64_t key[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const int64_t k = k0 + 64 * u + lane;
            key[u] = k < cnt ? keys[k] : ~0ull;
        }
#pragma unroll
        for (int u = 0; u < UN; ++u) t[u] = key[u] != ~0ull ? row_join[(int64_t)(key[u] >> jbits)] : -1;
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            local[u] = 0;
            if (SCATTER && t[u] >= 0) {
                const int64_t ig = (int64_t)(key[u] >> jbits);
                int64_t jl = (int64_t)(key[u] & jmask);
                const int64_t gb = joins[t[u]].gbase;
                int64_t il = ig - gb;
                if (perm) {
                    il = perm_of(perm, ig, jmask);
                    jl = perm_of(perm, gb + jl, jmask);
                    if (il > jl) { const int64_t w = il; il = jl; jl = w; }
                }
                local[u] = ((uint64_t)il << 32) | (uint64_t)jl;
            }
        }
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            // runs of equal joins among the lanes (the valid lanes are a prefix)
            const int tu = t[u];
            const int prev = __shfl(tu, lane > 0 ? lane - 1 : 0);
            const bool head = tu >= 0 && (lane == 0 || prev != tu);
            const uint64_t H = __ballot(head);
            const int nvalid = __popcll(__ballot(tu >= 0));
            const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);  // lanes <= this one
            const uint64_t hb = H & upto, ha = H & ~upto;
            const int h = hb ? 63 - __clzll(hb) : 0;          // this lane's run head
            const int nxt = ha ? __builtin_ctzll(ha) : nvalid;  // the next run's head
            uint32_t base = 0;
            if (head) {
                const uint32_t len = (uint32_t)(nxt - lane);
                if (SCATTER) base = atomicSub(&jcnt[tu], len) - len;  // the run takes [c - len, c) of the join
                else atomicAdd(&jcnt[tu], len);
            }
            if (SCATTER) {
                base = (uint32_t)__shfl((int)base, h);
                if (tu >= 0) bucket[joff[tu] + base + (uint32_t)(lane - h)] = local[u];
            }
        }
    }
}

// Runs of equal values among a wave's lanes: whether this lane starts one,
// the run's first lane h and (at the head) its length.
struct RowRun {
    bool head;
    int h;
    uint32_t len;
};
__device__ inline RowRun row_run(uint32_t v, int lane) {
    const uint32_t prev = (uint32_t)__shfl((int)v, lane > 0 ? lane - 1 : 0);
    const bool head = lane == 0 || prev != v;
    const uint64_t H = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);  // lanes <= this one
    const uint64_t hb = H & upto, ha = H & ~upto;
    RowRun r;
    r.head = head;
    r.h = 63 - __clzll(hb);                       // lane 0 is always a head
    const int nxt = ha ? __builtin_ctzll(ha) : 64;  // the next run's head
    r.len = (uint32_t)(nxt - lane);
    return r;
}

// Step 2 for the joins listed in order[0, njl): LDS counters (joins of at
// most ROWS_LDS rows) or global ones (gcnt, one per batch row).  roff[r] =
// first output position of batch row r; the first workgroup also writes
// roff[rows], the total.
template <bool LDS>
__global__ void __launch_bounds__(JR_THREADS) batch_join_rows_kernel(const int32_t* __restrict__ order, int njl,
                                                              const BatchJoin* __restrict__ joins, int nj, int64_t rows,
                                                              const uint64_t* __restrict__ joff,
                                                              const uint64_t* __restrict__ bucket,
                                                              uint32_t* __restrict__ gcnt, uint64_t* __restrict__ roff,
                                                              uint32_t* __restrict__ outj) {
    __shared__ uint32_t sc[LDS ? ROWS_LDS : 1];
    __shared__ uint32_t wsum[JR_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (LDS && blockIdx.x == 0 && tid == 0) roff[rows] = joff[nj];
    for (int q = blockIdx.x; q < njl; q += gridDim.x) {
        const int t = order[q];
        const int64_t gb = joins[t].gbase;
        const int n = joins[t].n;
        uint32_t* c = LDS ? sc : gcnt + gb;
        for (int a = tid; a < n; a += JR_THREADS) c[a] = 0;
        __syncthreads();
        const uint64_t k0 = joff[t], k1 = joff[t + 1];
        // UN keys per thread per step, loaded before any is used (their
        // latencies overlap)
        constexpr int UN = 8;
        // (k - lane is wave-uniform: every lane of a wave runs the same
        // iterations, as the run detection below needs)
        for (uint64_t k = k0 + tid; k - (uint64_t)lane < k1; k += (uint64_t)JR_THREADS * UN) {
            uint32_t rows_u[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const uint64_t kk = k + (uint64_t)JR_THREADS * u;
                rows_u[u] = kk < k1 ? (uint32_t)(bucket[kk] >> 32) : 0xFFFFFFFFu;
            }
            // one LDS atomic per run of equal rows among a wave's lanes (the
            // survivors of one register share their row: same-address
            // atomics of a wave would serialise)
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const uint32_t rw = rows_u[u];
                const RowRun rr = row_run(rw, lane);
                if (rr.head && rw != 0xFFFFFFFFu) atomicAdd(&c[rw], rr.len);
            }
        }
        __syncthreads();
        // exclusive scan of the row counts, JR_THREADS rows at a time: each
        // row's first position in the join's range
        uint32_t run = 0;
        for (int base = 0; base < n; base += JR_THREADS) {
            const int a = base + tid;
            const uint32_t v = a < n ? c[a] : 0;
            uint32_t x = v;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up((int)x, o);
                if (lane >= o) x += y;
            }
            if (lane == 63) wsum[wave] = x;
            __syncthreads();
            uint32_t before = run;
            for (int w = 0; w < wave; ++w) before += wsum[w];
            if (a < n) {
                c[a] = before + x - v;
                roff[gb + a] = k0 + before + x - v;
            }
            for (int w = 0; w < JR_THREADS / 64; ++w) run += wsum[w];
            __syncthreads();
        }
        for (uint64_t k = k0 + tid; k - (uint64_t)lane < k1; k += (uint64_t)JR_THREADS * UN) {
            uint64_t v[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const uint64_t kk = k + (uint64_t)JR_THREADS * u;
                v[u] = kk < k1 ? bucket[kk] : ~0ull;
            }
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const uint32_t rw = v[u] == ~0ull ? 0xFFFFFFFFu : (uint32_t)(v[u] >> 32);
                const RowRun rr = row_run(rw, lane);
                uint32_t pos = 0;
                if (rr.head && rw != 0xFFFFFFFFu) pos = atomicAdd(&c[rw], rr.len);
                pos = (uint32_t)__shfl((int)pos, rr.h) + (uint32_t)(lane - rr.h);
                if (rw != 0xFFFFFFFFu) outj[k0 + pos] = (uint32_t)v[u];
            }
        }
        __syncthreads();  // the counters are reused by the next join
    }
}

// ----------------------------------------------------------- row metadata
// One unit (64 / GW rows of one join) per wave, written into its group's
// arrays.
__global__ void __launch_bounds__(256) batch_rowmeta_kernel(const float* const* __restrict__ xs,
                                                            const BatchJoin* __restrict__ joins,
                                                            const BatchGroup* __restrict__ groups,
                                                            const int64_t* __restrict__ cstart, int nj,
                                                            const JoinParams* __restrict__ params, uint8_t* ws,
                                                            const uint64_t* __restrict__ perm, int jbits) {
    const int64_t c = (int64_t)blockIdx.x * 4 + tjscreen::uni((int)(threadIdx.x >> 6));
    if (c >= cstart[nj]) return;
    const int lane = threadIdx.x & 63;
    const int t = find_join(cstart, nj, c);
    const BatchJoin J = joins[t];
    const BatchGroup G = groups[J.group];
    const JoinParams P = params[t];
    uint8_t* q = ws + G.q_off;
    RowMeta* meta = (RowMeta*)(ws + G.meta_off);
    const int gw = tjmeta::meta_width(J.dp);
    const int64_t r0 = (c - cstart[t]) * meta_unit_rows(J.dp);
    const float* x = xs[t];
    const uint64_t jm = (1ull << jbits) - 1;
    const int64_t gb = J.gbase;
    const int d = J.d;
    // the caller's row behind position r (rows are read through the spatial
    // order when there is one)
    auto src = [=](int64_t r) { return x + (perm ? perm_of(perm, gb + r, jm) : r) * (int64_t)d; };
    using S = decltype(src);
    if (J.prec == P_FP8) {
        if (gw == 16) tjmeta::rowmeta_wave<P_FP8, 16, S, 64>(src, J.n, r0, d, J.dp, P, lane, q, q, meta, J.grow);
        else if (gw == 32) tjmeta::rowmeta_wave<P_FP8, 32, S, 32>(src, J.n, r0, d, J.dp, P, lane, q, q, meta, J.grow);
        else tjmeta::rowmeta_wave<P_FP8, 64, S, 16>(src, J.n, r0, d, J.dp, P, lane, q, q, meta, J.grow);
    } else if (gw == 16) {
        tjmeta::rowmeta_wave<P_FP16, 16, S, 64>(src, J.n, r0, d, J.dp, P, lane, q, q, meta, J.grow);
    } else if (gw == 32) {
        tjmeta::rowmeta_wave<P_FP16, 32, S, 32>(src, J.n, r0, d, J.dp, P, lane, q, q, meta, J.grow);
    } else {
        tjmeta::rowmeta_wave<P_FP16, 64, S, 16>(src, J.n, r0, d, J.dp, P, lane, q, q, meta, J.grow);
    }
}

// ------------------------------------------------------------------ screen
struct ScreenBatchArgs {
    const int2* items;  // (join, panel), largest first
    const int32_t* ranges;  // per item: column ranges left after tile pruning (NULL: [panel start, n))
    int64_t n_items;
    const BatchJoin* joins;
    const BatchGroup* groups;
    const uint8_t* ws;
    tjscreen::CandOut co;
    uint64_t* cursor;
    int jbits;
    int fast_accept;
};

// One persistent launch over the panels of every join of one precision
// family.  The body of an item is the single-join screen's panel sweep,
// instantiated per (precision, width) and selected by the item's group.
template <int FAM>
__global__ void __launch_bounds__(256, 2) batch_screen_kernel(const ScreenBatchArgs A) {
    using namespace tjscreen;
    __shared__ __attribute__((aligned(16))) uint8_t sB[RING_MAX];
    __shared__ __attribute__((aligned(16))) float sNp[NSTAGE_MAX * 64];
    __shared__ __attribute__((aligned(16))) float sW[NSTAGE_MAX * 64];
    __shared__ __attribute__((aligned(16))) uint32_t sSc[NSTAGE_MAX * 64];
    __shared__ uint64_t sCK[4][CCAP_BATCH];
    __shared__ float sCP[4][CCAP_BATCH];
    __shared__ __attribute__((aligned(16))) float sT[4][4];  // row levels of unseeded bodies (none here)
    __shared__ int64_t sItem;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    int nC = 0;
    uint32_t nAcc = 0;  // fast accepts of this wave (wave-uniform)
    for (;;) {
        if (tid == 0) sItem = (int64_t)atomicAdd((unsigned long long*)A.cursor, 1ull);
        __syncthreads();
        const int64_t item = (int64_t)uni64((uint64_t)sItem);  // uniform: scalar loads of the tables below
        __syncthreads();
        if (item >= A.n_items) break;
        const int2 it = A.items[item];
        const BatchJoin J = A.joins[it.x];
        const BatchGroup G = A.groups[J.group];
        PanelIO io;
        const uint8_t* q = A.ws + G.q_off;
        const RowMeta* meta = (const RowMeta*)(A.ws + G.meta_off);
        io.qa[0] = io.qa[1] = io.qb[0] = io.qb[1] = q;
        io.ma = io.mb = meta;
        io.na = io.nb = J.grow + J.n;
        io.koff_a = J.gbase - J.grow;  // i -> batch row
        io.koff_b = -J.grow;           // j -> local row
        io.jbits = A.jbits;
        const int64_t r0 = J.grow + (int64_t)it.y * J.bm;
        const int32_t* rg = A.ranges ? A.ranges + item * RSTRIDE : nullptr;
        const int nr = rg ? uni(rg[0]) : 1;
        for (int r = 0; r < nr; ++r) {
        const int64_t cb = rg ? J.grow + (int64_t)uni(rg[1 + 2 * r]) : r0;
        const int64_t ce = rg ? J.grow + (int64_t)uni(rg[2 + 2 * r]) : J.grow + (int64_t)J.n;
#define X(ID, PR, DPV)                                                                                          \
    case ID:                                                                                                    \
        screen_panel<PR, DPV, rb_batch<PR, DPV>(), true, CCAP_BATCH, true>(io, A.co, r0, cb, ce, sB, sNp, sW, sSc, sCK[wave], \
                                                       sCP[wave], sT[wave], nC, nAcc, A.fast_accept);                     \
        break;
        if constexpr (FAM == 0) {
            switch (uni(G.combo)) {
                TJ_BATCH_F16A(X)
                default: break;
            }
        } else if constexpr (FAM == 1) {
            switch (uni(G.combo)) {
                TJ_BATCH_F16B(X)
                default: break;
            }
        } else if constexpr (FAM == 2) {
            switch (uni(G.combo)) {
                TJ_BATCH_F8A(X)
                default: break;
            }
        } else {
            switch (uni(G.combo)) {
                TJ_BATCH_F8B(X)
                default: break;
            }
        }
#undef X
        }
    }
    cand_finish(A.co, sCK[wave], sCP[wave], nC, nAcc, lane);
}

// ----------------------------------------------------------- certificate
struct BatchSrc {
    const float* const* xs;
    const BatchJoin* joins;
    const BatchGroup* groups;
    const int32_t* row_join;
    const JoinParams* params;
    const uint8_t* ws;
    int jbits;
    const uint64_t* perm;  // permuted position -> caller's row (NULL: identity)
    __device__ inline void fetch(uint64_t key, tjcert::PairRef& r) const {
        const uint64_t jmask = (1ull << jbits) - 1;
        const int64_t ig = (int64_t)(key >> jbits);
        const int64_t jl = (int64_t)(key & jmask);
        const int t = row_join[ig];
        const BatchJoin J = joins[t];
        const int64_t il = ig - J.gbase;
        const RowMeta* meta = (const RowMeta*)(ws + groups[J.group].meta_off);
        r.P = params + t;
        r.mi = meta[J.grow + il];
        r.mj = meta[J.grow + jl];
        const int64_t si = perm ? perm_of(perm, ig, jmask) : il;
        const int64_t sj = perm ? perm_of(perm, J.gbase + jl, jmask) : jl;
        r.xi = xs[t] + si * (int64_t)J.d;
        r.xj = xs[t] + sj * (int64_t)J.d;
        r.d = J.d;
    }
};

}  // namespace tjbatch

using namespace tj;
using namespace tjbatch;

namespace tj {
int screen_batch_occupancy(int fam) {
    int blocks = 0;
    hipError_t e;
    switch (fam) {
        case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batch_screen_kernel<0>, 256, 0); break;
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batch_screen_kernel<1>, 256, 0); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batch_screen_kernel<2>, 256, 0); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, batch_screen_kernel<3>, 256, 0); break;
    }
    if (e != hipSuccess) return 1;
    return blocks > 0 ? blocks : 1;
}
}  // namespace tj

struct tj_batch {
    int nj;
    int64_t rows, max_n, n_items, pairs_total;
    int ibits, jbits, keybits;
    int64_t out_cap, cand_cap;
    unsigned flags;
    int n_groups;
    int64_t fam_begin[N_FAMILIES + 1];  // items of family f: [fam_begin[f], fam_begin[f + 1])
    int fam_grid[N_FAMILIES];
    // device tables (owned)
    void* dev;
    BatchJoin* d_joins;
    BatchGroup* d_groups;
    int2* d_items;
    int32_t* d_row_join;
    int32_t* d_row_local;  // row -> its number within its join
    int64_t* d_cmeta;  // row-metadata units before join t
    int64_t* d_cscan;  // scale-scan units before join t
    int64_t* d_corder;  // spatial-order units (ORDER_ROWS rows) before join t
    int64_t units_meta, units_scan, units_order;
    // tile pruning (off unless prune)
    bool prune;
    int order_bits;
    int64_t ball_floats;
    // workspace layout
    size_t off_maxabs, off_cursor, off_params, off_cand, off_aux, off_out, off_sort, total;
    size_t off_pivots, off_perm, off_balls, off_cend, off_rows;
    size_t off_jcnt, off_jpart, off_bucket, off_rcnt, off_roff, off_long;  // row-grouped output
    int32_t* d_jorder;  // joins of at most ROWS_LDS rows, then the others, each largest first
    int64_t n_small;
    hipEvent_t ev[5];
    bool timing;
    hipGraphExec_t graph;
    bool has_graph;
};

static size_t bcarve(size_t& cur, size_t bytes) {
    size_t at = (cur + 255) / 256 * 256;
    cur = at + bytes;
    return at;
}

namespace tj {
int capture_graph(hipGraphExec_t* exec, int (*fn)(void*, hipStream_t), void* ctx);
}

struct BatchCall {
    tj_batch* b;
    const float* const* x;
    const double* T;
    void* ws;
    int32_t* pairs;
    int64_t* offsets;
    uint64_t* stats;
    bool marks;
};

static int batch_enqueue(const BatchCall& c, hipStream_t s) {
    tj_batch* b = c.b;
    uint8_t* ws = (uint8_t*)c.ws;
    uint32_t* maxabs = (uint32_t*)(ws + b->off_maxabs);
    uint64_t* cursors = (uint64_t*)(ws + b->off_cursor);
    JoinParams* params = (JoinParams*)(ws + b->off_params);
    uint64_t* cand = (uint64_t*)(ws + b->off_cand);
    float* aux = (float*)(ws + b->off_aux);
    uint64_t* out_keys = (uint64_t*)(ws + b->off_out);
    uint64_t* stats = c.stats;
    int rc;
#define TJ_MARK(k) \
    if (c.marks && b->timing) TJ_HIP(hipEventRecord(b->ev[k], s))
    if ((rc = launch_fill_u64(stats, TJ_STATS_SLOTS, 0, s))) return rc;
    if ((rc = launch_fill_u64(stats + TJ_ST_PAIRS, 1, (uint64_t)b->pairs_total, s))) return rc;
    if ((rc = launch_fill_u64((uint64_t*)maxabs, (b->nj + 1) / 2, 0, s))) return rc;
    if ((rc = launch_fill_u64(cursors, N_FAMILIES, 0, s))) return rc;
    const uint64_t* perm = nullptr;
    int32_t* ranges = nullptr;
    if (b->prune) {
        perm = (const uint64_t*)(ws + b->off_perm);
        ranges = (int32_t*)(ws + b->off_cend);
        if ((rc = launch_fill_u64((uint64_t*)(ws + b->off_rows), 1, (uint64_t)b->rows, s))) return rc;
    }
    TJ_MARK(0);
    if (b->units_scan > 0) {
        int64_t blocks = (b->units_scan + 3) / 4;  // one unit per wave, 4 waves per block
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(batch_scan_kernel, dim3((unsigned)blocks), dim3(256), 0, s, c.x, b->d_joins, b->d_cscan,
                           b->nj, maxabs, stats);
        TJ_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(batch_params_kernel, dim3((unsigned)((b->nj + 255) / 256)), dim3(256), 0, s, b->d_joins, b->nj,
                       c.T, maxabs, params, stats);
    TJ_HIP(hipGetLastError());
    if (b->prune && b->units_scan > 0) {
        // spatial order of each join's rows (sorted order keys = the permutation)
        int32_t* pivots = (int32_t*)(ws + b->off_pivots);
        hipLaunchKernelGGL(batch_pivot_kernel, dim3((unsigned)b->nj), dim3(256), 0, s, c.x, b->d_joins, pivots);
        TJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(batch_order_kernel, dim3((unsigned)((b->units_order + 3) / 4)), dim3(256), 0, s, c.x,
                           b->d_joins, b->d_corder, b->nj, pivots, b->jbits, (uint64_t*)perm);
        TJ_HIP(hipGetLastError());
        if ((rc = launch_sort(ws + b->off_sort, (uint64_t*)perm, (const uint64_t*)(ws + b->off_rows), b->rows,
                              b->order_bits, s)))
            return rc;
    }
    if (b->units_meta > 0) {
        hipLaunchKernelGGL(batch_rowmeta_kernel, dim3((unsigned)((b->units_meta + 3) / 4)), dim3(256), 0, s, c.x,
                           b->d_joins, b->d_groups, b->d_cmeta, b->nj, params, ws, perm, b->jbits);
        TJ_HIP(hipGetLastError());
    }
    if (b->prune && b->n_items > 0) {
        float* balls = (float*)(ws + b->off_balls);
        hipLaunchKernelGGL(batch_ball_kernel, dim3((unsigned)b->n_items), dim3(256), 0, s, c.x, b->d_joins,
                           b->d_items, perm, b->jbits, balls);
        TJ_HIP(hipGetLastError());
        // grid-stride over the items, a few workgroups per CU (one stats
        // atomic pair each)
        const int64_t pgrid = std::max<int64_t>(1, std::min<int64_t>((b->n_items + 3) / 4, (int64_t)num_cus() * 8));
        hipLaunchKernelGGL(batch_prune_kernel, dim3((unsigned)pgrid), dim3(256), 0, s, b->d_joins, b->d_items,
                           b->n_items, params, balls, ranges, stats);
        TJ_HIP(hipGetLastError());
    }
    TJ_MARK(1);
    for (int f = 0; f < N_FAMILIES; ++f) {
        const int64_t n_items = b->fam_begin[f + 1] - b->fam_begin[f];
        if (n_items == 0) continue;
        ScreenBatchArgs a{};
        a.items = b->d_items + b->fam_begin[f];
        a.ranges = ranges ? ranges + b->fam_begin[f] * RSTRIDE : nullptr;
        a.n_items = n_items;
        a.joins = b->d_joins;
        a.groups = b->d_groups;
        a.ws = ws;
        a.co = tjscreen::CandOut{cand, aux, b->cand_cap, &stats[TJ_ST_CANDIDATES],
                                 out_keys, b->out_cap, &stats[TJ_ST_OUT], &stats[TJ_ST_CERT_ACCEPT]};
        a.cursor = cursors + f;
        a.jbits = b->jbits;
        a.fast_accept = (b->flags & TJ_FLAG_AUDIT) ? 0 : 1;
        const dim3 grid((unsigned)b->fam_grid[f]);
        switch (f) {
            case 0: hipLaunchKernelGGL(batch_screen_kernel<0>, grid, dim3(256), 0, s, a); break;
            case 1: hipLaunchKernelGGL(batch_screen_kernel<1>, grid, dim3(256), 0, s, a); break;
            case 2: hipLaunchKernelGGL(batch_screen_kernel<2>, grid, dim3(256), 0, s, a); break;
            default: hipLaunchKernelGGL(batch_screen_kernel<3>, grid, dim3(256), 0, s, a); break;
        }
        TJ_HIP(hipGetLastError());
    }
    TJ_MARK(2);
    if (b->cand_cap > 0 && b->n_items > 0) {
        int64_t blocks = (b->cand_cap + 255) / 256;
        const int64_t maxb = (int64_t)num_cus() * 8;
        if (blocks > maxb) blocks = maxb;
        const BatchSrc src{c.x, b->d_joins, b->d_groups, b->d_row_join, params, ws, b->jbits, perm};
        hipLaunchKernelGGL(tjcert::certify_kernel<BatchSrc>, dim3((unsigned)blocks), dim3(256), 0, s, src, cand, aux,
                           b->cand_cap, out_keys, b->out_cap, stats, (b->flags & TJ_FLAG_AUDIT) ? 1 : 0);
        TJ_HIP(hipGetLastError());
    }
    TJ_MARK(3);
    {
        // output: keys by join (into the bucket), rows per join, per-row
        // sort + emit; the join offsets are the caller's offsets array
        uint32_t* jcnt = (uint32_t*)(ws + b->off_jcnt);
        uint64_t* joff = (uint64_t*)c.offsets;
        uint64_t* bucket = (uint64_t*)(ws + b->off_bucket);
        uint64_t* roff = (uint64_t*)(ws + b->off_roff);
        uint32_t* outj = (uint32_t*)out_keys;  // the keys are in the bucket by then
        uint32_t* outj2 = outj + b->out_cap;
        if ((rc = launch_fill_u64((uint64_t*)jcnt, (b->nj + 1) / 2, 0, s))) return rc;
        int64_t blocks = (b->out_cap + 255) / 256;
        if (blocks > (int64_t)num_cus() * 16) blocks = (int64_t)num_cus() * 16;
        if (blocks < 1) blocks = 1;
        hipLaunchKernelGGL(batch_bucket_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, s, out_keys,
                           stats + TJ_ST_OUT, b->out_cap, b->jbits, b->d_row_join, b->d_joins, perm, jcnt,
                           (const uint64_t*)nullptr, (uint64_t*)nullptr);
        TJ_HIP(hipGetLastError());
        if ((rc = launch_row_offsets(jcnt, b->nj, (uint64_t*)(ws + b->off_jpart), joff, nullptr, s))) return rc;
        hipLaunchKernelGGL(batch_bucket_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, s, out_keys,
                           stats + TJ_ST_OUT, b->out_cap, b->jbits, b->d_row_join, b->d_joins, perm, jcnt, joff,
                           bucket);
        TJ_HIP(hipGetLastError());
        const int g0 = (int)std::max<int64_t>(1, std::min<int64_t>(b->n_small, 2 * (int64_t)num_cus()));
        hipLaunchKernelGGL(batch_join_rows_kernel<true>, dim3((unsigned)g0), dim3(JR_THREADS), 0, s, b->d_jorder,
                           (int)b->n_small, b->d_joins, b->nj, b->rows, joff, bucket, (uint32_t*)(ws + b->off_rcnt),
                           roff, outj);
        TJ_HIP(hipGetLastError());
        if (b->nj > b->n_small) {
            const int g1 = (int)std::min<int64_t>(b->nj - b->n_small, 2 * (int64_t)num_cus());
            hipLaunchKernelGGL(batch_join_rows_kernel<false>, dim3((unsigned)g1), dim3(JR_THREADS), 0, s,
                               b->d_jorder + b->n_small, (int)(b->nj - b->n_small), b->d_joins, b->nj, b->rows, joff,
                               bucket, (uint32_t*)(ws + b->off_rcnt), roff, outj);
            TJ_HIP(hipGetLastError());
        }
        if (b->rows > 0 && (rc = launch_row_sort(roff, b->rows, outj, outj2, b->out_cap, GRID_EMIT_I32, c.pairs,
                                                 b->d_row_local, (uint32_t*)(ws + b->off_long), nullptr, s)))
            return rc;
    }
    TJ_MARK(4);
#undef TJ_MARK
    return TJ_OK;
}

extern "C" {

int tj_batch_create(tj_batch** out, int joins, const int64_t* n, const int* d, const int* precision,
                    int64_t out_capacity, int64_t cand_capacity, unsigned flags) {
    TJ_CHECK(out != nullptr, TJ_ERR_INVALID, "batch pointer is NULL");
    *out = nullptr;
    TJ_CHECK(joins >= 1, TJ_ERR_INVALID, "a batch needs at least one join");
    TJ_CHECK(n && d && precision, TJ_ERR_INVALID, "NULL shape arrays");
    TJ_CHECK(out_capacity >= 0 && out_capacity < (1ll << 32), TJ_ERR_INVALID, "out_capacity must be in [0, 2^32)");
    TJ_CHECK(cand_capacity >= 0 && cand_capacity < (1ll << 32), TJ_ERR_INVALID, "cand_capacity must be in [0, 2^32)");
    std::vector<BatchJoin> J(joins);
    int64_t rows = 0, max_n = 0, pairs = 0;
    for (int t = 0; t < joins; ++t) {
        TJ_CHECK(n[t] >= 0 && n[t] < (1ll << 31), TJ_ERR_INVALID, "join %d: bad size %lld", t, (long long)n[t]);
        TJ_CHECK(d[t] >= 1, TJ_ERR_INVALID, "join %d: bad width %d", t, d[t]);
        TJ_CHECK(precision[t] == P_FP16 || precision[t] == P_FP8, TJ_ERR_UNSUPPORTED,
                 "join %d: the batch screen serves TJ_PREC_FP16 and TJ_PREC_FP8 (got %d)", t, precision[t]);
        const int dp = padded_dim(d[t], precision[t]);
        const int combo = dp > 0 ? combo_of(precision[t], dp) : -1;
        TJ_CHECK(combo >= 0, TJ_ERR_UNSUPPORTED, "join %d: the batch screen serves d <= 256 (d=%d)", t, d[t]);
        J[t].gbase = rows;
        J[t].n = (int32_t)n[t];
        J[t].d = d[t];
        J[t].dp = dp;
        J[t].prec = precision[t];
        J[t].bm = rows_per_wg(combo);
        J[t].group = combo;  // remapped to a dense group index below
        J[t].ball_off = 0;
        rows += n[t];
        max_n = std::max<int64_t>(max_n, n[t]);
        pairs += n[t] * (n[t] - 1) / 2;
    }
    // groups: one per (precision, width) present, in combo order
    int gid[16];
    std::vector<BatchGroup> G;
    for (int c = 0; c < 16; ++c) gid[c] = -1;
    std::vector<int64_t> grows;
    for (int t = 0; t < joins; ++t) {
        const int c = J[t].group;
        if (gid[c] < 0) {
            gid[c] = (int)G.size();
            BatchGroup g{};
            g.combo = c;
            g.prec = J[t].prec;
            g.dp = J[t].dp;
            G.push_back(g);
            grows.push_back(0);
        }
        const int g = gid[c];
        J[t].group = g;
        J[t].grow = grows[g];
        grows[g] += J[t].n;
    }
    tj_batch* b = new (std::nothrow) tj_batch();
    TJ_CHECK(b != nullptr, TJ_ERR_INVALID, "out of host memory");
    b->nj = joins;
    b->rows = rows;
    b->max_n = max_n;
    b->pairs_total = pairs;
    b->out_cap = out_capacity;
    b->cand_cap = cand_capacity;
    b->flags = flags;
    b->n_groups = (int)G.size();
    if (rows >= (1ll << 31)) {
        delete b;
        set_error("a batch holds fewer than 2^31 rows in total");
        return TJ_ERR_UNSUPPORTED;
    }
    b->ibits = bits_for(rows > 1 ? rows : 2);
    b->jbits = bits_for(max_n > 1 ? max_n : 2);
    b->keybits = b->ibits + b->jbits;
    b->prune = (flags & TJ_FLAG_PRUNE) != 0;
    b->order_bits = bits_for(joins > 1 ? joins : 2) + PIV_BITS + b->jbits;
    b->ball_floats = 0;
    for (int t = 0; t < joins; ++t) {
        J[t].ball_off = b->ball_floats;
        if (J[t].n >= 2) b->ball_floats += (int64_t)((J[t].n + J[t].bm - 1) / J[t].bm) * (J[t].d + 1);
    }
    if (b->order_bits > 63) b->prune = false;
    if (b->keybits > 63) {
        delete b;
        set_error("pair keys need more than 63 bits");
        return TJ_ERR_UNSUPPORTED;
    }
    // workspace layout
    size_t cur = 0;
    b->off_maxabs = bcarve(cur, 4 * (size_t)joins + 8);
    b->off_cursor = bcarve(cur, 8 * N_FAMILIES);
    b->off_params = bcarve(cur, sizeof(JoinParams) * (size_t)joins);
    for (size_t g = 0; g < G.size(); ++g) {
        const int es = prec_esize(G[g].prec);
        G[g].q_off = (int64_t)bcarve(cur, (size_t)grows[g] * G[g].dp * es + 16);
        G[g].meta_off = (int64_t)bcarve(cur, sizeof(RowMeta) * (size_t)grows[g] + 16);
    }
    b->off_cand = bcarve(cur, 8 * (size_t)std::max<int64_t>(cand_capacity, 1));
    b->off_aux = bcarve(cur, 4 * (size_t)std::max<int64_t>(cand_capacity, 1));
    b->off_out = bcarve(cur, 8 * (size_t)std::max<int64_t>(out_capacity, 1));
    // row-grouped output: per-join counts, the keys bucketed by join, the
    // global row counters of joins beyond ROWS_LDS rows, the row offsets,
    // the list of long rows.  The ids in row order and the long rows' merge
    // buffer reuse the output keys (two uint32 per key).
    b->off_jcnt = bcarve(cur, 4 * (size_t)(joins + 2));
    b->off_jpart = bcarve(cur, row_scan_part_bytes(joins));
    b->off_bucket = bcarve(cur, 8 * (size_t)std::max<int64_t>(out_capacity, 1));
    b->off_rcnt = bcarve(cur, 4 * (size_t)(rows + 2));
    b->off_roff = bcarve(cur, 8 * (size_t)(rows + 1));
    b->off_long = bcarve(cur, row_sort_scratch_bytes(rows));
    b->off_sort = b->prune ? bcarve(cur, sort_temp_bytes(std::max<int64_t>(rows, 1))) : 0;  // the spatial order
    b->off_pivots = b->off_perm = b->off_balls = b->off_cend = b->off_rows = 0;
    if (b->prune) {
        b->off_pivots = bcarve(cur, 4 * (size_t)NPIV * joins);
        b->off_perm = bcarve(cur, 8 * (size_t)std::max<int64_t>(rows, 1));
        b->off_balls = bcarve(cur, 4 * (size_t)std::max<int64_t>(b->ball_floats, 1));
        b->off_rows = bcarve(cur, 8);
    }
    b->total = (cur + 255) / 256 * 256;

    // work list: every panel of every join with at least two rows, largest
    // first (estimated MFMA steps plus a per-block overhead), ties in
    // (join, panel) order
    struct Item { int fam; double cost; int t, p; };
    std::vector<Item> items;
    for (int t = 0; t < joins; ++t) {
        if (J[t].n < 2) continue;
        const int bm = J[t].bm;
        const int ks = J[t].dp / prec_kstep(J[t].prec);
        const int rb = bm / 128;
        const int panels = (J[t].n + bm - 1) / bm;
        for (int p = 0; p < panels; ++p) {
            const int64_t cols = (int64_t)J[t].n - (int64_t)p * bm;
            const double nblk = (double)((cols + 31) / 32);
            items.push_back({family_of(J[t].prec, J[t].dp), nblk * (16.0 * ks * rb + 48.0), t, p});
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& c) {
        return a.fam != c.fam ? a.fam < c.fam : a.cost > c.cost;
    });
    b->n_items = (int64_t)items.size();
    if (b->prune) {
        b->off_cend = bcarve(cur, 4 * (size_t)RSTRIDE * std::max<int64_t>(b->n_items, 1));
        b->total = (cur + 255) / 256 * 256;
    }
    for (int f = 0; f <= N_FAMILIES; ++f) b->fam_begin[f] = 0;
    for (const Item& x : items) b->fam_begin[x.fam + 1]++;
    for (int f = 0; f < N_FAMILIES; ++f) b->fam_begin[f + 1] += b->fam_begin[f];
    std::vector<int2> it(items.size() > 0 ? items.size() : 1);
    for (size_t k = 0; k < items.size(); ++k) it[k] = make_int2(items[k].t, items[k].p);
    std::vector<int32_t> rj(rows > 0 ? rows : 1);
    std::vector<int32_t> jo(joins);
    for (int t = 0; t < joins; ++t) jo[t] = t;
    std::stable_sort(jo.begin(), jo.end(), [&](int x, int y) {
        const bool sx = J[x].n <= ROWS_LDS, sy = J[y].n <= ROWS_LDS;
        return sx != sy ? sx : J[x].n > J[y].n;
    });
    int64_t n_small = 0;
    for (int t = 0; t < joins; ++t) n_small += J[t].n <= ROWS_LDS ? 1 : 0;
    std::vector<int32_t> rl(rows > 0 ? rows : 1);
    for (int t = 0; t < joins; ++t)
        for (int64_t r = 0; r < J[t].n; ++r) {
            rj[J[t].gbase + r] = t;
            rl[J[t].gbase + r] = (int32_t)r;
        }
    std::vector<int64_t> cmeta(joins + 1, 0), cscan(joins + 1, 0), corder(joins + 1, 0);
    for (int t = 0; t < joins; ++t) {
        const int rpw = 64 / tjmeta::group_width(J[t].dp);
        const int64_t mrows = meta_unit_rows(J[t].dp);
        cmeta[t + 1] = cmeta[t] + (J[t].n + mrows - 1) / mrows;
        cscan[t + 1] = cscan[t] + (J[t].n + (int64_t)SCAN_SUB * rpw - 1) / ((int64_t)SCAN_SUB * rpw);
        corder[t + 1] = corder[t] + (J[t].n + ORDER_ROWS - 1) / ORDER_ROWS;
    }
    b->units_meta = cmeta[joins];
    b->units_scan = cscan[joins];
    b->units_order = corder[joins];
    for (int f = 0; f < N_FAMILIES; ++f) {
        const int64_t nf = b->fam_begin[f + 1] - b->fam_begin[f];
        b->fam_grid[f] = (int)std::max<int64_t>(1, std::min<int64_t>(nf, (int64_t)screen_batch_occupancy(f) * num_cus()));
    }

    // device tables
    size_t tcur = 0;
    const size_t o_j = bcarve(tcur, sizeof(BatchJoin) * J.size());
    const size_t o_g = bcarve(tcur, sizeof(BatchGroup) * G.size());
    const size_t o_i = bcarve(tcur, sizeof(int2) * it.size());
    const size_t o_r = bcarve(tcur, sizeof(int32_t) * rj.size());
    const size_t o_l = bcarve(tcur, sizeof(int32_t) * rl.size());
    const size_t o_o = bcarve(tcur, sizeof(int32_t) * jo.size());
    const size_t o_cm = bcarve(tcur, sizeof(int64_t) * cmeta.size());
    const size_t o_cs = bcarve(tcur, sizeof(int64_t) * cscan.size());
    const size_t o_co = bcarve(tcur, sizeof(int64_t) * corder.size());
    hipError_t e = hipMalloc(&b->dev, tcur);
    if (e != hipSuccess) {
        delete b;
        return hip_fail(e, "hipMalloc(batch tables)");
    }
    uint8_t* base = (uint8_t*)b->dev;
    b->d_joins = (BatchJoin*)(base + o_j);
    b->d_groups = (BatchGroup*)(base + o_g);
    b->d_items = (int2*)(base + o_i);
    b->d_row_join = (int32_t*)(base + o_r);
    b->d_row_local = (int32_t*)(base + o_l);
    b->d_jorder = (int32_t*)(base + o_o);
    b->n_small = n_small;
    b->d_cmeta = (int64_t*)(base + o_cm);
    b->d_cscan = (int64_t*)(base + o_cs);
    b->d_corder = (int64_t*)(base + o_co);
    if ((e = hipMemcpy(b->d_joins, J.data(), sizeof(BatchJoin) * J.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_groups, G.data(), sizeof(BatchGroup) * G.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_items, it.data(), sizeof(int2) * it.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_row_join, rj.data(), sizeof(int32_t) * rj.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_row_local, rl.data(), sizeof(int32_t) * rl.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_jorder, jo.data(), sizeof(int32_t) * jo.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_cmeta, cmeta.data(), sizeof(int64_t) * cmeta.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_cscan, cscan.data(), sizeof(int64_t) * cscan.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->d_corder, corder.data(), sizeof(int64_t) * corder.size(), hipMemcpyHostToDevice)) != hipSuccess) {
        (void)hipFree(b->dev);
        delete b;
        return hip_fail(e, "hipMemcpy(batch tables)");
    }
    b->timing = false;
    if (flags & TJ_FLAG_TIMING) {
        for (int k = 0; k < 5; ++k) {
            if ((e = hipEventCreate(&b->ev[k])) != hipSuccess) {
                for (int q = 0; q < k; ++q) (void)hipEventDestroy(b->ev[q]);
                (void)hipFree(b->dev);
                delete b;
                return hip_fail(e, "hipEventCreate");
            }
        }
        b->timing = true;
    }
    b->has_graph = false;
    *out = b;
    return TJ_OK;
}

int tj_batch_destroy(tj_batch* b) {
    if (!b) return TJ_OK;
    if (b->has_graph) (void)hipGraphExecDestroy(b->graph);
    if (b->timing)
        for (int k = 0; k < 5; ++k) (void)hipEventDestroy(b->ev[k]);
    (void)hipFree(b->dev);
    delete b;
    return TJ_OK;
}

int tj_batch_workspace_bytes(const tj_batch* b, size_t* bytes) {
    TJ_CHECK(b && bytes, TJ_ERR_INVALID, "NULL argument");
    *bytes = b->total;
    return TJ_OK;
}

int tj_batch_describe(const tj_batch* b, char* buf, size_t len) {
    TJ_CHECK(b && buf, TJ_ERR_INVALID, "NULL argument");
    snprintf(buf, len,
             "{\"joins\":%d,\"rows\":%lld,\"max_n\":%lld,\"pairs\":%lld,\"groups\":%d,\"items\":%lld,"
             "\"items_fp16\":%lld,\"items_fp16b\":%lld,\"items_fp8\":%lld,\"items_fp8b\":%lld,\"launches\":%d,"
             "\"ibits\":%d,\"jbits\":%d,\"keybits\":%d,\"out_capacity\":%lld,\"cand_capacity\":%lld,"
             "\"workspace\":%zu}",
             b->nj, (long long)b->rows, (long long)b->max_n, (long long)b->pairs_total, b->n_groups,
             (long long)b->n_items, (long long)(b->fam_begin[1] - b->fam_begin[0]),
             (long long)(b->fam_begin[2] - b->fam_begin[1]), (long long)(b->fam_begin[3] - b->fam_begin[2]),
             (long long)(b->fam_begin[4] - b->fam_begin[3]),
             (int)(b->fam_begin[1] > b->fam_begin[0]) + (int)(b->fam_begin[2] > b->fam_begin[1]) +
                 (int)(b->fam_begin[3] > b->fam_begin[2]) + (int)(b->fam_begin[4] > b->fam_begin[3]),
             b->ibits, b->jbits, b->keybits, (long long)b->out_cap,
             (long long)b->cand_cap, b->total);
    return TJ_OK;
}

int tj_batch_stage_times(tj_batch* b, float* ms, int n) {
    TJ_CHECK(b && ms, TJ_ERR_INVALID, "NULL argument");
    TJ_CHECK(b->timing, TJ_ERR_INVALID, "batch was created without TJ_FLAG_TIMING");
    TJ_HIP(hipEventSynchronize(b->ev[4]));
    for (int k = 0; k < n && k < 4; ++k) TJ_HIP(hipEventElapsedTime(&ms[k], b->ev[k], b->ev[k + 1]));
    return TJ_OK;
}

static int batch_check(const tj_batch* b, const float* const* x, const double* T, void* ws, int32_t* pairs,
                       int64_t* offsets, uint64_t* stats) {
    TJ_CHECK(b != nullptr, TJ_ERR_INVALID, "batch is NULL");
    TJ_CHECK(x && T && ws && offsets && stats, TJ_ERR_INVALID, "NULL argument");
    TJ_CHECK(pairs != nullptr || b->out_cap == 0, TJ_ERR_INVALID, "pairs buffer is NULL");
    return TJ_OK;
}

int tj_batch_join(tj_batch* b, const float* const* x, const double* T, void* workspace, int32_t* pairs,
                  int64_t* offsets, uint64_t* stats, void* stream) {
    int rc = batch_check(b, x, T, workspace, pairs, offsets, stats);
    if (rc) return rc;
    const BatchCall c{b, x, T, workspace, pairs, offsets, stats, true};
    return batch_enqueue(c, (hipStream_t)stream);
}

static int batch_capture_fn(void* ctx, hipStream_t s) { return batch_enqueue(*(const BatchCall*)ctx, s); }

int tj_batch_capture(tj_batch* b, const float* const* x, const double* T, void* workspace, int32_t* pairs,
                     int64_t* offsets, uint64_t* stats) {
    int rc = batch_check(b, x, T, workspace, pairs, offsets, stats);
    if (rc) return rc;
    BatchCall c{b, x, T, workspace, pairs, offsets, stats, false};
    hipGraphExec_t exec;
    if ((rc = capture_graph(&exec, batch_capture_fn, &c))) return rc;
    if (b->has_graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = exec;
    b->has_graph = true;
    return TJ_OK;
}

int tj_batch_replay(tj_batch* b, void* stream) {
    TJ_CHECK(b != nullptr, TJ_ERR_INVALID, "batch is NULL");
    TJ_CHECK(b->has_graph, TJ_ERR_INVALID, "no graph captured for this batch");
    TJ_HIP(hipGraphLaunch(b->graph, (hipStream_t)stream));
    return TJ_OK;
}

}  // extern "C"
