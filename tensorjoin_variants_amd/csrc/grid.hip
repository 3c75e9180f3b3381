// grid.hip -- uniform-grid epsilon-join for low-dimensional points (d <= 3).
//
// Counterpart of GDS-Join (GDS/main.cu:690-950 index build on the CPU,
// GDS/kernel.cu:132-486 one query per thread over 3^k neighbour cells,
// GDS/GPU.cu:775-872 batches with thrust::sort_by_key of the results).
// Same index idea -- cells of side >= eps, a query only visits its own and
// the adjacent cells -- rebuilt for MI355X:
//  * the index is built on the device: bounds, cell keys (cell id << index
//    bits | point index), the radix sort of the keys (the digits wholly
//    inside the index need no pass: the keys are written in index order),
//    and the points written in key order -- coordinates widened to binary64
//    (exact) and the original index, 32 bytes a row, so the probe converts
//    nothing per candidate; S's non-empty cells go into an occupancy bitmap
//    with a prefix count per word (grids of at most 4096 cells per S point
//    and 2^28 cells) or
//    an open-addressing hash of [start, end) runs (sparser grids).  (A
//    counting sort over the cells, one device atomic per point, measured
//    slower: scattered device atomics cost more than the sort's passes);
//  * with the bitmap the three x-adjacent cells of a neighbour row are
//    consecutive ids and consecutive ranks, so their S rows are one
//    contiguous range: nine ranges per query, found from one bitmap word,
//    one prefix count and two run bounds each; sparser grids look every
//    neighbour cell up in the hash (27 ranges);
//  * queries run in cell order, so the 64 lanes of a wave visit the same
//    neighbour cells and the same S rows;
//  * one probe pass: each wave sums its lanes' candidate ranges (an upper
//    bound of their neighbours), reserves that much of its workgroup's
//    slice of a candidate buffer with one atomic, and every lane writes its
//    neighbour ids into its own segment as it finds them, storing its count
//    and segment start; an
//    exclusive scan of the counts gives each row's final range, and the
//    per-row sort (rowsort.hip) reads each segment, orders it and emits the
//    (i, j) rows -- or the neighbour table of GDS-Join (counts per point +
//    ids grouped by point, GDS/main.cu:507-540) -- at the row's final range.
//    Waves whose candidates overflow that buffer (8 ids per output slot + 2
//    per query) count their neighbours first and reserve exactly that many
//    in a region of out_cap ids, so every output that fits its capacity is
//    written in one run;
//  * the predicate is the sequential binary64 difference form of the CPU
//    oracle (d2 <= T; GDS used sqrt(d2) <= eps, GDS/kernel.cu:368-371).
// The cell side is sqrt(T) * (1 + 2^-10) (or larger when an axis would
// span more than 2^38 cells or the sort keys would not fit 63 bits): every
// pair with d2_seq <= T has all coordinate differences below one cell side,
// so it lies in adjacent cells.
#include <algorithm>
#include "tj_internal.h"

namespace {

struct GridParams {
    double lo[3];     // grid origin per dimension (one cell below the minimum)
    double inv_side;  // 1 / side
    double side;
    uint64_t ncell[3];
    uint64_t mul[3];  // linear id = sum c_k * mul[k] (mul[0] = 1: x-adjacent cells are adjacent ids)
    uint64_t cells;   // product of ncell
    int keybits, d, use_bm;  // use_bm: cells fit the occupancy bitmap
    int xb;  // bits of the sub-cell x position in S's sort keys (rows of a cell ordered by x)
};

// A point of the index: coordinates widened to binary64 (exact) and the
// point's index in its input array.
struct GPoint {
    double x, y, z;
    uint32_t idx, pad;
};

__device__ inline uint32_t f2ord(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float ord2f(uint32_t o) {
    const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __builtin_bit_cast(float, u);
}

// Per block: part[6 b + k] = ordered min, part[6 b + 3 + k] = ordered max
// of the finite coordinates of dimension k (reduced by bounds_reduce_kernel:
// one same-address atomic per block had serialised into most of the time);
// the non-finite ones are counted in stats[TJ_ST_INVALID] (the join is then
// rejected).  Three-dimensional points go four at a time as three 16-byte
// loads.
__global__ void __launch_bounds__(256) bounds_kernel(const float* __restrict__ x, int64_t n, int d, uint32_t* part,
                                                     uint64_t* __restrict__ stats) {
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0, 0, 0};
    uint32_t bad = 0;
    auto take = [&](int k, float f) {
        if ((__builtin_bit_cast(uint32_t, f) & 0x7FFFFFFFu) >= 0x7F800000u) { bad++; return; }
        const uint32_t o = f2ord(f);
        mn[k] = o < mn[k] ? o : mn[k];
        mx[k] = o > mx[k] ? o : mx[k];
    };
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n4 = (d == 3 && ((uintptr_t)x & 15) == 0) ? n / 4 : 0;
    const float4* xv = (const float4*)x;
    for (int64_t g = tid; g < n4; g += stride) {
        const float4 a = xv[3 * g], b = xv[3 * g + 1], c = xv[3 * g + 2];
        take(0, a.x); take(1, a.y); take(2, a.z);
        take(0, a.w); take(1, b.x); take(2, b.y);
        take(0, b.z); take(1, b.w); take(2, c.x);
        take(0, c.y); take(1, c.z); take(2, c.w);
    }
    for (int64_t t = 4 * n4 + tid; t < n; t += stride)
        for (int k = 0; k < d; ++k) take(k, x[t * d + k]);
    // wave, then block reduction
    __shared__ uint32_t smn[3][4], smx[3][4], sbad[4];
    const int wave = threadIdx.x >> 6;
    for (int k = 0; k < d; ++k) {
        for (int s = 32; s > 0; s >>= 1) {
            const uint32_t a = __shfl_xor(mn[k], s), b = __shfl_xor(mx[k], s);
            mn[k] = a < mn[k] ? a : mn[k];
            mx[k] = b > mx[k] ? b : mx[k];
        }
        if ((threadIdx.x & 63) == 0) { smn[k][wave] = mn[k]; smx[k][wave] = mx[k]; }
    }
    for (int s = 32; s > 0; s >>= 1) bad += __shfl_xor(bad, s);
    if ((threadIdx.x & 63) == 0) sbad[wave] = bad;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t nb = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) nb += sbad[w];
        if (nb) atomicAdd((unsigned long long*)&stats[TJ_ST_INVALID], (unsigned long long)nb);
        for (int k = 0; k < 3; ++k) {
            uint32_t a = 0xFFFFFFFFu, b = 0;
            if (k < d)
                for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                    a = smn[k][w] < a ? smn[k][w] : a;
                    b = smx[k][w] > b ? smx[k][w] : b;
                }
            part[6 * blockIdx.x + k] = a;
            part[6 * blockIdx.x + 3 + k] = b;
        }
    }
}

// bounds[0..5] from the nparts block partials.
__global__ void __launch_bounds__(256) bounds_reduce_kernel(const uint32_t* __restrict__ part, int nparts,
                                                            uint32_t* __restrict__ bounds) {
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
        for (int k = 0; k < 3; ++k) {
            const uint32_t a = part[6 * b + k], c = part[6 * b + 3 + k];
            v[k] = a < v[k] ? a : v[k];
            v[3 + k] = c > v[3 + k] ? c : v[3 + k];
        }
    __shared__ uint32_t sv[6][256];
    for (int k = 0; k < 6; ++k) sv[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 6; ++k) {
                const uint32_t a = sv[k][threadIdx.x], c = sv[k][threadIdx.x + o];
                sv[k][threadIdx.x] = k < 3 ? (c < a ? c : a) : (c > a ? c : a);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) bounds[threadIdx.x] = sv[threadIdx.x][0];
}

__global__ void grid_params_kernel(const uint32_t* bounds, double T, int d, int idxbits, uint64_t bm_bits,
                                   GridParams* gp) {
    if (threadIdx.x != 0) return;
    GridParams g;
    double side = sqrt(T) * (1.0 + 0x1p-10);
    if (!(side > 0.0)) side = 0x1p-100;  // T = 0: only identical points pair up
    double lo[3] = {0, 0, 0}, ext[3] = {0, 0, 0};
    for (int k = 0; k < d; ++k) {
        const double a = (double)ord2f(bounds[k]), b = (double)ord2f(bounds[3 + k]);
        lo[k] = a;
        ext[k] = b > a ? b - a : 0.0;
    }
    // Widen the cell until (1) no axis has more than 2^38 cells and (2) the
    // linear id of every (padded) cell fits beside the index bits of the
    // sort keys.  (1) keeps cell_of
    // exact enough: its binary64 value (x - lo) / side carries a relative
    // error below 2^-51, i.e. below 2^-13 cells at 2^38 cells, far inside the
    // 2^-10 slack of the side, so two points closer than sqrt(T) on an axis
    // always land in the same or adjacent cells.  Without it one far outlier
    // stretches an axis to ~2^55 cells, where rounding moves a point by
    // several cells and the probe misses pairs.
    for (;;) {
        int bits = 0;
        bool wide = false;
        for (int k = 0; k < d; ++k) {
            const double nc = floor(ext[k] / side) + 4.0;
            if (nc > 0x1p38) wide = true;
            int b = 0;
            while (b < 62 && ldexp(1.0, b) < nc) ++b;
            bits += b;
        }
        if (!wide && bits + idxbits <= 63) break;
        side *= 2.0;
    }
    uint64_t mul = 1;
    int bits = 0;
    for (int k = 0; k < 3; ++k) {
        if (k < d) {
            g.lo[k] = lo[k] - side;  // cell 0 stays empty: neighbours c-1 >= 0
            g.ncell[k] = (uint64_t)(floor(ext[k] / side) + 4.0);
        } else {
            g.lo[k] = 0.0;
            g.ncell[k] = 1;
        }
        g.mul[k] = mul;
        mul *= g.ncell[k];
    }
    while (bits < 63 && (1ull << bits) < mul) ++bits;
    g.side = side;
    g.inv_side = 1.0 / side;
    g.keybits = bits;
    g.d = d;
    g.cells = mul;
    g.use_bm = mul <= bm_bits ? 1 : 0;
    // S's keys order each cell's rows by their x position, quantized to xb
    // bits (as many as fit, up to 8)
    g.xb = 63 - bits - idxbits < 8 ? 63 - bits - idxbits : 8;
    *gp = g;
}

// Clear the words of the occupancy bitmap the grid uses.
__global__ void __launch_bounds__(256) clear_bitmap_kernel(const GridParams* __restrict__ gp, uint64_t* __restrict__ bm) {
    const GridParams g = *gp;
    const int64_t words = g.use_bm ? (int64_t)((g.cells + 63) / 64) : 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < words; t += (int64_t)gridDim.x * blockDim.x)
        bm[t] = 0;
}

__device__ inline void cell_of(const double* x, const GridParams& g, int64_t c[3]) {
    // unrolled over the three axes (static indices keep c and x in registers)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c[k] = 0;
        if (k < g.d) {
            int64_t v = (int64_t)floor((x[k] - g.lo[k]) * g.inv_side);
            if (v < 1) v = 1;
            if (v > (int64_t)g.ncell[k] - 2) v = (int64_t)g.ncell[k] - 2;
            c[k] = v;
        }
    }
}

// Global quantized x of a point: its x cell << xb | its position within
// the cell in 2^xb steps -- the same binary64 expression as cell_of, so it
// is a non-decreasing function of x, and S's rows are ordered by it within
// every row of cells.
__device__ inline uint64_t xpos(double x, const GridParams& g) {
    const double v = (x - g.lo[0]) * g.inv_side;
    int64_t c = (int64_t)floor(v);
    if (c < 1) c = 1;
    if (c > (int64_t)g.ncell[0] - 2) c = (int64_t)g.ncell[0] - 2;
    const double f = floor((v - (double)c) * (double)(1 << g.xb));
    const int64_t q = f < 0.0 ? 0 : (f > (double)((1 << g.xb) - 1) ? (int64_t)((1 << g.xb) - 1) : (int64_t)f);
    return ((uint64_t)c << g.xb) | (uint64_t)q;
}

// Sort keys: cell id << idxbits | point index; with sub (S's keys) the cell's
// x position sits between them, so each cell's rows are ordered by x.  Three-dimensional points go
// four at a time as three 16-byte loads and two 16-byte stores.
__global__ void __launch_bounds__(256) keys_kernel(const float* __restrict__ x, int64_t n, const GridParams* __restrict__ gp,
                                                   int idxbits, int sub, uint64_t* __restrict__ keys) {
    const GridParams g = *gp;
    const int xb = sub ? g.xb : 0;
    auto key_of = [&](double x0, double x1, double x2, int64_t t) {
        const double q[3] = {x0, x1, x2};
        int64_t c[3];
        cell_of(q, g, c);
        const uint64_t id = (uint64_t)c[0] * g.mul[0] + (uint64_t)c[1] * g.mul[1] + (uint64_t)c[2] * g.mul[2];
        const uint64_t xs = xb ? xpos(x0, g) & ((1ull << xb) - 1) : 0;
        return (((id << xb) | xs) << idxbits) | (uint64_t)t;
    };
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n4 = (g.d == 3 && ((uintptr_t)x & 15) == 0 && ((uintptr_t)keys & 15) == 0) ? n / 4 : 0;
    const float4* xv = (const float4*)x;
    for (int64_t q = tid; q < n4; q += stride) {
        const float4 a = xv[3 * q], b = xv[3 * q + 1], c = xv[3 * q + 2];
        const int64_t t = 4 * q;
        ulonglong2* kv = (ulonglong2*)(keys + t);
        kv[0] = make_ulonglong2(key_of(a.x, a.y, a.z, t), key_of(a.w, b.x, b.y, t + 1));
        kv[1] = make_ulonglong2(key_of(b.z, b.w, c.x, t + 2), key_of(c.y, c.z, c.w, t + 3));
    }
    for (int64_t t = 4 * n4 + tid; t < n; t += stride) {
        const float* p = x + t * g.d;
        keys[t] = key_of(p[0], g.d > 1 ? p[1] : 0.0, g.d > 2 ? p[2] : 0.0, t);
    }
}

__device__ inline uint64_t hmix(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
    return k;
}
constexpr uint64_t EMPTY = ~0ull;
constexpr int GRID_SHARDS = 64;  // slices of the probe's candidate buffer

// The points in sorted-key (cell) order as index rows: coordinates widened
// to binary64 and the original index.  For S (hash != NULL) also the
// occupied cells: a bitmap bit per run start, or, on sparse grids, the
// run's [start, end) inserted into the hash.
__global__ void __launch_bounds__(256) rows_kernel(const float* __restrict__ x, int64_t n, const GridParams* __restrict__ gp,
                                                   int idxbits, const uint64_t* __restrict__ keys, GPoint* __restrict__ rows,
                                                   uint64_t* __restrict__ hkey, uint2* __restrict__ hval, uint64_t hmask,
                                                   uint64_t* __restrict__ bm) {
    const GridParams g = *gp;
    const uint64_t imask = (1ull << idxbits) - 1;
    const int csh = idxbits + (hkey ? g.xb : 0);  // cell id = key >> csh (S's keys carry x positions)
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t kk = keys[p];
        const uint64_t j = kk & imask, cell = kk >> csh;
        const float* y = x + j * g.d;
        GPoint r;
        r.x = y[0];
        r.y = g.d > 1 ? (double)y[1] : 0.0;
        r.z = g.d > 2 ? (double)y[2] : 0.0;
        r.idx = (uint32_t)j;
        r.pad = 0;
        rows[p] = r;
        if (!hkey || (p > 0 && (keys[p - 1] >> csh) == cell)) continue;
        if (g.use_bm) {
            atomicOr((unsigned long long*)&bm[cell >> 6], 1ull << (cell & 63));
            continue;
        }
        // end of the run: galloping search over the sorted keys
        int64_t lo = p, step = 1;  // keys[lo] is in the cell
        while (lo + step < n && (keys[lo + step] >> csh) == cell) { lo += step; step <<= 1; }
        int64_t hi = lo + step < n ? lo + step : n;  // keys[hi] is past the cell (or hi = n)
        while (hi - lo > 1) {
            const int64_t mid = lo + (hi - lo) / 2;
            if ((keys[mid] >> csh) == cell) lo = mid; else hi = mid;
        }
        uint64_t h = hmix(cell) & hmask;
        for (;;) {
            const unsigned long long prev = atomicCAS((unsigned long long*)&hkey[h], (unsigned long long)EMPTY,
                                                      (unsigned long long)cell);
            if (prev == EMPTY || prev == cell) break;
            h = (h + 1) & hmask;
        }
        hval[h] = make_uint2((uint32_t)p, (uint32_t)hi);
    }
}
__device__ inline uint2 lookup(const uint64_t* __restrict__ hkey, const uint2* __restrict__ hval, uint64_t hmask,
                               uint64_t cell) {
    uint64_t h = hmix(cell) & hmask;
    for (;;) {
        const uint64_t k = hkey[h];
        if (k == cell) return hval[h];
        if (k == EMPTY) return make_uint2(0, 0);
        h = (h + 1) & hmask;
    }
}

// Exclusive prefix of the bitmap words' popcounts: bmpre[w] = occupied cells
// below word w.  Three passes (block sums of 1024 words, one block over the
// sums, apply); the word count comes from the device-side grid parameters.
constexpr int BMB = 1024;
__device__ inline int64_t bm_words(const GridParams& g) { return g.use_bm ? (int64_t)((g.cells + 63) / 64) : 0; }
__global__ void __launch_bounds__(256) bm_block_sums(const GridParams* __restrict__ gp, const uint64_t* __restrict__ bm,
                                                     uint32_t* __restrict__ bsum) {
    const int64_t W = bm_words(*gp);
    const int64_t b0 = (int64_t)blockIdx.x * BMB;
    if (b0 >= W) return;
    uint32_t s = 0;
    for (int k = threadIdx.x; k < BMB; k += 256) if (b0 + k < W) s += (uint32_t)__popcll(bm[b0 + k]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ uint32_t part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}
__global__ void __launch_bounds__(1024) bm_top_scan(const GridParams* __restrict__ gp, uint32_t* __restrict__ bsum) {
    const int64_t nb = (bm_words(*gp) + BMB - 1) / BMB;
    __shared__ uint32_t part[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t a = threadIdx.x * per, b = a + per < nb ? a + per : nb;
    uint32_t s = 0;
    for (int64_t t = a; t < b; ++t) s += bsum[t];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int64_t t = a; t < b; ++t) {
        const uint32_t v = bsum[t];
        bsum[t] = run;
        run += v;
    }
}
__global__ void __launch_bounds__(256) bm_apply(const GridParams* __restrict__ gp, const uint64_t* __restrict__ bm,
                                                const uint32_t* __restrict__ bsum, uint32_t* __restrict__ bmpre) {
    const int64_t W = bm_words(*gp);
    const int64_t b0 = (int64_t)blockIdx.x * BMB;
    if (b0 >= W) return;
    uint32_t v[4], s = 0;  // each thread owns 4 consecutive words
    for (int u = 0; u < 4; ++u) {
        const int64_t t = b0 + threadIdx.x * 4 + u;
        v[u] = t < W ? (uint32_t)__popcll(bm[t]) : 0;
        s += v[u];
    }
    __shared__ uint32_t part[256];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t w = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += w;
        __syncthreads();
    }
    uint32_t run = bsum[blockIdx.x] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
    for (int u = 0; u < 4; ++u) {
        const int64_t t = b0 + threadIdx.x * 4 + u;
        if (t < W) bmpre[t] = run;
        run += v[u];
    }
}
__device__ inline uint32_t cell_rank(const uint64_t* __restrict__ bm, const uint32_t* __restrict__ bmpre, uint64_t cell) {
    const uint64_t w = bm[cell >> 6];
    const uint32_t b = (uint32_t)(cell & 63);
    return bmpre[cell >> 6] + (uint32_t)__popcll(b ? (w & ((1ull << b) - 1)) : 0ull);
}
// runs[rank(cell)] = first sorted S position of each non-empty cell; the
// entry after the last run holds n.
__global__ void __launch_bounds__(256) runs_kernel(int64_t n, const GridParams* __restrict__ gp, int idxbits,
                                                   const uint64_t* __restrict__ keys, const uint64_t* __restrict__ bm,
                                                   const uint32_t* __restrict__ bmpre, uint32_t* __restrict__ runs) {
    const GridParams g = *gp;
    if (!g.use_bm) return;
    const int csh = idxbits + g.xb;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t cell = keys[p] >> csh;
        if (p == 0 || (keys[p - 1] >> csh) != cell) runs[cell_rank(bm, bmpre, cell)] = (uint32_t)p;
        if (p == n - 1) runs[cell_rank(bm, bmpre, cell) + 1] = (uint32_t)n;
    }
}

// The sequential binary64 difference form over D <= 3 coordinates.
template <int D>
__device__ inline double d2_point(const double* q, const GPoint& y) {
    double s = 0.0, t;
    t = q[0] - y.x; s = s + t * t;
    if (D > 1) { t = q[1] - y.y; s = s + t * t; }
    if (D > 2) { t = q[2] - y.z; s = s + t * t; }
    return s;
}

// A lane's 64-bit value broadcast to the wave (lane k wave-uniform).
__device__ inline uint64_t rl_u64(uint64_t v, int k) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), k);
    return ((uint64_t)hi << 32) | lo;
}
__device__ inline double rl_f64(double v, int k) { return __builtin_bit_cast(double, rl_u64(__builtin_bit_cast(uint64_t, v), k)); }

__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}
__device__ inline uint64_t shfl_up_u64(uint64_t v, int o) {
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, o);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), o);
    return ((uint64_t)hi << 32) | lo;
}

// One pass over the queries, in cell order, 64 queries per workgroup step
// (query l on lane l of every wave) and PW waves sharing them: wave w finds
// the candidate ranges of neighbour rows w, w + PW, ... and the workgroup's
// waves split the sweeps (the short ranges of a row by the row's wave, the
// chunks of the long ranges round robin), so a step over a dense blob, whose
// 64 queries hold a few thousand candidates each, runs on PW waves at once.
// Each query's neighbour count so far lives in LDS (sMine, LDS atomics).
// After the ranges, wave 0 reserves the step's candidate total in its
// workgroup's slice of the candidate buffer outj (one atomic on
// cursor[slice]; same-address atomics serialise) and each query's hits are
// written into its own segment as they are found (ascending within each run
// of S rows, unordered across runs); cnt[i] and src[i] receive the query's
// neighbour count and segment start.  The buffer's last out_cap ids (from
// fast_cap on) take the steps whose candidates no longer fit: they count
// first and reserve their exact neighbour total there (cursor[GRID_SHARDS]),
// so any output that fits its capacity is written whatever the candidate
// total; cand_total_kernel adds the candidate total to stats[TJ_ST_GRID_CAND].
constexpr int PW = 4;  // waves per probe workgroup
// The probe waits on memory most of its time: registers are planned for
// seven workgroups per CU (at most 72 VGPRs; six at the 75 the compiler
// would take: 1.18 -> 1.11 ms on the 4.2 M x 4.2 M join; eight spill).
constexpr int PROBE_WG_PER_CU = 7;
template <int D, bool UPPER>
__global__ void __launch_bounds__(64 * PW, PROBE_WG_PER_CU) probe_kernel(const GPoint* __restrict__ rrow, int64_t nR,
                                                       const GridParams* __restrict__ gp, const GPoint* __restrict__ srow,
                                                       const uint64_t* __restrict__ hkey, const uint2* __restrict__ hval,
                                                       uint64_t hmask, const uint64_t* __restrict__ bm,
                                                       const uint32_t* __restrict__ bmpre, const uint32_t* __restrict__ runs,
                                                       double T, uint32_t* __restrict__ cnt, uint64_t* __restrict__ src,
                                                       uint32_t* __restrict__ outj, int64_t fast_cap, int64_t cand_cap,
                                                       uint64_t* __restrict__ cursor, uint64_t* __restrict__ stats) {
    constexpr int NROWS = D == 3 ? 9 : D == 2 ? 3 : 1;  // neighbour rows (dy, dz)
    const GridParams g = *gp;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ uint2 sRun[NROWS][64];  // S range of each query per neighbour row (bitmap grids)
    __shared__ uint32_t sCand[PW][64];  // candidates per query found by each wave
    __shared__ uint32_t sMine[64];      // neighbours per query so far
    __shared__ uint64_t sSeg[64];       // segment start per query
    __shared__ int sPhase;
    // this workgroup's slice of the fast region (one cursor per slice)
    const int shard = (int)(blockIdx.x % GRID_SHARDS);
    const int64_t slice = fast_cap / GRID_SHARDS;
    const int64_t stride = (int64_t)gridDim.x * 64;
    for (int64_t tb = (int64_t)blockIdx.x * 64; tb < nR; tb += stride) {
        const int64_t t = tb + lane;
        const bool valid = t < nR;
        GPoint qp;
        if (valid) qp = rrow[t];
        else qp = GPoint{0.0, 0.0, 0.0, 0u, 0u};
        const uint32_t i = qp.idx;
        const double q[3] = {qp.x, qp.y, qp.z};
        int64_t c[3];
        cell_of(q, g, c);
        // neighbour row rr -> its middle cell (the x-adjacent cells mid - 1,
        // mid, mid + 1; cell_of keeps c in [1, ncell - 2], so all exist)
        auto row_mid = [&](int rr) {
            const int dz = D == 3 ? rr / 3 - 1 : 0, dy = D >= 2 ? rr % 3 - 1 : 0;
            return (uint64_t)c[0] + (uint64_t)(c[1] + dy) * g.mul[1] + (uint64_t)(c[2] + dz) * g.mul[2];
        };
        // Reach of neighbour row rr: dd <= the squared yz distance from the
        // query to any point of the row's cells (the cell faces, pulled in by
        // 2^-12 of a side: more than the binary64 cell assignment can move a
        // point).  A row with dd > Tw holds no neighbour (Tw exceeds T by far
        // more than the gap between d2_seq and the exact d2); otherwise a long
        // range keeps only the rows whose x lies within w = sqrt(Tw - dd) of
        // the query's (a cell's S rows are ordered by x position, xpos): the
        // rest are at least w apart in x alone.
        const double mg = g.side * 0x1p-12, Tw = T * (1.0 + 0x1p-30);
        auto trim = [&](int rr, uint2& run) {
            const int dz = D == 3 ? rr / 3 - 1 : 0, dy = D >= 2 ? rr % 3 - 1 : 0;
            double dd = 0.0;
            if (dy) {
                const double f = g.lo[1] + (double)(c[1] + (dy > 0 ? 1 : 0)) * g.side;
                const double t = (dy > 0 ? f - q[1] : q[1] - f) - mg;
                if (t > 0.0) dd += t * t;
            }
            if (dz) {
                const double f = g.lo[2] + (double)(c[2] + (dz > 0 ? 1 : 0)) * g.side;
                const double t = (dz > 0 ? f - q[2] : q[2] - f) - mg;
                if (t > 0.0) dd += t * t;
            }
            if (dd > Tw) {
                run = make_uint2(0, 0);
                return;
            }
            if (!g.xb || run.y - run.x <= 32) return;
            const double w = sqrt(Tw - dd) + mg;
            const uint64_t xl = xpos(q[0] - w, g), xh = xpos(q[0] + w, g);
            uint32_t lo = run.x, hi = run.y;  // first row with xpos >= xl
            while (lo < hi) {
                const uint32_t m = lo + (hi - lo) / 2;
                if (xpos(srow[m].x, g) < xl) lo = m + 1;
                else hi = m;
            }
            uint32_t lo2 = lo, hi2 = run.y;  // first row with xpos > xh
            while (lo2 < hi2) {
                const uint32_t m = lo2 + (hi2 - lo2) / 2;
                if (xpos(srow[m].x, g) <= xh) lo2 = m + 1;
                else hi2 = m;
            }
            run = make_uint2(lo, lo2);
        };
        // ---- candidate ranges of this wave's rows and their total
        uint32_t mycand = 0;
        // (bitmap grids: this wave's rows -- at most three -- in three
        // rounds of independent loads: bitmap words and prefix counts, run
        // bounds, then the x-window trims)
        constexpr int MYR = (NROWS + PW - 1) / PW;
        if (g.use_bm) {
            uint64_t w0[MYR], b0[MYR];
            uint32_t pre[MYR];
#pragma unroll
            for (int u = 0; u < MYR; ++u) {
                const int rr = wave + u * PW;
                const bool use = valid && rr < NROWS;
                b0[u] = row_mid(use ? rr : 0) - 1;
                w0[u] = use ? bm[b0[u] >> 6] : 0ull;
                pre[u] = use ? bmpre[b0[u] >> 6] : 0u;
            }
            uint2 run[MYR];
#pragma unroll
            for (int u = 0; u < MYR; ++u) {
                const int rr = wave + u * PW;
                const bool use = valid && rr < NROWS;
                const uint64_t b2 = b0[u] + 2;
                const uint32_t sh = (uint32_t)(b0[u] & 63);
                uint32_t occ = use ? (uint32_t)((w0[u] >> sh) & 7) : 0u;
                if (use && sh > 61 && (b2 >> 6) != (b0[u] >> 6)) occ |= (uint32_t)(bm[b2 >> 6] << (64 - sh)) & 7;
                run[u] = make_uint2(0, 0);
                if (occ) {
                    const uint32_t rk = pre[u] + (uint32_t)__popcll(sh ? (w0[u] & ((1ull << sh) - 1)) : 0ull);
                    run[u] = make_uint2(runs[rk], runs[rk + (uint32_t)__popc(occ)]);
                }
            }
#pragma unroll
            for (int u = 0; u < MYR; ++u) {
                const int rr = wave + u * PW;
                if (rr >= NROWS) break;
                if (run[u].y > run[u].x) trim(rr, run[u]);
                sRun[rr][lane] = run[u];
                mycand += run[u].y - run[u].x;
            }
        } else if (valid) {
            for (int rr = wave; rr < NROWS; rr += PW)
                for (int dx = -1; dx <= 1; ++dx) {
                    const uint2 run = lookup(hkey, hval, hmask, row_mid(rr) + (uint64_t)(int64_t)dx);
                    mycand += run.y - run.x;
                }
        }
        sCand[wave][lane] = mycand;
        __syncthreads();
        // ---- one reservation per step: the candidate total in the fast
        // region; phase 2 sweeps and writes, phase 0 only counts
        if (wave == 0) {
            uint32_t cand = 0;
#pragma unroll
            for (int w = 0; w < PW; ++w) cand += sCand[w][lane];
            uint64_t inc = cand;
            for (int o = 1; o < 64; o <<= 1) {
                const uint64_t v = shfl_up_u64(inc, o);
                if (lane >= o) inc += v;
            }
            const uint64_t W = shfl_u64(inc, 63);
            uint64_t base = 0;
            if (lane == 0 && W) base = atomicAdd((unsigned long long*)&cursor[shard], (unsigned long long)W);
            base = shfl_u64(base, 0);
            sSeg[lane] = (uint64_t)shard * (uint64_t)slice + base + (inc - cand);
            sMine[lane] = 0;
            if (lane == 0) sPhase = base + W <= (uint64_t)slice ? 2 : 0;
        }
        __syncthreads();
        int phase = sPhase;
        for (;;) {
            // a hit of query k (own lane or broadcast): its position in the
            // segment from the LDS count
            auto sweep_own = [&](uint2 run) {
                const uint64_t seg = sSeg[lane];
                for (uint32_t p = run.x; p < run.y; ++p) {
                    const GPoint y = srow[p];
                    if (UPPER && y.idx <= i) continue;
                    if (d2_point<D>(q, y) <= T) {
                        const uint32_t at = atomicAdd(&sMine[lane], 1u);
                        if (phase) outj[seg + at] = y.idx;
                    }
                }
            };
            for (int rr = 0; rr < NROWS; ++rr) {
                const int own = rr % PW;  // the wave of this row's short ranges
                if (!g.use_bm) {
                    // sparse grid: each lane sweeps its own three cells
                    if (wave == own && valid)
                        for (int dx = -1; dx <= 1; ++dx)
                            sweep_own(lookup(hkey, hval, hmask, row_mid(rr) + (uint64_t)(int64_t)dx));
                    continue;
                }
                const uint2 run = sRun[rr][lane];
                const uint32_t len = run.y - run.x;
                const bool longr = len > 32;
                if (wave == own && !longr && len) sweep_own(run);
                if (__ballot(longr) == 0) continue;
                // Long ranges (a dense blob).  The queries of a step sit in the
                // same few cells, so their ranges overlap: the long ranges are
                // walked in chunks of 128 rows, two rows per lane (coalesced
                // loads), skipping rows no long range holds; chunk t of row rr
                // belongs to wave (t + rr) % PW.  The queries whose range meets
                // the chunk are taken one at a time, broadcast from their lane
                // (readlane), and every lane tests its two rows against the
                // query; the query's hits in the chunk go to consecutive
                // positions of its segment (ballot + mbcnt + one LDS atomic):
                // coalesced stores.  A query whose range covers the whole chunk
                // needs no range test.
                uint32_t ch = longr ? run.x : 0xFFFFFFFFu;
                for (int o = 32; o > 0; o >>= 1) ch = min(ch, (uint32_t)__shfl_xor((int)ch, o));
                ch = __builtin_amdgcn_readfirstlane(ch);
                for (int tch = rr; ch != 0xFFFFFFFFu; ++tch) {
                    uint32_t cend = longr && run.x < ch + 128 && run.y > ch ? min(run.y, ch + 128) : 0u;
                    for (int o = 32; o > 0; o >>= 1) cend = max(cend, (uint32_t)__shfl_xor((int)cend, o));
                    cend = __builtin_amdgcn_readfirstlane(cend);
                    if (tch % PW == wave) {
                        const uint32_t p0 = ch + (uint32_t)lane, p1 = p0 + 64;
                        const bool h0 = p0 < cend, h1 = p1 < cend;
                        GPoint y0, y1;
                        if (h0) y0 = srow[p0];
                        else y0 = GPoint{0.0, 0.0, 0.0, 0u, 0u};
                        if (h1) y1 = srow[p1];
                        else y1 = GPoint{0.0, 0.0, 0.0, 0u, 0u};
                        // wait for the rows here: the query loop below then has no
                        // load in flight, and its stores need no wait
                        __builtin_amdgcn_s_waitcnt(0);
                        uint64_t qm = __ballot(longr && run.x < cend && run.y > ch);
                        while (qm) {
                            const int k = __builtin_ctzll(qm);
                            qm &= qm - 1;
                            const double qk[3] = {rl_f64(q[0], k), rl_f64(q[1], k), rl_f64(q[2], k)};
                            const uint32_t rx = (uint32_t)__builtin_amdgcn_readlane((int)run.x, k);
                            const uint32_t ry = (uint32_t)__builtin_amdgcn_readlane((int)run.y, k);
                            const uint32_t ik = UPPER ? (uint32_t)__builtin_amdgcn_readlane((int)i, k) : 0u;
                            bool t0, t1;
                            if (rx <= ch && ry >= cend) {
                                t0 = h0 && !(UPPER && y0.idx <= ik) && d2_point<D>(qk, y0) <= T;
                                t1 = h1 && !(UPPER && y1.idx <= ik) && d2_point<D>(qk, y1) <= T;
                            } else {
                                t0 = p0 >= rx && p0 < ry && !(UPPER && y0.idx <= ik) && d2_point<D>(qk, y0) <= T;
                                t1 = p1 >= rx && p1 < ry && !(UPPER && y1.idx <= ik) && d2_point<D>(qk, y1) <= T;
                            }
                            const uint64_t m0 = __ballot(t0), m1 = __ballot(t1);
                            if (m0 | m1) {
                                const uint32_t n0 = (uint32_t)__popcll(m0), n = n0 + (uint32_t)__popcll(m1);
                                uint32_t at = 0;
                                if (lane == 0) at = atomicAdd(&sMine[k], n);
                                at = (uint32_t)__builtin_amdgcn_readfirstlane((int)at);
                                if (phase) {
                                    const uint64_t sk = sSeg[k] + at;
                                    const uint32_t r0 = __builtin_amdgcn_mbcnt_hi(
                                        (uint32_t)(m0 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m0, 0u));
                                    const uint32_t r1 = __builtin_amdgcn_mbcnt_hi(
                                        (uint32_t)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u));
                                    if (t0) outj[sk + r0] = y0.idx;
                                    if (t1) outj[sk + n0 + r1] = y1.idx;
                                }
                            }
                        }
                    }
                    // the next chunk starts at the first row past this one that a
                    // long range holds
                    uint32_t nx = longr && run.y > ch + 128 ? max(run.x, ch + 128) : 0xFFFFFFFFu;
                    for (int o = 32; o > 0; o >>= 1) nx = min(nx, (uint32_t)__shfl_xor((int)nx, o));
                    ch = __builtin_amdgcn_readfirstlane(nx);
                }
            }
            __syncthreads();
            if (phase) break;
            // exact reservation of the counted neighbours
            if (wave == 0) {
                const uint32_t h = sMine[lane];
                uint64_t hinc = h;
                for (int o = 1; o < 64; o <<= 1) {
                    const uint64_t v = shfl_up_u64(hinc, o);
                    if (lane >= o) hinc += v;
                }
                const uint64_t H = shfl_u64(hinc, 63);
                uint64_t b2 = 0;
                if (lane == 0 && H) b2 = atomicAdd((unsigned long long*)&cursor[GRID_SHARDS], (unsigned long long)H);
                b2 = shfl_u64(b2, 0);
                const bool fits = b2 + H <= (uint64_t)(cand_cap - fast_cap);
                // when the output passes its capacity the counts stand and
                // nothing is written (segment 0 keeps the row sort's reads in
                // bounds)
                sSeg[lane] = fits ? (uint64_t)fast_cap + b2 + (hinc - h) : 0;
                sMine[lane] = fits ? 0 : h;
                if (lane == 0) sPhase = fits ? 2 : 1;
            }
            __syncthreads();
            phase = sPhase;
            if (phase == 1) break;
        }
        if (wave == 0 && valid) {
            cnt[i] = sMine[lane];
            src[i] = sSeg[lane];
        }
        __syncthreads();
    }
}

// stats[TJ_ST_GRID_CAND] = the candidates of all steps: every step adds its
// total to its slice's cursor, whether it fits or not (one atomic per step
// on a shared counter would serialise).
__global__ void cand_total_kernel(const uint64_t* __restrict__ cursor, uint64_t* __restrict__ stats) {
    if (threadIdx.x != 0) return;
    uint64_t t = 0;
    for (int s = 0; s < GRID_SHARDS; ++s) t += cursor[s];
    stats[TJ_ST_GRID_CAND] += t;
}

}  // namespace

namespace tj {

// Blocks of one bounds pass (a partial per block, no atomics): four per CU.
static int64_t bounds_blocks() { return 4 * (int64_t)num_cus(); }

// The candidate buffer: a fast region of 8 ids per output slot + 2 per
// query, then out_cap ids for the waves that reserve exactly.
int64_t grid_cand_capacity(int64_t nR, int64_t out_cap) { return 9 * out_cap + 2 * nR + 1024; }

GridLayout grid_layout(int64_t nR, int64_t nS, int64_t out_cap) {
    GridLayout L{};
    size_t cur = 0;
    auto carve = [&](size_t bytes) { size_t at = (cur + 255) / 256 * 256; cur = at + bytes; return at; };
    const int64_t nmax = std::max<int64_t>(std::max<int64_t>(nR, nS), 1);
    uint64_t hs = 1;
    while (hs < (uint64_t)(2 * std::max<int64_t>(nS, 1))) hs <<= 1;
    L.hslots = hs;
    L.idxbits = bits_for(nmax > 1 ? nmax : 2);
    // occupancy bitmap of S: up to 4096 cells per S point, at least 2^20
    // and at most 2^28 cells (32 MiB; only the words the grid uses are
    // cleared and scanned); sparser grids go without it
    L.bm_bits = std::max<uint64_t>(1ull << 20,
                                   std::min<uint64_t>(1ull << 28, 4096ull * (uint64_t)std::max<int64_t>(nS, 1)));
    L.bounds = carve(64);
    L.bpart = carve(4 * 6 * 2 * (size_t)bounds_blocks());
    L.params = carve(sizeof(GridParams));
    L.count_nr = carve(8);
    L.count_ns = carve(8);
    L.rkeys = carve(8 * (size_t)std::max<int64_t>(nR, 1));
    L.skeys = carve(8 * (size_t)std::max<int64_t>(nS, 1));
    L.sort_tmp = carve(sort_temp_bytes(nmax));
    L.srow = carve(sizeof(GPoint) * (size_t)std::max<int64_t>(nS, 1));
    L.rrow = carve(sizeof(GPoint) * (size_t)std::max<int64_t>(nR, 1));
    L.hkey = carve(8 * hs);
    L.hval = carve(8 * hs);
    L.bitmap = carve(L.bm_bits / 8);
    L.bmpre = carve(4 * (L.bm_bits / 64));
    L.bmsum = carve(4 * (L.bm_bits / 64 / BMB + 1));
    L.runs = carve(4 * (size_t)(std::max<int64_t>(nS, 1) + 1));
    // output: the slice cursors, per-query counts, segment starts and final
    // offsets, the candidate buffer, the long rows' merge buffer, the row
    // sort's scratch
    L.cand_cap = grid_cand_capacity(nR, out_cap);
    L.cursor = carve(8 * (GRID_SHARDS + 1));
    L.cnt = carve(4 * (size_t)std::max<int64_t>(nR, 1));
    L.src = carve(8 * (size_t)std::max<int64_t>(nR, 1));
    L.off = carve(8 * (size_t)(nR + 1));
    L.scan_part = carve(row_scan_part_bytes(nR));
    L.outj = carve(4 * (size_t)L.cand_cap);
    L.outj2 = carve(4 * (size_t)(out_cap > 0 ? out_cap : 1));
    L.longrows = carve(row_sort_scratch_bytes(nR));
    L.total = (cur + 255) / 256 * 256;
    return L;
}

static unsigned blocks_for(int64_t n, int64_t cap) {
    int64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

int launch_grid_join(const float* R, int64_t nR, const float* S, int64_t nS, int d, double T, bool self_join,
                     bool upper, uint8_t* ws, const GridLayout& L, int64_t out_cap, int emit, void* pairs,
                     uint32_t* counts, uint64_t* stats, hipStream_t s) {
    TJ_CHECK(d >= 1 && d <= 3, TJ_ERR_UNSUPPORTED, "grid join serves d <= 3 (d=%d)", d);
    TJ_CHECK(nR < (1ll << 32) && nS < (1ll << 32), TJ_ERR_UNSUPPORTED, "grid join: point sets beyond 2^32");
    if (self_join) { S = R; nS = nR; }
    if (nR == 0 || nS == 0) return TJ_OK;
    uint32_t* bounds = (uint32_t*)(ws + L.bounds);
    GridParams* gp = (GridParams*)(ws + L.params);
    uint64_t* rkeys = (uint64_t*)(ws + L.rkeys);
    uint64_t* skeys = (uint64_t*)(ws + L.skeys);
    GPoint* srow = (GPoint*)(ws + L.srow);
    uint64_t* hkey = (uint64_t*)(ws + L.hkey);
    uint2* hval = (uint2*)(ws + L.hval);
    uint64_t* cR = (uint64_t*)(ws + L.count_nr);
    uint64_t* cS = (uint64_t*)(ws + L.count_ns);
    const uint64_t hmask = L.hslots - 1;
    int rc;
    const unsigned gb = (unsigned)(num_cus() * 4);
    uint32_t* bpart = (uint32_t*)(ws + L.bpart);
    const unsigned bbR = blocks_for((nR + 3) / 4, bounds_blocks());
    hipLaunchKernelGGL(bounds_kernel, dim3(bbR), dim3(256), 0, s, R, nR, d, bpart, stats);
    unsigned nparts = bbR;
    if (!self_join) {
        const unsigned bbS = blocks_for((nS + 3) / 4, bounds_blocks());
        hipLaunchKernelGGL(bounds_kernel, dim3(bbS), dim3(256), 0, s, S, nS, d, bpart + 6 * (size_t)bbR, stats);
        nparts += bbS;
    }
    hipLaunchKernelGGL(bounds_reduce_kernel, dim3(1), dim3(256), 0, s, bpart, (int)nparts, bounds);
    hipLaunchKernelGGL(grid_params_kernel, dim3(1), dim3(64), 0, s, bounds, T, d, L.idxbits, L.bm_bits, gp);
    uint64_t* bm = (uint64_t*)(ws + L.bitmap);
    hipLaunchKernelGGL(clear_bitmap_kernel, dim3(blocks_for((int64_t)(L.bm_bits / 64), gb)), dim3(256), 0, s, gp, bm);
    if ((rc = launch_fill_u64(cR, 1, (uint64_t)nR, s))) return rc;
    if ((rc = launch_fill_u64(cS, 1, (uint64_t)nS, s))) return rc;
    if ((rc = launch_fill_u64(hkey, (int64_t)L.hslots, EMPTY, s))) return rc;
    // R and S sorted by cell (key = cell id << idxbits | index); a self-join
    // sorts once and reads its queries from S's rows.  The keys are written
    // in index order, so the digits wholly inside the index need no pass.
    const int lo_digit = L.idxbits / 8 * 8;
    GPoint* rrow = self_join ? srow : (GPoint*)(ws + L.rrow);
    if (!self_join) {
        hipLaunchKernelGGL(keys_kernel, dim3(blocks_for((nR + 3) / 4, gb)), dim3(256), 0, s, R, nR, gp, L.idxbits, 0, rkeys);
        if ((rc = launch_sort(ws + L.sort_tmp, rkeys, cR, nR, 63, s, nullptr, lo_digit))) return rc;
        hipLaunchKernelGGL(rows_kernel, dim3(blocks_for(nR, gb)), dim3(256), 0, s, R, nR, gp, L.idxbits,
                           (const uint64_t*)rkeys, rrow, (uint64_t*)nullptr, (uint2*)nullptr, hmask, (uint64_t*)nullptr);
    }
    hipLaunchKernelGGL(keys_kernel, dim3(blocks_for((nS + 3) / 4, gb)), dim3(256), 0, s, S, nS, gp, L.idxbits, 1, skeys);
    if ((rc = launch_sort(ws + L.sort_tmp, skeys, cS, nS, 63, s, nullptr, lo_digit))) return rc;
    hipLaunchKernelGGL(rows_kernel, dim3(blocks_for(nS, gb)), dim3(256), 0, s, S, nS, gp, L.idxbits,
                       (const uint64_t*)skeys, srow, hkey, hval, hmask, bm);
    // rank structure of the bitmap (its kernels return at once without one)
    uint32_t* bmpre = (uint32_t*)(ws + L.bmpre);
    uint32_t* bmsum = (uint32_t*)(ws + L.bmsum);
    uint32_t* runs = (uint32_t*)(ws + L.runs);
    const unsigned nbm = (unsigned)((L.bm_bits / 64 + BMB - 1) / BMB);
    hipLaunchKernelGGL(bm_block_sums, dim3(nbm), dim3(256), 0, s, gp, bm, bmsum);
    hipLaunchKernelGGL(bm_top_scan, dim3(1), dim3(1024), 0, s, gp, bmsum);
    hipLaunchKernelGGL(bm_apply, dim3(nbm), dim3(256), 0, s, gp, bm, bmsum, bmpre);
    hipLaunchKernelGGL(runs_kernel, dim3(blocks_for(nS, gb)), dim3(256), 0, s, nS, gp, L.idxbits, skeys, bm, bmpre, runs);
    // probe -> offsets -> per-row sort and emit
    uint32_t* cnt = counts ? counts : (uint32_t*)(ws + L.cnt);
    uint64_t* src = (uint64_t*)(ws + L.src);
    uint64_t* off = (uint64_t*)(ws + L.off);
    uint32_t* outj = (uint32_t*)(ws + L.outj);
    uint32_t* outj2 = (uint32_t*)(ws + L.outj2);
    uint64_t* cursor = (uint64_t*)(ws + L.cursor);
    if ((rc = launch_fill_u64(cursor, GRID_SHARDS + 1, 0, s))) return rc;
    // (one instantiation per dimension and triangle flag: the distance has
    // no runtime branches)
    auto kern = upper ? (d == 1 ? probe_kernel<1, true> : d == 2 ? probe_kernel<2, true> : probe_kernel<3, true>)
                      : (d == 1 ? probe_kernel<1, false> : d == 2 ? probe_kernel<2, false> : probe_kernel<3, false>);
    // one workgroup per group of 64 queries (up to 2^20 workgroups, then
    // they stride): the groups' work is very uneven (a blob core's queries
    // hold thousands of candidates, the uniform field's about one), and the
    // dispatcher hands each freed CU slot the next group
    const unsigned pb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((nR + 63) / 64, (int64_t)1 << 20));
    hipLaunchKernelGGL(kern, dim3(pb), dim3(64 * PW), 0, s, (const GPoint*)rrow, nR, gp, (const GPoint*)srow,
                       (const uint64_t*)hkey, (const uint2*)hval, hmask, (const uint64_t*)bm, (const uint32_t*)bmpre,
                       (const uint32_t*)runs, T, cnt, src, outj, (int64_t)(L.cand_cap - out_cap), (int64_t)L.cand_cap,
                       cursor, stats);
    hipLaunchKernelGGL(cand_total_kernel, dim3(1), dim3(64), 0, s, (const uint64_t*)cursor, stats);
    if ((rc = launch_row_offsets(cnt, nR, (uint64_t*)(ws + L.scan_part), off, stats + TJ_ST_OUT, s))) return rc;
    if ((rc = launch_row_sort(off, nR, outj, outj2, out_cap, emit, pairs, nullptr, (uint32_t*)(ws + L.longrows), src, s)))
        return rc;
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

}  // namespace tj
