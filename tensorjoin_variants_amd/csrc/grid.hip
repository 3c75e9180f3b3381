// grid.hip -- uniform-grid epsilon-join for low-dimensional points (d <= 3).
//
// Counterpart of GDS-Join (GDS/main.cu:690-950 index build on the CPU,
// GDS/kernel.cu:132-486 one query per thread over 3^k neighbour cells,
// GDS/GPU.cu:775-872 batches with thrust::sort_by_key of the results).
// Same index idea -- cells of side >= eps, a query only visits its own and
// the adjacent cells -- rebuilt for MI355X:
//  * the index is built on the device: bounds (order-preserving integer
//    atomics), cell keys, the radix sort by cell, S gathered into cell order
//    as 16-byte rows, and an open-addressing hash of the non-empty cells
//    (memory bounded by 2|S| slots however sparse the grid is);
//  * when the grid has at most 64 |S| cells, no hash at all: a bitmap of
//    the non-empty cells with a prefix count per 64-bit word ranks every
//    cell, and runs[rank] is where its S rows start.  The three x-adjacent
//    cells of a neighbour row are consecutive ids, so their S rows are one
//    contiguous range [runs[rank(mid - 1)], runs[rank(mid - 1) + occupied]):
//    a query reads nine bitmap words (shared by the cell-ordered lanes of a
//    wave), nine prefix counts and 18 run bounds, where hashing cost a probe
//    sequence of random 64-byte lines per neighbour cell, empty or not;
//  * queries run in cell order (R is gathered into cell order as 16-byte
//    rows too; a self-join reuses S's), so the 64 lanes of a wave visit the
//    same neighbour cells and the same S rows, and read their queries,
//    counts and offsets contiguously;
//  * output without a global sort: a count pass stores each query's number
//    of neighbours, an exclusive scan turns the counts into row offsets, a
//    write pass stores every neighbour id in its row's final range, and a
//    per-row sort orders each row (by its lane, its wave or, for rows beyond
//    512 ids, its workgroup; rowsort.hip) while it emits the (i, j) rows -- the same
//    shape as GDS-Join's neighbour table (counts per point + ids grouped by
//    point, GDS/main.cu:507-540), which the table output gives directly;
//  * the predicate is the sequential binary64 difference form of the CPU
//    oracle (d2 <= T; GDS used sqrt(d2) <= eps, GDS/kernel.cu:368-371).
// The cell side is sqrt(T) * (1 + 2^-10) (or larger when an axis would
// span more than 2^38 cells or the cell ids would not fit a 64-bit key):
// every pair with d2_seq <= T has all coordinate differences below one cell
// side, so it lies in adjacent cells.
#include <type_traits>
#include "tj_internal.h"

namespace {

struct GridParams {
    double lo[3];     // grid origin per dimension (one cell below the minimum)
    double inv_side;  // 1 / side
    double side;
    uint64_t ncell[3];
    uint64_t mul[3];  // linear id = sum c_k * mul[k] (mul[0] = 1: x-adjacent cells are adjacent ids)
    uint64_t cells;   // product of ncell
    int keybits, idxbits, d, use_bm;  // use_bm: cells fit the occupancy bitmap
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
    // linear id of every (padded) cell fits beside the index bits.  (1) keeps
    // cell_of exact enough: its binary64 value (x - lo) / side carries a
    // relative error below 2^-51, i.e. below 2^-13 cells at 2^38 cells, far
    // inside the 2^-10 slack of the side, so two points closer than sqrt(T)
    // on an axis always land in the same or adjacent cells.  Without it one
    // far outlier stretches an axis to ~2^55 cells, where rounding moves a
    // point by several cells and the probe misses pairs.
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
    g.idxbits = idxbits;
    g.d = d;
    g.cells = mul;
    g.use_bm = mul <= bm_bits ? 1 : 0;
    *gp = g;
}

// Clear the words of the occupancy bitmap the grid uses.
__global__ void __launch_bounds__(256) clear_bitmap_kernel(const GridParams* __restrict__ gp, uint64_t* __restrict__ bm) {
    const GridParams g = *gp;
    const int64_t words = g.use_bm ? (int64_t)((g.cells + 63) / 64) : 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < words; t += (int64_t)gridDim.x * blockDim.x)
        bm[t] = 0;
}

__device__ inline void cell_of(const float* x, const GridParams& g, int64_t c[3]) {
    // unrolled over the three axes (static indices keep c and x in registers)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c[k] = 0;
        if (k < g.d) {
            int64_t v = (int64_t)floor(((double)x[k] - g.lo[k]) * g.inv_side);
            if (v < 1) v = 1;
            if (v > (int64_t)g.ncell[k] - 2) v = (int64_t)g.ncell[k] - 2;
            c[k] = v;
        }
    }
}

__global__ void __launch_bounds__(256) keys_kernel(const float* __restrict__ x, int64_t n, const GridParams* __restrict__ gp,
                                                   uint64_t* __restrict__ keys) {
    const GridParams g = *gp;
    auto key_of = [&](const float* p, int64_t t) {
        int64_t c[3];
        cell_of(p, g, c);
        const uint64_t id = (uint64_t)c[0] * g.mul[0] + (uint64_t)c[1] * g.mul[1] + (uint64_t)c[2] * g.mul[2];
        return (id << g.idxbits) | (uint64_t)t;
    };
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    // three-dimensional points four at a time: three 16-byte loads, two
    // 16-byte stores
    const int64_t n4 = (g.d == 3 && ((uintptr_t)x & 15) == 0 && ((uintptr_t)keys & 15) == 0) ? n / 4 : 0;
    const float4* xv = (const float4*)x;
    for (int64_t q = tid; q < n4; q += stride) {
        const float4 a = xv[3 * q], b = xv[3 * q + 1], c = xv[3 * q + 2];
        const float p0[3] = {a.x, a.y, a.z}, p1[3] = {a.w, b.x, b.y}, p2[3] = {b.z, b.w, c.x}, p3[3] = {c.y, c.z, c.w};
        const int64_t t = 4 * q;
        ulonglong2* kv = (ulonglong2*)(keys + t);
        kv[0] = make_ulonglong2(key_of(p0, t), key_of(p1, t + 1));
        kv[1] = make_ulonglong2(key_of(p2, t + 2), key_of(p3, t + 3));
    }
    for (int64_t t = 4 * n4 + tid; t < n; t += stride) keys[t] = key_of(x + t * g.d, t);
}

__device__ inline uint64_t hmix(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
    return k;
}
constexpr uint64_t EMPTY = ~0ull;

// R in cell order: 16-byte rows (x, y, z, original index as bits), so the
// probes read their queries and counts contiguously.
__global__ void __launch_bounds__(256) rows_kernel(const float* __restrict__ x, int64_t n, const GridParams* __restrict__ gp,
                                                   const uint64_t* __restrict__ keys, float4* __restrict__ rows) {
    const GridParams g = *gp;
    const uint64_t imask = (1ull << g.idxbits) - 1;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t j = keys[p] & imask;
        float4 r = make_float4(0.f, 0.f, 0.f, __builtin_bit_cast(float, (uint32_t)j));
        const float* y = x + j * g.d;
        r.x = y[0];
        if (g.d > 1) r.y = y[1];
        if (g.d > 2) r.z = y[2];
        rows[p] = r;
    }
}

// S in cell order: 16-byte rows (x, y, z, original index as bits); run starts
// of each non-empty cell inserted into the hash (slot: key, start, end).
__global__ void __launch_bounds__(256) gather_kernel(const float* __restrict__ S, int64_t n, const GridParams* __restrict__ gp,
                                                     const uint64_t* __restrict__ keys, float4* __restrict__ srow,
                                                     uint64_t* __restrict__ hkey, uint2* __restrict__ hval, uint64_t hmask,
                                                     uint64_t* __restrict__ bm) {
    const GridParams g = *gp;
    const uint64_t imask = (1ull << g.idxbits) - 1;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t kk = keys[p];
        const uint64_t j = kk & imask, cell = kk >> g.idxbits;
        float4 r = make_float4(0.f, 0.f, 0.f, __builtin_bit_cast(float, (uint32_t)j));
        const float* y = S + j * g.d;
        r.x = y[0];
        if (g.d > 1) r.y = y[1];
        if (g.d > 2) r.z = y[2];
        srow[p] = r;
        if (p == 0 || (keys[p - 1] >> g.idxbits) != cell) {
            if (g.use_bm) {
                atomicOr((unsigned long long*)&bm[cell >> 6], 1ull << (cell & 63));
                continue;
            }
            // end of the run: galloping search over the sorted keys
            int64_t lo = p, step = 1;  // keys[lo] is in the cell
            while (lo + step < n && (keys[lo + step] >> g.idxbits) == cell) { lo += step; step <<= 1; }
            int64_t hi = lo + step < n ? lo + step : n;  // keys[hi] is past the cell (or hi = n)
            while (hi - lo > 1) {
                const int64_t mid = lo + (hi - lo) / 2;
                if ((keys[mid] >> g.idxbits) == cell) lo = mid; else hi = mid;
            }
            const int64_t e = hi;
            uint64_t h = hmix(cell) & hmask;
            for (;;) {
                const unsigned long long prev = atomicCAS((unsigned long long*)&hkey[h], (unsigned long long)EMPTY,
                                                          (unsigned long long)cell);
                if (prev == EMPTY || prev == cell) break;
                h = (h + 1) & hmask;
            }
            hval[h] = make_uint2((uint32_t)p, (uint32_t)e);
        }
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
__global__ void __launch_bounds__(256) runs_kernel(int64_t n, const GridParams* __restrict__ gp,
                                                   const uint64_t* __restrict__ keys, const uint64_t* __restrict__ bm,
                                                   const uint32_t* __restrict__ bmpre, uint32_t* __restrict__ runs) {
    const GridParams g = *gp;
    if (!g.use_bm) return;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t cell = keys[p] >> g.idxbits;
        if (p == 0 || (keys[p - 1] >> g.idxbits) != cell) runs[cell_rank(bm, bmpre, cell)] = (uint32_t)p;
        if (p == n - 1) runs[cell_rank(bm, bmpre, cell) + 1] = (uint32_t)n;
    }
}

__device__ inline uint2 lookup(const uint64_t* __restrict__ hkey, const uint2* __restrict__ hval, uint64_t hmask, uint64_t cell) {
    uint64_t h = hmix(cell) & hmask;
    for (;;) {
        const uint64_t k = hkey[h];
        if (k == cell) return hval[h];
        if (k == EMPTY) return make_uint2(0, 0);
        h = (h + 1) & hmask;
    }
}

// The sequential binary64 difference form over D <= 3 coordinates.
template <int D>
__device__ inline double d2_row(const float* q, const float4& y) {
    double s = 0.0, t;
    t = (double)q[0] - (double)y.x; s = s + t * t;
    if (D > 1) { t = (double)q[1] - (double)y.y; s = s + t * t; }
    if (D > 2) { t = (double)q[2] - (double)y.z; s = s + t * t; }
    return s;
}

__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// One query per lane, queries in cell order (rrow: x, y, z, original index).
// The same traversal runs twice: the count pass stores each query's number
// of neighbours at cnt[i] (and counts candidate evaluations); after an
// exclusive scan of the counts, the write pass stores each neighbour id j
// at outj[off[i] + k], k the query's running count, so every row's ids land
// in their final range without any global atomic.  A row's ids arrive in
// cell order, not sorted: the row sort kernels order each row afterwards.
// Positions at or past out_cap are skipped (the caller retries larger).
enum { PROBE_COUNT = 0, PROBE_WRITE = 1 };

template <int MODE, int D>
__global__ void __launch_bounds__(256) probe_kernel(const float4* __restrict__ rrow, int64_t nR, const GridParams* __restrict__ gp,
                                                    const float4* __restrict__ srow,
                                                    const uint64_t* __restrict__ hkey, const uint2* __restrict__ hval,
                                                    uint64_t hmask, const uint64_t* __restrict__ bm,
                                                    const uint32_t* __restrict__ bmpre, const uint32_t* __restrict__ runs,
                                                    double T, int upper, uint32_t* __restrict__ cnt,
                                                    const uint64_t* __restrict__ off, uint32_t* __restrict__ outj,
                                                    int64_t out_cap, uint64_t* __restrict__ stats) {
    const GridParams g = *gp;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // neighbours found for each lane's query by other lanes (wave-shared sweep)
    __shared__ uint32_t sCnt[4][64];
    // each lane's S range per neighbour row
    __shared__ uint2 sRun[4][9][64];
    sCnt[wave][lane] = 0;
    uint64_t cand = 0;
    // whole waves step together (tb is the wave's first query), so the
    // shared sweep always has all 64 lanes
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t tb = (int64_t)blockIdx.x * blockDim.x + wave * 64; tb < nR; tb += stride) {
        const int64_t t = tb + lane;
        const bool valid = t < nR;
        const float4 qr = valid ? rrow[t] : make_float4(0.f, 0.f, 0.f, 0.f);
        const uint32_t i = __builtin_bit_cast(uint32_t, qr.w);
        const float q[3] = {qr.x, qr.y, qr.z};
        const uint64_t base = (MODE == PROBE_WRITE && valid) ? off[i] : 0;
        uint32_t mine = 0;  // neighbours of this lane's query so far
        int64_t c[3];
        cell_of(q, g, c);
        auto visit = [&](const float4& y) {
            const uint32_t j = __builtin_bit_cast(uint32_t, y.w);
            if (upper && j <= i) return;
            if (MODE == PROBE_COUNT) ++cand;
            if (d2_row<D>(q, y) <= T) {
                if (MODE == PROBE_WRITE) {
                    const uint64_t at = base + mine;
                    if (at < (uint64_t)out_cap) outj[at] = j;
                }
                ++mine;
            }
        };
        auto sweep = [&](uint2 run) {
            // eight rows in flight at a time
            uint32_t p = run.x;
            for (; p + 8 <= run.y; p += 8) {
                float4 y[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) y[u] = srow[p + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) visit(y[u]);
            }
            for (; p < run.y; ++p) visit(srow[p]);
        };
        const int r1 = g.d > 1 ? 1 : 0, r2 = g.d > 2 ? 1 : 0;
        // ids of the x-adjacent cells of neighbour row (dy, dz): mid - 1,
        // mid, mid + 1 (cell_of keeps c in [1, ncell - 2], so all exist)
        auto row_mid = [&](int dy, int dz) {
            return (uint64_t)c[0] + (uint64_t)(c[1] + dy) * g.mul[1] + (uint64_t)(c[2] + dz) * g.mul[2];
        };
        if (!g.use_bm) {
            // sparse grid: one hash run per cell
            if (valid)
                for (int dz = -r2; dz <= r2; ++dz)
                    for (int dy = -r1; dy <= r1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx)
                            sweep(lookup(hkey, hval, hmask, row_mid(dy, dz) + (uint64_t)(int64_t)dx));
            if (MODE == PROBE_COUNT && valid) cnt[i] = mine;
            continue;
        }
        // The S rows of each neighbour row's three cells are one range.  The
        // ranges of all (up to nine) rows are found first, with their loads
        // in two independent rounds (bitmap word + its prefix count, then the
        // run bounds) instead of nine dependent chains, and kept in LDS.
        {
            uint64_t w0[9];
            uint32_t pre[9];
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const int dz = r / 3 - 1, dy = r % 3 - 1;
                const bool use = valid && (dz == 0 || r2) && (dy == 0 || r1);
                const uint64_t b0 = row_mid(use ? dy : 0, use ? dz : 0) - 1;
                w0[r] = use ? bm[b0 >> 6] : 0ull;
                pre[r] = use ? bmpre[b0 >> 6] : 0u;
            }
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                const int dz = r / 3 - 1, dy = r % 3 - 1;
                const bool use = valid && (dz == 0 || r2) && (dy == 0 || r1);
                const uint64_t b0 = row_mid(use ? dy : 0, use ? dz : 0) - 1, b2 = b0 + 2;
                const uint32_t sh = (uint32_t)(b0 & 63);
                uint32_t occ = use ? (uint32_t)((w0[r] >> sh) & 7) : 0u;
                if (use && sh > 61 && (b2 >> 6) != (b0 >> 6)) occ |= (uint32_t)(bm[b2 >> 6] << (64 - sh)) & 7;
                uint2 run = make_uint2(0, 0);
                if (occ) {
                    const uint32_t r0 = pre[r] + (uint32_t)__popcll(sh ? (w0[r] & ((1ull << sh) - 1)) : 0ull);
                    run = make_uint2(runs[r0], runs[r0 + (uint32_t)__popc(occ)]);
                }
                sRun[wave][r][lane] = run;
            }
        }
        for (int r = 0; r < 9; ++r) {
            const int dz = r / 3 - 1, dy = r % 3 - 1;
            if ((dz != 0 && !r2) || (dy != 0 && !r1)) continue;
                const uint2 run = sRun[wave][r][lane];
                const uint32_t len = run.y - run.x;
                if (__ballot(len > 32) == 0) {
                    sweep(run);
                    continue;
                }
                // Long ranges (a dense blob).  Lanes of a wave hold queries of
                // the same few cells, so their ranges mostly overlap: the wave
                // walks the union [u0, u1) of its ranges row by row, every lane
                // reading the same row (one broadcast load) and evaluating it
                // only if it lies in its own range -- each pair is still met
                // once, and in ascending row order per lane.
                uint32_t u0 = len ? run.x : 0xFFFFFFFFu, u1 = len ? run.y : 0u, mx = len;
                for (int o = 32; o > 0; o >>= 1) {
                    u0 = min(u0, (uint32_t)__shfl_xor((int)u0, o));
                    u1 = max(u1, (uint32_t)__shfl_xor((int)u1, o));
                    mx = max(mx, (uint32_t)__shfl_xor((int)mx, o));
                }
                if (u1 - u0 <= 2 * mx + 64) {
                    const uint32_t a = __builtin_amdgcn_readfirstlane(u0), b = __builtin_amdgcn_readfirstlane(u1);
                    uint32_t pr = a;
                    for (; pr + 8 <= b; pr += 8) {
                        float4 y[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) y[u] = srow[pr + u];
#pragma unroll
                        for (int u = 0; u < 8; ++u)
                            if (pr + u >= run.x && pr + u < run.y) visit(y[u]);
                    }
                    for (; pr < b; ++pr) {
                        const float4 y = srow[pr];
                        if (pr >= run.x && pr < run.y) visit(y);
                    }
                    continue;
                }
                // A lane near a dense blob holds hundreds of rows while the
                // rest hold a few: the wave sweeps the concatenation of its
                // lanes' ranges together, 64 rows a step (row k belongs to
                // the lane whose range prefix holds k); hits are counted per
                // owner in LDS and the owner's next free position is its
                // own count so far plus that LDS count.
                uint32_t inc = len;
                for (int o = 1; o < 64; o <<= 1) {
                    const uint32_t v = __shfl_up(inc, o);
                    if (lane >= o) inc += v;
                }
                const uint32_t excl = inc - len;
                const uint32_t total = __shfl(inc, 63);
                for (uint32_t k0 = 0; k0 < total; k0 += 64) {
                    const uint32_t k = k0 + (uint32_t)lane;
                    int o = 0;  // the largest lane whose prefix is <= k: the owner of row k
                    for (int step = 32; step > 0; step >>= 1)
                        if (__shfl(excl, o + step) <= k) o += step;
                    const uint32_t p = __shfl(run.x, o) + (k - __shfl(excl, o));
                    const float qx = __shfl(qr.x, o), qy = __shfl(qr.y, o), qz = __shfl(qr.z, o);
                    const uint32_t io = __builtin_bit_cast(uint32_t, __shfl(qr.w, o));
                    const uint64_t obase = MODE == PROBE_WRITE ? shfl_u64(base, o) + (uint64_t)__shfl(mine, o) : 0;
                    if (k < total) {
                        const float4 y = srow[p];
                        const uint32_t j = __builtin_bit_cast(uint32_t, y.w);
                        if (!(upper && j <= io)) {
                            if (MODE == PROBE_COUNT) ++cand;
                            const float qo[3] = {qx, qy, qz};
                            if (d2_row<D>(qo, y) <= T) {
                                const uint32_t slot = atomicAdd(&sCnt[wave][o], 1u);
                                if (MODE == PROBE_WRITE) {
                                    const uint64_t at = obase + slot;
                                    if (at < (uint64_t)out_cap) outj[at] = j;
                                }
                            }
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                mine += sCnt[wave][lane];
                sCnt[wave][lane] = 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        if (MODE == PROBE_COUNT && valid) cnt[i] = mine;
    }
    if (MODE == PROBE_COUNT) {
        for (int s = 32; s > 0; s >>= 1) cand += __shfl_xor(cand, s);
        if (lane == 0 && cand) atomicAdd((unsigned long long*)&stats[TJ_ST_GRID_CAND], (unsigned long long)cand);
    }
}

}  // namespace

namespace tj {


// Blocks of one bounds pass (a partial per block, no atomics): four per CU.
static int64_t bounds_blocks() { return 4 * (int64_t)num_cus(); }

GridLayout grid_layout(int64_t nR, int64_t nS, int64_t out_cap) {
    GridLayout L{};
    size_t cur = 0;
    auto carve = [&](size_t bytes) { size_t at = (cur + 255) / 256 * 256; cur = at + bytes; return at; };
    const int64_t nmax = nR > nS ? nR : nS;
    uint64_t hs = 1;
    while (hs < (uint64_t)(2 * (nS > 0 ? nS : 1))) hs <<= 1;
    L.hslots = hs;
    L.idxbits = bits_for(nmax > 1 ? nmax : 2);
    L.bounds = carve(64);
    L.bpart = carve(4 * 6 * 2 * (size_t)bounds_blocks());
    L.params = carve(sizeof(GridParams));
    L.count_nr = carve(8);
    L.count_ns = carve(8);
    L.rkeys = carve(8 * (size_t)(nR > 0 ? nR : 1));
    L.skeys = carve(8 * (size_t)(nS > 0 ? nS : 1));
    L.sort_tmp = carve(sort_temp_bytes(std::max<int64_t>(nmax, 1)));
    L.srow = carve(16 * (size_t)(nS > 0 ? nS : 1));
    L.rrow = carve(16 * (size_t)(nR > 0 ? nR : 1));
    L.hkey = carve(8 * hs);
    L.hval = carve(8 * hs);
    // occupancy bitmap: 64 bits per S point (at least 2^20); sparser grids
    // go without it
    L.bm_bits = std::max<uint64_t>(1ull << 20, 64ull * (uint64_t)(nS > 0 ? nS : 1));
    L.bitmap = carve(L.bm_bits / 8);
    L.bmpre = carve(4 * (L.bm_bits / 64));
    L.bmsum = carve(4 * (L.bm_bits / 64 / BMB + 1));
    L.runs = carve(4 * (size_t)((nS > 0 ? nS : 1) + 1));
    // output: per-row counts and offsets, the ids in row order (and a
    // ping-pong copy for the long-row merge), the list of long rows
    L.cnt = carve(4 * (size_t)(nR > 0 ? nR : 1));
    L.off = carve(8 * (size_t)(nR + 1));
    L.scan_part = carve(row_scan_part_bytes(nR));
    L.outj = carve(4 * (size_t)(out_cap > 0 ? out_cap : 1));
    L.outj2 = carve(4 * (size_t)(out_cap > 0 ? out_cap : 1));
    L.longrows = carve(row_sort_scratch_bytes(nR));  // the long rows' count and list, the row classes
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
    float4* srow = (float4*)(ws + L.srow);
    uint64_t* hkey = (uint64_t*)(ws + L.hkey);
    uint2* hval = (uint2*)(ws + L.hval);
    uint64_t* cR = (uint64_t*)(ws + L.count_nr);
    uint64_t* cS = (uint64_t*)(ws + L.count_ns);
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
    float4* rrow = self_join ? srow : (float4*)(ws + L.rrow);
    if (!self_join) {
        hipLaunchKernelGGL(keys_kernel, dim3(blocks_for((nR + 3) / 4, gb)), dim3(256), 0, s, R, nR, gp, rkeys);
        if ((rc = launch_sort(ws + L.sort_tmp, rkeys, cR, nR, 63, s, nullptr, lo_digit))) return rc;
        hipLaunchKernelGGL(rows_kernel, dim3(blocks_for(nR, gb)), dim3(256), 0, s, R, nR, gp, rkeys, rrow);
    }
    hipLaunchKernelGGL(keys_kernel, dim3(blocks_for((nS + 3) / 4, gb)), dim3(256), 0, s, S, nS, gp, skeys);
    if ((rc = launch_sort(ws + L.sort_tmp, skeys, cS, nS, 63, s, nullptr, lo_digit))) return rc;
    hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(nS, gb)), dim3(256), 0, s, S, nS, gp, skeys, srow, hkey, hval,
                       (uint64_t)(L.hslots - 1), bm);
    // rank structure of the bitmap (its kernels return at once without one)
    uint32_t* bmpre = (uint32_t*)(ws + L.bmpre);
    uint32_t* bmsum = (uint32_t*)(ws + L.bmsum);
    uint32_t* runs = (uint32_t*)(ws + L.runs);
    const unsigned nbm = (unsigned)((L.bm_bits / 64 + BMB - 1) / BMB);
    hipLaunchKernelGGL(bm_block_sums, dim3(nbm), dim3(256), 0, s, gp, bm, bmsum);
    hipLaunchKernelGGL(bm_top_scan, dim3(1), dim3(1024), 0, s, gp, bmsum);
    hipLaunchKernelGGL(bm_apply, dim3(nbm), dim3(256), 0, s, gp, bm, bmsum, bmpre);
    hipLaunchKernelGGL(runs_kernel, dim3(blocks_for(nS, gb)), dim3(256), 0, s, nS, gp, skeys, bm, bmpre, runs);
    // count pass -> offsets -> write pass -> per-row sort and emit
    uint32_t* cnt = counts ? counts : (uint32_t*)(ws + L.cnt);
    uint64_t* off = (uint64_t*)(ws + L.off);
    uint64_t* part = (uint64_t*)(ws + L.scan_part);
    uint32_t* outj = (uint32_t*)(ws + L.outj);
    uint32_t* outj2 = (uint32_t*)(ws + L.outj2);
    const unsigned pb = blocks_for(nR, (int64_t)num_cus() * 16);
    const uint64_t hmask = (uint64_t)(L.hslots - 1);
    // (one instantiation per dimension: the distance has no runtime branches)
    auto probe = [&](auto mode, uint32_t* c, const uint64_t* o, uint32_t* oj) {
        constexpr int M = decltype(mode)::value;
        auto kern = d == 1 ? probe_kernel<M, 1> : d == 2 ? probe_kernel<M, 2> : probe_kernel<M, 3>;
        hipLaunchKernelGGL(kern, dim3(pb), dim3(256), 0, s, rrow, nR, gp, srow, hkey, hval, hmask, bm, bmpre, runs, T,
                           (int)upper, c, o, oj, out_cap, stats);
    };
    probe(std::integral_constant<int, PROBE_COUNT>{}, cnt, (const uint64_t*)nullptr, (uint32_t*)nullptr);
    if ((rc = launch_row_offsets(cnt, nR, part, off, stats + TJ_ST_OUT, s))) return rc;
    probe(std::integral_constant<int, PROBE_WRITE>{}, (uint32_t*)nullptr, (const uint64_t*)off, outj);
    if ((rc = launch_row_sort(off, nR, outj, outj2, out_cap, emit, pairs, nullptr, (uint32_t*)(ws + L.longrows), s)))
        return rc;
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

}  // namespace tj
