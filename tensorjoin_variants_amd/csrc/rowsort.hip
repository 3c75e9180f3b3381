// rowsort.hip -- output grouped by row without a global key sort: the
// exclusive scan that turns per-row counts into row offsets, and the per-row
// sort of the ids written into each row's range, which emits the (i, j) rows
// (or the bare ids of a neighbour table) in (i, j) order.  Used by the grid
// join (counts from its count pass) and by batched joins (counts of the
// screen's output keys).  Replaces the global sort of the results
// (thrust::sort_by_key, GDS/GPU.cu:775-872; output_sort, OA/src/output.cu:19-31)
// for outputs whose keys are grouped by a row: every row sorts only its own
// ids, in registers or LDS, and the rows' ranges are fixed before any id is
// written.
#include "tj_internal.h"

namespace {

__device__ inline uint64_t shfl_u64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// Exclusive scan of the row counts: off[r] = sum of cnt[0..r), off[n] = the
// total (also stored to *total when given).  Three steps: sums of blocks of
// SCB counts, one block scanning those sums, one pass writing the offsets.
constexpr int SCB = 4096;
__global__ void __launch_bounds__(256) scan_sums_kernel(const uint32_t* __restrict__ cnt, int64_t n,
                                                        uint64_t* __restrict__ part) {
    const int64_t b0 = (int64_t)blockIdx.x * SCB;
    uint64_t s = 0;
    for (int k = threadIdx.x; k < SCB; k += 256)
        if (b0 + k < n) s += cnt[b0 + k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __shared__ uint64_t w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}
__global__ void __launch_bounds__(1024) scan_top_kernel(uint64_t* __restrict__ part, int64_t nb,
                                                        uint64_t* __restrict__ off_end, uint64_t* __restrict__ total) {
    __shared__ uint64_t sp[1024];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t a = threadIdx.x * per, b = a + per < nb ? a + per : nb;
    uint64_t s = 0;
    for (int64_t t = a; t < b; ++t) s += part[t];
    sp[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint64_t v = threadIdx.x >= (unsigned)o ? sp[threadIdx.x - o] : 0;
        __syncthreads();
        sp[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = threadIdx.x ? sp[threadIdx.x - 1] : 0;
    for (int64_t t = a; t < b; ++t) {
        const uint64_t v = part[t];
        part[t] = run;
        run += v;
    }
    if (threadIdx.x == 1023) {
        *off_end = sp[1023];
        if (total) *total = sp[1023];
    }
}
__global__ void __launch_bounds__(256) scan_apply_kernel(const uint32_t* __restrict__ cnt, int64_t n,
                                                         const uint64_t* __restrict__ part, uint64_t* __restrict__ off) {
    const int64_t b0 = (int64_t)blockIdx.x * SCB;
    constexpr int PER = SCB / 256;  // consecutive counts per thread
    uint32_t v[PER];
    uint64_t s = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int64_t t = b0 + threadIdx.x * PER + u;
        v[u] = t < n ? cnt[t] : 0;
        s += v[u];
    }
    __shared__ uint64_t sp[256];
    sp[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint64_t w = threadIdx.x >= (unsigned)o ? sp[threadIdx.x - o] : 0;
        __syncthreads();
        sp[threadIdx.x] += w;
        __syncthreads();
    }
    uint64_t run = part[blockIdx.x] + (threadIdx.x ? sp[threadIdx.x - 1] : 0);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int64_t t = b0 + threadIdx.x * PER + u;
        if (t < n) off[t] = run;
        run += v[u];
    }
}

// Emit the sorted ids of row i at output positions [o0, o0 + n): (i, j) rows
// of int32 / int64, or the bare ids of the neighbour table.  The i written
// is ilocal[i] when a table of local row numbers is given (joins batched in
// one row space), else i.
__device__ inline void emit_one(int emit, void* out, uint64_t at, uint32_t i, uint32_t j) {
    if (emit == tj::GRID_EMIT_TABLE) {
        ((uint32_t*)out)[at] = j;
    } else if (emit == tj::GRID_EMIT_I32) {
        ((int2*)out)[at] = make_int2((int)i, (int)j);
    } else {
        ((longlong2*)out)[at] = make_longlong2((long long)i, (long long)j);
    }
}

// Per-row sort of the neighbour ids (distinct within a row) + emit, by row
// length: rows of at most LANE_MAX ids are sorted by their own lane in
// registers (segsort_lane_kernel); rows of up to P ids, P a power of two from
// 16 to 512, by a group of G lanes holding E = P / G ids each
// (midsort_kernel<G, E>: 16 lanes up to 128 ids, then 32 and 64, so no lane
// holds more than 8); longer rows are listed for longsort_kernel.  Both sorts are
// bitonic networks unrolled at compile time: compare-exchanges between the
// registers of a lane, or between lanes of a group (shuffles), no LDS and no
// barriers; partners within 16 lanes come by DPP moves.  Ids are distinct
// within a row, so any correct sort gives the same order.
constexpr int LANE_MAX = 8;
constexpr int GROUP_MAX = 512;  // 16 lanes x 32 ids

__device__ __forceinline__ void cx(uint32_t& a, uint32_t& b, bool asc) {
    const uint32_t lo = min(a, b), hi = max(a, b);
    a = asc ? lo : hi;
    b = asc ? hi : lo;
}

// Class code of a row length, stored per row by segsort_lane_kernel for the
// later kernels: 0 for rows of at most LANE_MAX ids (sorted by their lane),
// log2(P) - 3 for the rows of midsort_kernel with P = G * E (16 .. GROUP_MAX),
// LONG_CODE above.
constexpr int LONG_CODE = 7;
__device__ inline uint8_t row_code(uint32_t n) {
    if (n <= (uint32_t)LANE_MAX) return 0;
    if (n > (uint32_t)GROUP_MAX) return (uint8_t)LONG_CODE;
    int c = 1, p = 16;
    while ((uint32_t)p < n) { p <<= 1; ++c; }
    return (uint8_t)c;
}
template <int P>
constexpr int code_of() { return P == 16 ? 1 : P == 32 ? 2 : P == 64 ? 3 : P == 128 ? 4 : P == 256 ? 5 : 6; }

__global__ void __launch_bounds__(256) segsort_lane_kernel(const uint64_t* __restrict__ off, int64_t nR,
                                                           const uint32_t* __restrict__ outj, int64_t out_cap, int emit,
                                                           void* __restrict__ out, const int32_t* __restrict__ ilocal,
                                                           uint32_t* __restrict__ longrows,
                                                           uint32_t* __restrict__ nlong, uint8_t* __restrict__ codes,
                                                           const uint64_t* __restrict__ src) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nR; i += stride) {
        uint64_t o0 = off[i], o1 = off[i + 1];
        if (o1 > (uint64_t)out_cap) o1 = (uint64_t)out_cap;  // overflowed run: skip what has no room
        if (o0 > o1) o0 = o1;
        const uint32_t n = (uint32_t)(o1 - o0);
        const uint8_t code = row_code(n);
        codes[i] = code;
        if (code == LONG_CODE) longrows[atomicAdd(nlong, 1u)] = (uint32_t)i;
        if (n > (uint32_t)LANE_MAX || n == 0) continue;
        const uint64_t r0 = src ? src[i] : o0;
        uint32_t v[LANE_MAX];
#pragma unroll
        for (int k = 0; k < LANE_MAX; ++k) v[k] = (uint32_t)k < n ? outj[r0 + k] : 0xFFFFFFFFu;
#pragma unroll
        for (int kk = 2; kk <= LANE_MAX; kk <<= 1)
#pragma unroll
            for (int jj = kk >> 1; jj > 0; jj >>= 1)
#pragma unroll
                for (int x = 0; x < LANE_MAX; ++x)
                    if ((x ^ jj) > x) cx(v[x], v[x ^ jj], (x & kk) == 0);
        const uint32_t io = ilocal ? (uint32_t)ilocal[i] : (uint32_t)i;
#pragma unroll
        for (int k = 0; k < LANE_MAX; ++k)
            if ((uint32_t)k < n) emit_one(emit, out, o0 + k, io, v[k]);
    }
}

// Rows of class P = G * E ids: each wave scans 64 rows at a time and sorts
// the ones of its class 64 / G at a time, one group of G lanes per row.  Lane
// l of a group holds ids x = G e + l (e < E) of its row: steps with a
// distance of at least G are between registers of a lane, shorter ones
// between lanes (shuffles within the group).
template <int G, int E>
__global__ void __launch_bounds__(256) midsort_kernel(const uint64_t* __restrict__ off, int64_t nR,
                                                      const uint32_t* __restrict__ outj, int64_t out_cap, int emit,
                                                      void* __restrict__ out, const int32_t* __restrict__ ilocal,
                                                      const uint8_t* __restrict__ codes, const uint64_t* __restrict__ src) {
    constexpr int P = G * E;
    constexpr int NG = 64 / G;  // rows per wave step
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane / G, l = lane % G;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t tb = (int64_t)blockIdx.x * blockDim.x + wave * 64; tb < nR; tb += stride) {
        const int64_t i = tb + lane;
        uint64_t mine = __ballot(i < nR && codes[i] == code_of<P>());
        while (mine) {
            // the next NG rows of the class, one per group
            int row = -1;
#pragma unroll
            for (int q = 0; q < NG; ++q) {
                const int b = mine ? __builtin_ctzll(mine) : -1;
                if (b >= 0) mine &= mine - 1;
                if (q == g) row = b;
            }
            const int s = row < 0 ? 0 : row;  // tb + s < nR
            uint64_t r0 = 0, w0 = 0;
            uint32_t rn = 0;
            if (row >= 0) {
                uint64_t o0 = off[tb + s], o1 = off[tb + s + 1];
                if (o1 > (uint64_t)out_cap) o1 = (uint64_t)out_cap;
                if (o0 > o1) o0 = o1;
                r0 = o0;
                w0 = src ? src[tb + s] : o0;
                rn = (uint32_t)(o1 - o0);
            }
            uint32_t v[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t x = (uint32_t)(G * e + l);
                v[e] = x < rn ? outj[w0 + x] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int kk = 2; kk <= P; kk <<= 1)
#pragma unroll
                for (int jj = kk >> 1; jj > 0; jj >>= 1) {
                    if (jj >= G) {
                        // x and x ^ jj differ in e; kk > jj >= G, so the
                        // direction depends on e only
#pragma unroll
                        for (int e = 0; e < E; ++e) {
                            const int f = e ^ (jj / G);
                            if (f > e) cx(v[e], v[f], ((G * e) & kk) == 0);
                        }
                    } else {
                        // keep the low value when this lane is the lower index
                        // of an ascending pair (or the upper of a descending one)
#pragma unroll
                        for (int e = 0; e < E; ++e) {
                            const uint32_t p = tj::xor_lane(v[e], jj, lane);
                            const int x = G * e + l;
                            const bool asc = (x & kk) == 0;
                            const bool lower = (l & jj) == 0;
                            v[e] = (asc == lower) ? min(v[e], p) : max(v[e], p);
                        }
                    }
                }
            const uint32_t ri = ilocal ? (uint32_t)ilocal[tb + s] : (uint32_t)(tb + s);  // tb + s < nR
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t x = (uint32_t)(G * e + l);
                if (x < rn) emit_one(emit, out, r0 + x, ri, v[e]);
            }
        }
    }
}

// Rows longer than GROUP_MAX (dense clusters, large eps): one workgroup per
// row.  Chunks of CHUNK ids are sorted in LDS (bitonic), then sorted runs are
// merged pairwise (merge path: each thread finds where its slice of the
// output splits the two runs by a binary search, then merges sequentially),
// ping-ponging between outj and outj2 over the row's own range, until one run
// remains; then emitted.  The list of long rows and its length come from
// segsort_lane_kernel through device memory.
constexpr int CHUNK = 4096;
__global__ void __launch_bounds__(256) longsort_kernel(const uint64_t* __restrict__ off, const uint32_t* __restrict__ longrows,
                                                       const uint32_t* __restrict__ nlong, uint32_t* __restrict__ outj,
                                                       uint32_t* __restrict__ outj2, int64_t out_cap, int emit,
                                                       void* __restrict__ out, const int32_t* __restrict__ ilocal,
                                                       const uint64_t* __restrict__ rsrc) {
    __shared__ uint32_t S[CHUNK];
    const uint32_t count = *nlong;
    for (uint32_t r = blockIdx.x; r < count; r += gridDim.x) {
        const uint32_t i = longrows[r];
        uint64_t o0 = off[i], o1 = off[i + 1];
        if (o1 > (uint64_t)out_cap) o1 = (uint64_t)out_cap;
        if (o0 > o1) o0 = o1;
        const uint64_t n = o1 - o0;
        // the row's ids ping-pong between its segment of outj and its final
        // range of outj2
        uint32_t* src = outj + (rsrc ? rsrc[i] : o0);
        uint32_t* dst = outj2 + o0;
        // sorted chunks; a row shorter than CHUNK runs the network of the
        // next power of two at or above its length
        uint32_t net = 2 * GROUP_MAX;
        while ((uint64_t)net < n && net < (uint32_t)CHUNK) net <<= 1;
        for (uint64_t c0 = 0; c0 < n; c0 += CHUNK) {
            const uint32_t cn = (uint32_t)(n - c0 < (uint64_t)CHUNK ? n - c0 : (uint64_t)CHUNK);
            for (uint32_t k = threadIdx.x; k < net; k += 256) S[k] = k < cn ? src[c0 + k] : 0xFFFFFFFFu;
            for (uint32_t kk = 2; kk <= net; kk <<= 1)
                for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
                    __syncthreads();
                    for (uint32_t x = threadIdx.x; x < net; x += 256) {
                        const uint32_t y = x ^ jj;
                        if (y > x) {
                            const uint32_t a = S[x], b = S[y];
                            const bool up = (x & kk) == 0;
                            if ((a > b) == up) {
                                S[x] = b;
                                S[y] = a;
                            }
                        }
                    }
                }
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < cn; k += 256) src[c0 + k] = S[k];
            __syncthreads();
        }
        // pairwise merges of runs of width w
        for (uint64_t w = CHUNK; w < n; w <<= 1) {
            for (uint64_t a0 = 0; a0 < n; a0 += 2 * w) {
                const uint64_t a1 = a0 + w < n ? a0 + w : n;
                const uint64_t b1 = a0 + 2 * w < n ? a0 + 2 * w : n;
                const uint64_t na = a1 - a0, nb = b1 - a1, tot = na + nb;
                const uint64_t per = (tot + 255) / 256;
                const uint64_t d0 = (uint64_t)threadIdx.x * per;
                if (d0 < tot) {
                    const uint64_t d1 = d0 + per < tot ? d0 + per : tot;
                    // split of the first d0 outputs: ia from A, d0 - ia from B
                    uint64_t lo = d0 > nb ? d0 - nb : 0, hi = d0 < na ? d0 : na;
                    while (lo < hi) {
                        const uint64_t m = (lo + hi) >> 1;
                        if (src[a0 + m] < src[a1 + (d0 - m - 1)]) lo = m + 1; else hi = m;
                    }
                    uint64_t ia = lo, ib = d0 - lo;
                    for (uint64_t dd = d0; dd < d1; ++dd) {
                        const bool takeA = ib >= nb || (ia < na && src[a0 + ia] < src[a1 + ib]);
                        dst[a0 + dd] = takeA ? src[a0 + ia++] : src[a1 + ib++];
                    }
                }
            }
            __syncthreads();
            uint32_t* t = src;
            src = dst;
            dst = t;
        }
        const uint32_t io = ilocal ? (uint32_t)ilocal[i] : i;
        for (uint64_t k = threadIdx.x; k < n; k += 256) emit_one(emit, out, o0 + k, io, src[k]);
        __syncthreads();
    }
}

}  // namespace

namespace tj {

static unsigned row_blocks(int64_t n, int64_t cap) {
    int64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

size_t row_scan_part_bytes(int64_t n) { return 8 * (size_t)((n + SCB - 1) / SCB + 1); }
// the long-row count and list (n + 2 words), then a class code per row
static size_t codes_at(int64_t n) { return (4 * (size_t)(n + 2) + 255) / 256 * 256; }
size_t row_sort_scratch_bytes(int64_t n) { return codes_at(n) + (size_t)(n > 0 ? n : 1); }

int launch_row_offsets(const uint32_t* cnt, int64_t n, uint64_t* part, uint64_t* off, uint64_t* total,
                       hipStream_t s) {
    const int64_t nsb = (n + SCB - 1) / SCB;
    if (nsb > 0) hipLaunchKernelGGL(scan_sums_kernel, dim3((unsigned)nsb), dim3(256), 0, s, cnt, n, part);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(1024), 0, s, part, nsb, off + n, total);
    if (nsb > 0) hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)nsb), dim3(256), 0, s, cnt, n, part, off);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

int launch_row_sort(const uint64_t* off, int64_t n, uint32_t* outj, uint32_t* outj2, int64_t out_cap, int emit,
                    void* out, const int32_t* ilocal, uint32_t* longlist, const uint64_t* src, hipStream_t s) {
    if (n <= 0) return TJ_OK;
    uint32_t* nlong = longlist;
    uint32_t* longrows = longlist + 1;
    uint8_t* codes = (uint8_t*)longlist + codes_at(n);
    int rc;
    if ((rc = launch_fill_u64((uint64_t*)nlong, 1, 0, s))) return rc;  // (clears the count and its neighbour word)
    const dim3 sg(row_blocks(n, (int64_t)num_cus() * 8));
    hipLaunchKernelGGL(segsort_lane_kernel, sg, dim3(256), 0, s, off, n, outj, out_cap, emit, out, ilocal, longrows,
                       nlong, codes, src);
    hipLaunchKernelGGL((midsort_kernel<16, 1>), sg, dim3(256), 0, s, off, n, outj, out_cap, emit, out, ilocal,
                       (const uint8_t*)codes, src);
    hipLaunchKernelGGL((midsort_kernel<16, 2>), sg, dim3(256), 0, s, off, n, outj, out_cap, emit, out, ilocal,
                       (const uint8_t*)codes, src);
    hipLaunchKernelGGL((midsort_kernel<16, 4>), sg, dim3(256), 0, s, off, n, outj, out_cap, emit, out, ilocal,
                       (const uint8_t*)codes, src);
    hipLaunchKernelGGL((midsort_kernel<16, 8>), sg, dim3(256), 0, s, off, n, outj, out_cap, emit, out, ilocal,
                       (const uint8_t*)codes, src);
    hipLaunchKernelGGL((midsort_kernel<32, 8>), sg, dim3(256), 0, s, off, n, outj, out_cap, emit, out, ilocal,
                       (const uint8_t*)codes, src);
    hipLaunchKernelGGL((midsort_kernel<64, 8>), sg, dim3(256), 0, s, off, n, outj, out_cap, emit, out, ilocal,
                       (const uint8_t*)codes, src);
    hipLaunchKernelGGL(longsort_kernel, dim3((unsigned)num_cus()), dim3(256), 0, s, off, longrows, nlong, outj, outj2,
                       out_cap, emit, out, ilocal, src);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

}  // namespace tj
