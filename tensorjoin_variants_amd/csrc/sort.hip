// sort.hip -- deterministic final pair order.
//
// Counterpart of output_mirror + output_sort (OA/src/output.cu:8-31: mirror
// upper IDs, CUB 64-bit radix sort with a two-phase scratch query) and of
// the host canonical() sort (TG/src/retained_ops.py:17-19).  The pair keys
// are (i << jbits) | j, so ascending keys are ascending (i, j); since every
// key is distinct the sorted output is unique whatever order the screen and
// refinement appended pairs in.
//
// MI355X design: LSD radix sort over only the bits the key range needs
// (8 bits per pass), written here (no rocPRIM): per-block digit histograms,
// a two-step exclusive scan (one workgroup per digit row, then the 256 digit
// totals), and a stable scatter in which each wave
// ranks its 64 keys by an 8-ballot multisplit and the 4 waves of a block
// combine through LDS.  The element count is read from device memory, so
// the whole sort is capturable in a hipGraph behind a producer whose output
// size is unknown at launch.
#include "tj_internal.h"

namespace {

constexpr int SB = 256;     // threads per block = keys per tile
constexpr int MAXNB = 1024; // histogram blocks

__device__ inline int64_t clamp_count(const uint64_t* count, int64_t cap) {
    int64_t n = (int64_t)*count;
    return n < cap ? n : cap;
}

__global__ void __launch_bounds__(SB) hist_kernel(const uint64_t* __restrict__ keys, const uint64_t* count, int64_t cap,
                                                  int shift, uint32_t* __restrict__ hist, int nb) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t n = clamp_count(count, cap);
    const int64_t chunk = ((n + nb - 1) / nb + SB - 1) / SB * SB;
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    const int64_t b1 = b0 + chunk < n ? b0 + chunk : n;
    // eight keys per thread in flight (one load at a time left the pass
    // waiting on memory latency)
    constexpr int U = 8;
    int64_t t = b0 + threadIdx.x;
    for (; t + (int64_t)(U - 1) * SB < b1; t += (int64_t)U * SB) {
        uint32_t dg[U];
#pragma unroll
        for (int u = 0; u < U; ++u) dg[u] = (uint32_t)(keys[t + (int64_t)u * SB] >> shift) & 255u;
#pragma unroll
        for (int u = 0; u < U; ++u) atomicAdd(&h[dg[u]], 1u);
    }
    for (; t < b1; t += SB) atomicAdd(&h[(keys[t] >> shift) & 255u], 1u);
    __syncthreads();
    hist[(int64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the 256 x nb digit-major counters in two small steps:
// one workgroup per digit scans its row (block offsets within the digit) and
// records the digit total; one workgroup turns the totals into digit bases,
// added to every row entry by the scatter.
__global__ void __launch_bounds__(256) scan_rows_kernel(uint32_t* hist, int nb, uint32_t* tot) {
    __shared__ uint32_t part[256];
    uint32_t* row = hist + (int64_t)blockIdx.x * nb;
    const int per = (nb + 255) / 256;
    const int a = threadIdx.x * per;
    const int b = a + per < nb ? a + per : nb;
    uint32_t s = 0;
    for (int t = a; t < b; ++t) s += row[t];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int t = a; t < b; ++t) {
        const uint32_t v = row[t];
        row[t] = run;
        run += v;
    }
    if (threadIdx.x == 255) tot[blockIdx.x] = part[255];
}

__global__ void __launch_bounds__(256) scan_totals_kernel(uint32_t* tot) {
    __shared__ uint32_t part[256];
    part[threadIdx.x] = tot[threadIdx.x];
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    tot[256 + threadIdx.x] = threadIdx.x ? part[threadIdx.x - 1] : 0;  // digit bases
}

__device__ inline int lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Stable scatter of one block's chunk, in tiles of TILE = SB x KPT keys.
// Tile order is key order (striped: key base + k SB + tid is slot k of
// thread tid).  Per slot, each wave ranks its 64 keys among equal digits by
// an 8-ballot multisplit; one pass over (slot, wave) per digit turns the
// per-wave counts into offsets; the tile is counting-sorted into LDS and
// written out as contiguous runs per digit (coalesced stores).
constexpr int KPT = 8;
constexpr int TILE = SB * KPT;

__global__ void __launch_bounds__(SB) scatter_kernel(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst,
                                                     const uint64_t* count, int64_t cap, int shift,
                                                     const uint32_t* __restrict__ offs, const uint32_t* __restrict__ bases,
                                                     int nb) {
    __shared__ uint32_t run[256];         // next global position per digit
    __shared__ uint32_t tstart[256];      // first tile position per digit
    __shared__ uint32_t ttot[256];        // keys per digit in the tile
    __shared__ uint32_t wsum[4];          // per-wave sums of the digit-total scan
    // per (slot, wave) digit counts, then offsets within the tile (<= TILE:
    // 16 bits; the smaller LDS footprint keeps four blocks per CU resident)
    __shared__ uint16_t wcnt[KPT][4][256];
    __shared__ uint64_t sorted[TILE];
    const int tid = threadIdx.x, wave = tid >> 6;
    const int64_t n = clamp_count(count, cap);
    const int64_t chunk = ((n + nb - 1) / nb + SB - 1) / SB * SB;
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    const int64_t b1 = b0 + chunk < n ? b0 + chunk : n;
    run[tid] = bases[tid] + offs[(int64_t)tid * nb + blockIdx.x];
    for (int64_t base = b0; base < b1; base += TILE) {
        for (int e = tid; e < KPT * 4 * 256 / 2; e += SB) ((uint32_t*)&wcnt[0][0][0])[e] = 0;
        __syncthreads();
        uint64_t key[KPT];
        uint32_t dg[KPT];
        int rank[KPT];
        bool valid[KPT];
#pragma unroll
        for (int k = 0; k < KPT; ++k) {
            const int64_t t = base + (int64_t)k * SB + tid;
            valid[k] = t < b1;
            key[k] = valid[k] ? src[t] : 0;
            dg[k] = (uint32_t)(key[k] >> shift) & 255u;
        }
#pragma unroll
        for (int k = 0; k < KPT; ++k) {
            uint64_t peers = __ballot(valid[k]);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (dg[k] >> b) & 1u;
                const uint64_t bm = __ballot(bit);
                peers &= bit ? bm : ~bm;
            }
            rank[k] = lanes_below(peers);
            if (valid[k] && rank[k] == 0) wcnt[k][wave][dg[k]] = (uint16_t)__popcll(peers);
        }
        __syncthreads();
        // per digit: exclusive offsets over (slot, wave) in tile order, then
        // the exclusive scan of the tile's digit totals: each wave scans its
        // 64 digits in registers, one barrier joins the four wave sums
        {
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < KPT; ++k)
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const uint32_t c = wcnt[k][w][tid];
                    wcnt[k][w][tid] = (uint16_t)acc;
                    acc += c;
                }
            ttot[tid] = acc;
            uint32_t v = acc;
            const int lane = tid & 63;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t u = __shfl_up(v, o);
                if (lane >= o) v += u;
            }
            if (lane == 63) wsum[wave] = v;
            __syncthreads();
            uint32_t add = 0;
            for (int w = 0; w < wave; ++w) add += wsum[w];
            tstart[tid] = add + v - ttot[tid];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KPT; ++k)
            if (valid[k]) sorted[tstart[dg[k]] + wcnt[k][wave][dg[k]] + (uint32_t)rank[k]] = key[k];
        __syncthreads();
        const int64_t m = b1 - base < TILE ? b1 - base : TILE;
        for (int idx = tid; idx < m; idx += SB) {
            const uint64_t kk = sorted[idx];
            const uint32_t d = (uint32_t)(kk >> shift) & 255u;
            dst[run[d] + (uint32_t)idx - tstart[d]] = kk;
        }
        __syncthreads();
        run[tid] += ttot[tid];
        __syncthreads();
    }
}

__global__ void copy_kernel(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst, const uint64_t* count, int64_t cap) {
    const int64_t n = clamp_count(count, cap);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) dst[t] = src[t];
}

__global__ void unpack_kernel(const uint64_t* __restrict__ keys, const uint64_t* count, int64_t cap, int jbits,
                              void* __restrict__ pairs, int fmt) {
    const int64_t n = clamp_count(count, cap);
    const uint64_t jmask = (1ull << jbits) - 1;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[t];
        if (fmt == TJ_OUT_I32) ((int2*)pairs)[t] = make_int2((int)(k >> jbits), (int)(k & jmask));
        else ((longlong2*)pairs)[t] = make_longlong2((long long)(k >> jbits), (long long)(k & jmask));
    }
}

// upper pairs (i < j) of an n-point self-join -> i*n+j, j*n+i, plus the n diagonal ids
__global__ void mirror_kernel(const int64_t* __restrict__ pairs, const uint64_t* count, int64_t cap, int64_t n,
                              uint64_t* __restrict__ ids, uint64_t* total) {
    const int64_t m = clamp_count(count, cap);
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt == 0) *total = (uint64_t)(2 * m + n);
    for (int64_t t = gt; t < m; t += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t i = (uint64_t)pairs[2 * t], j = (uint64_t)pairs[2 * t + 1];
        ids[2 * t] = i * (uint64_t)n + j;
        ids[2 * t + 1] = j * (uint64_t)n + i;
    }
    for (int64_t t = gt; t < n; t += (int64_t)gridDim.x * blockDim.x) ids[2 * m + t] = (uint64_t)t * (uint64_t)n + (uint64_t)t;
}

int nblocks_for(int64_t cap) {
    int64_t nb = (cap + 8 * SB - 1) / (8 * SB);
    if (nb < 1) nb = 1;
    if (nb > MAXNB) nb = MAXNB;
    return (int)nb;
}

}  // namespace

namespace tj {

size_t sort_temp_bytes(int64_t capacity) {
    size_t a = ((size_t)(capacity > 0 ? capacity : 1) * 8 + 255) / 256 * 256;
    return a + (size_t)256 * nblocks_for(capacity) * 4 + 512 * 4 + 256;
}

int launch_sort(void* temp, uint64_t* keys, const uint64_t* count, int64_t capacity, int bits, hipStream_t s,
                uint64_t** result, int lo_bit) {
    if (result) *result = keys;
    if (capacity <= 0 || bits <= 0) return TJ_OK;
    const int nb = nblocks_for(capacity);
    uint64_t* alt = (uint64_t*)temp;
    uint32_t* hist = (uint32_t*)((uint8_t*)temp + ((size_t)capacity * 8 + 255) / 256 * 256);
    uint32_t* tot = hist + (size_t)256 * nb;  // [0, 256): digit totals, [256, 512): digit bases
    const int passes = (bits + 7) / 8;
    uint64_t* src = keys;
    uint64_t* dst = alt;
    for (int p = lo_bit / 8; p < passes; ++p) {
        hipLaunchKernelGGL(hist_kernel, dim3(nb), dim3(SB), 0, s, src, count, capacity, 8 * p, hist, nb);
        hipLaunchKernelGGL(scan_rows_kernel, dim3(256), dim3(256), 0, s, hist, nb, tot);
        hipLaunchKernelGGL(scan_totals_kernel, dim3(1), dim3(256), 0, s, tot);
        hipLaunchKernelGGL(scatter_kernel, dim3(nb), dim3(SB), 0, s, src, dst, count, capacity, 8 * p, hist, tot + 256, nb);
        uint64_t* t = src; src = dst; dst = t;
    }
    if (result) {
        *result = src;
    } else if (src != keys) {
        int64_t blocks = (capacity + 255) / 256;
        if (blocks > 2048) blocks = 2048;
        hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, keys, count, capacity);
    }
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

int launch_unpack(const uint64_t* keys, const uint64_t* count, int64_t cap, int jbits, void* pairs, int out_format,
                  hipStream_t s) {
    if (cap <= 0) return TJ_OK;
    int64_t blocks = (cap + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(unpack_kernel, dim3((unsigned)blocks), dim3(256), 0, s, keys, count, cap, jbits, pairs, out_format);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

int launch_mirror(const int64_t* pairs, const uint64_t* count, int64_t cap, int64_t n, uint64_t* ids, uint64_t* total,
                  hipStream_t s) {
    int64_t work = cap > n ? cap : n;
    int64_t blocks = (work + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(mirror_kernel, dim3((unsigned)blocks), dim3(256), 0, s, pairs, count, cap, n, ids, total);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

}  // namespace tj
