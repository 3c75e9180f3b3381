// plain.hip -- the plain GPU method: FP64 brute force, no certificate band.
//
// Every pair is decided by the sequential binary64 difference form
// (CEJ/src/audit_g2b_accepted_fp64.py:51-57; GDS/main.cu:964 for the brute
// force loop), so its output is exact by construction; it is the
// denominator of the certified method's speedup and the full-size checker
// on the GPU.  Counterpart of the GDS brute-force kernel
// (GDS/kernel.cu:493-522) and of the FP64 terminal applied to every pair
// (TS/src/reference.py:3-47).  Compiled with -ffp-contract=off: no FMA.
//
// MI355X design: 128x128 pair tile per 256-thread workgroup, 8x8 pairs per
// thread (64 independent binary64 chains in 128 VGPRs), K staged through
// LDS in chunks of 16 coordinates already widened to binary64, so a k step
// costs 16 LDS reads for 192 FP64 operations.  Padding coordinates are
// zeros and add exactly +0 to every chain.  Workgroups are mapped
// XCD-contiguously so tiles sharing S rows meet in one L2.
#include "tj_internal.h"

namespace {

constexpr int TM = 128, KC = 16;

__device__ void plain_tile(const float* __restrict__ R, int64_t nR, const float* __restrict__ S, int64_t nS, int d,
                           int jbits, double T, int self_join, int64_t bi, int64_t bj, uint64_t* __restrict__ out_keys,
                           int64_t out_cap, uint64_t* __restrict__ stats, double (*sR)[TM], double (*sS)[TM]);

__global__ void __launch_bounds__(256) plain_kernel(const float* __restrict__ R, int64_t nR, const float* __restrict__ S,
                                                    int64_t nS, int d, int jbits, double T, int self_join,
                                                    int64_t tiles_r, int64_t tiles_s, int64_t per_xcd,
                                                    uint64_t* __restrict__ out_keys, int64_t out_cap,
                                                    uint64_t* __restrict__ stats) {
    __shared__ double sR[KC][TM];
    __shared__ double sS[KC][TM];
    // persistent: workgroup g walks tiles g, g + grid, ...; the walk order is
    // XCD-contiguous (dispatch places workgroup g on XCD g % 8)
    for (int64_t vt = blockIdx.x; vt < per_xcd * 8; vt += gridDim.x) {
        const int64_t t = (vt % 8) * per_xcd + vt / 8;
        if (t >= tiles_r * tiles_s) continue;
        const int64_t bi = t / tiles_s, bj = t % tiles_s;
        if (self_join && bj < bi) continue;
        plain_tile(R, nR, S, nS, d, jbits, T, self_join, bi, bj, out_keys, out_cap, stats, sR, sS);
    }
}

__device__ void plain_tile(const float* __restrict__ R, int64_t nR, const float* __restrict__ S, int64_t nS, int d,
                           int jbits, double T, int self_join, int64_t bi, int64_t bj, uint64_t* __restrict__ out_keys,
                           int64_t out_cap, uint64_t* __restrict__ stats, double (*sR)[TM], double (*sS)[TM]) {
    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    const int64_t r0 = bi * TM, c0 = bj * TM;

    double acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = 0.0;

    // loader: thread -> (row = tid / 2, k half = tid % 2)
    const int lrow = tid >> 1, lk = (tid & 1) * 8;
    const int64_t gr = r0 + lrow, gs = c0 + lrow;
    for (int k0 = 0; k0 < d; k0 += KC) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = k0 + lk + u;
            sR[lk + u][lrow] = (gr < nR && k < d) ? (double)R[gr * d + k] : 0.0;
            sS[lk + u][lrow] = (gs < nS && k < d) ? (double)S[gs * d + k] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < KC; ++k) {
            double x[8], y[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) x[a] = sR[k][ty * 8 + a];
#pragma unroll
            for (int b = 0; b < 8; ++b) y[b] = sS[k][tx * 8 + b];
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const double delta = x[a] - y[b];
                    acc[a][b] = acc[a][b] + delta * delta;
                }
        }
        __syncthreads();
    }
    const int lane = tid & 63;
    uint64_t mine = 0;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int64_t i = r0 + ty * 8 + a, j = c0 + tx * 8 + b;
            const bool inside = i < nR && j < nS && (!self_join || i < j) && acc[a][b] <= T;
            const uint64_t m = __ballot(inside);
            if (m) {
                const int k = __popcll(m);
                uint64_t pos = 0;
                if (lane == 0) pos = atomicAdd((unsigned long long*)&stats[TJ_ST_OUT], (unsigned long long)k);
                pos = __shfl(pos, 0);
                if (inside) {
                    const uint64_t at = pos + (uint64_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                    if (at < (uint64_t)out_cap) out_keys[at] = key_pack((uint64_t)i, (uint64_t)j, jbits);
                }
                mine += k;
            }
        }
    (void)mine;
}

}  // namespace

namespace tj {
int launch_plain(const float* R, int64_t nR, const float* S, int64_t nS, int d, int jbits, double T, bool self_join,
                 uint64_t* out_keys, int64_t out_cap, uint64_t* stats, hipStream_t s) {
    if (nR == 0 || nS == 0) return TJ_OK;
    const int64_t tr = (nR + TM - 1) / TM, ts = (nS + TM - 1) / TM;
    const int64_t tiles = tr * ts;
    const int64_t per_xcd = (tiles + 7) / 8;
    int64_t grid = per_xcd * 8;
    const int64_t maxg = (int64_t)num_cus() * 4;
    if (grid > maxg) grid = maxg;
    hipLaunchKernelGGL(plain_kernel, dim3((unsigned)grid), dim3(256), 0, s, R, nR, S, nS, d, jbits, T, (int)self_join,
                       tr, ts, per_xcd, out_keys, out_cap, stats);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}
}  // namespace tj
