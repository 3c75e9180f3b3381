// meta.hip -- row metadata: the join-unit scale, the parameter block, and
// per-row quantized copies with the norm/residual bounds of the certificate.
//
// Counterparts in the reference: half_metadata (TG/src/gpu_kernels.py:33-51:
// fp16 split z=fp16(x), r=x-z, FP64 norms, outward FP32 bounds),
// gpu_quantize_metadata (G16/src/gpu_preparation.py:12-49: scale, codes,
// FP64 residual bound, invalid flag) and gpu_norm_metadata
// (G16/src/gpu_norms.py:11-25).  Bounds are re-derived here for general
// magnitudes (the reference admitted only |x| <= 1).
#include "meta_impl.h"

namespace {

// ---------------------------------------------------------------- scale scan
// max |x| over the finite coordinates (binary32 bits; non-negative floats
// order like their bit patterns) and a count of non-finite coordinates.
// 16-byte loads over the aligned body (scalar head and tail), a block
// reduction in LDS, and one atomic per block: per-wave atomics on the one
// address serialised into the kernel's largest cost.
__global__ void __launch_bounds__(256) scale_scan_kernel(const float* __restrict__ x, int64_t count,
                                                         uint32_t* maxabs, uint64_t* stats) {
    uint32_t m = 0;
    uint32_t bad = 0;
    auto take = [&](float f) {
        const uint32_t u = __builtin_bit_cast(uint32_t, f) & 0x7FFFFFFFu;
        if (u >= 0x7F800000u) bad++;
        else m = u > m ? u : m;
    };
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t head = (int64_t)(((16 - ((uintptr_t)x & 15)) & 15) / 4);  // floats before the first 16-byte boundary
    const int64_t h = head < count ? head : count;
    const int64_t nv = (count - h) / 4;
    const float4* xv = (const float4*)(x + h);
    for (int64_t t = tid; t < nv; t += stride) {
        const float4 v = xv[t];
        take(v.x); take(v.y); take(v.z); take(v.w);
    }
    for (int64_t t = tid; t < h; t += stride) take(x[t]);
    for (int64_t t = h + 4 * nv + tid; t < count; t += stride) take(x[t]);
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t om = __shfl_xor(m, o);
        m = om > m ? om : m;
        bad += __shfl_xor(bad, o);
    }
    __shared__ uint32_t sm[4], sb[4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sm[wave] = m; sb[wave] = bad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { m = sm[w] > m ? sm[w] : m; bad += sb[w]; }
        if (m) atomicMax(maxabs, m);
        if (bad) atomicAdd((unsigned long long*)&stats[TJ_ST_INVALID], (unsigned long long)bad);
    }
}

__global__ void params_kernel(const uint32_t* maxabs, double T, int d, int dp, int prec, JoinParams* p,
                              uint64_t* stats) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const JoinParams q = tjmeta::make_params(*maxabs, T, d, dp, prec);
    *p = q;
    stats[TJ_ST_SCALE_EXP] = (uint64_t)(int64_t)q.G;
}

template <int PREC, int GW>
__global__ void __launch_bounds__(256) rowmeta_kernel(const float* __restrict__ x, int64_t n, int d,
                                                      int dp, const JoinParams* __restrict__ pp,
                                                      uint8_t* __restrict__ qhi, uint8_t* __restrict__ qlo,
                                                      RowMeta* __restrict__ meta) {
    // ROWS = 16 steps of 64 / GW rows per wave, 4 waves per block: a wave's
    // steps run one after another (each waits for its rows), so short waves
    // and many of them hide the latency (64 rows a wave ran 3-4x longer)
    constexpr int ROWS = 16 * (64 / GW);
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * ROWS;
    if (r0 >= n) return;  // wave-uniform
    const JoinParams P = *pp;
    auto src = [=](int64_t r) { return x + r * (int64_t)d; };
    tjmeta::rowmeta_wave<PREC, GW, decltype(src), ROWS>(src, n, r0, d, dp, P, lane, qhi, qlo, meta, 0);
}

__global__ void fill_u64_kernel(uint64_t* p, int64_t n, uint64_t v) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

}  // namespace

namespace tj {

int launch_scale_scan(const float* x, int64_t n, int d, uint32_t* maxabs, uint64_t* stats, hipStream_t s) {
    int64_t count = n * (int64_t)d;
    if (count == 0) return TJ_OK;
    int64_t blocks = (count + 1023) / 1024;
    const int64_t cap = 2 * (int64_t)num_cus();
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(scale_scan_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, count, maxabs, stats);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

int launch_params(const uint32_t* maxabs, double T, int d, int dp, int prec, JoinParams* p, uint64_t* stats, hipStream_t s) {
    hipLaunchKernelGGL(params_kernel, dim3(1), dim3(64), 0, s, maxabs, T, d, dp, prec, p, stats);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

int launch_rowmeta(const float* x, int64_t n, int d, int dp, int prec, const JoinParams* p, void* qhi,
                   void* qlo, RowMeta* meta, uint64_t* stats, hipStream_t s) {
    (void)stats;
    if (n == 0) return TJ_OK;
    const int gw = tjmeta::meta_width(dp);
    const int64_t rows_per_block = 4 * 16 * (64 / gw);  // the kernel's 4 waves x ROWS
    dim3 grid((unsigned)((n + rows_per_block - 1) / rows_per_block)), block(256);
    uint8_t* h = (uint8_t*)qhi;
    uint8_t* l = (uint8_t*)qlo;
#define TJ_RM(PR)                                                                                              \
    case PR:                                                                                                   \
        if (gw == 16) hipLaunchKernelGGL((rowmeta_kernel<PR, 16>), grid, block, 0, s, x, n, d, dp, p, h, l, meta); \
        else if (gw == 32) hipLaunchKernelGGL((rowmeta_kernel<PR, 32>), grid, block, 0, s, x, n, d, dp, p, h, l, meta); \
        else hipLaunchKernelGGL((rowmeta_kernel<PR, 64>), grid, block, 0, s, x, n, d, dp, p, h, l, meta);         \
        break;
    switch (prec) {
        TJ_RM(P_FP16) TJ_RM(P_FP16X3) TJ_RM(P_BF16) TJ_RM(P_BF16X3) TJ_RM(P_FP8) TJ_RM(P_FP8X3)
        default: set_error("rowmeta: unknown precision %d", prec); return TJ_ERR_INVALID;
    }
#undef TJ_RM
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

int launch_fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t s) {
    if (n <= 0) return TJ_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}

}  // namespace tj
