// probe.hip -- measurement of the MFMA instructions the screen relies on.
//
// The certificate assumes |p - sum a b| <= gamma * sum |a b| + alpha for the
// FP32 accumulation inside v_mfma_f32_32x32x16_{f16,bf16} and
// v_mfma_scale_f32_32x32x64_f8f6f4, and the 32x32 output layout
// (lane l, register r) -> (row 8*(r/4) + 4*(l/32) + r%4, column l%32).
// Neither is fully specified by the ISA documentation, so both are checked
// on the device against exact dot products computed on the host (the
// reference faced the same gap and kept an explicit, tested premise:
// F16C/NUMERICAL_SCOPE.md:11-43).  Each workgroup (one wave) computes one
// 32 x 32 tile with `ksteps` chained MFMAs, exactly as the screen does.
#include "screen_impl.h"

namespace {

template <int PREC>
__global__ void __launch_bounds__(64) mfma_probe_kernel(const uint8_t* __restrict__ A, const uint8_t* __restrict__ B,
                                                        const int* __restrict__ sa, const int* __restrict__ sb,
                                                        const float* __restrict__ seed, int ksteps,
                                                        float* __restrict__ out) {
    using C = tjscreen::Cfg<PREC>;
    const int lane = threadIdx.x, hh = lane >> 5, c32 = lane & 31;
    const int64_t rowb = (int64_t)ksteps * C::KSTEP * C::ES;
    const uint8_t* a = A + ((int64_t)blockIdx.x * 32 + c32) * rowb + hh * C::CH;
    const uint8_t* b = B + ((int64_t)blockIdx.x * 32 + c32) * rowb + hh * C::CH;
    const int s_a = C::FP8 ? sa[blockIdx.x * 32 + c32] + 127 : 127;
    const int s_b = C::FP8 ? sb[blockIdx.x * 32 + c32] + 127 : 127;
    tjscreen::f32x16 acc = {};
    if (seed) {
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = seed[blockIdx.x * 32 + 8 * (r >> 2) + 4 * hh + (r & 3)];
    }
    for (int t = 0; t < ksteps; ++t) {
        tjscreen::Frag<C::CH> fa, fb;
        tjscreen::ld_frag(fa, a + 2 * t * C::CH);
        tjscreen::ld_frag(fb, b + 2 * t * C::CH);
        tjscreen::mma<PREC>(acc, fa, fb, s_a, s_b);
    }
    float* o = out + (int64_t)blockIdx.x * 1024;
#pragma unroll
    for (int r = 0; r < 16; ++r) o[(8 * (r >> 2) + 4 * hh + (r & 3)) * 32 + c32] = acc[r];
}

// Non-scaled FP8 instruction (v_mfma_f32_32x32x16_fp8_fp8, K = 16 per step),
// for comparison with the scaled f8f6f4 form; operands 8 bytes per lane.
__global__ void __launch_bounds__(64) mfma_probe_fp8_plain(const uint8_t* __restrict__ A, const uint8_t* __restrict__ B,
                                                           int ksteps, float* __restrict__ out) {
    const int lane = threadIdx.x, hh = lane >> 5, c32 = lane & 31;
    const int64_t rowb = (int64_t)ksteps * 16;
    const uint8_t* a = A + ((int64_t)blockIdx.x * 32 + c32) * rowb + hh * 8;
    const uint8_t* b = B + ((int64_t)blockIdx.x * 32 + c32) * rowb + hh * 8;
    tjscreen::f32x16 acc = {};
    for (int t = 0; t < ksteps; ++t) {
        long av = *(const long*)(a + 16 * t);
        long bv = *(const long*)(b + 16 * t);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(av, bv, acc, 0, 0, 0);
    }
    float* o = out + (int64_t)blockIdx.x * 1024;
#pragma unroll
    for (int r = 0; r < 16; ++r) o[(8 * (r >> 2) + 4 * hh + (r & 3)) * 32 + c32] = acc[r];
}

}  // namespace

extern "C" int tj_mfma_probe(int precision, const void* A, const void* B, const int* scale_a, const int* scale_b,
                             const float* seed, int ksteps, int nblocks, float* out, void* stream) {
    TJ_CHECK(A && B && out && ksteps > 0 && nblocks > 0, TJ_ERR_INVALID, "bad probe arguments");
    hipStream_t s = (hipStream_t)stream;
    const uint8_t* a = (const uint8_t*)A;
    const uint8_t* b = (const uint8_t*)B;
    switch (precision) {
        case P_FP16: hipLaunchKernelGGL(mfma_probe_kernel<P_FP16>, dim3(nblocks), dim3(64), 0, s, a, b, scale_a, scale_b, seed, ksteps, out); break;
        case P_BF16: hipLaunchKernelGGL(mfma_probe_kernel<P_BF16>, dim3(nblocks), dim3(64), 0, s, a, b, scale_a, scale_b, seed, ksteps, out); break;
        case P_FP8:
            TJ_CHECK(scale_a && scale_b, TJ_ERR_INVALID, "fp8 probe needs scales");
            hipLaunchKernelGGL(mfma_probe_kernel<P_FP8>, dim3(nblocks), dim3(64), 0, s, a, b, scale_a, scale_b, seed, ksteps, out);
            break;
        case 100:  // non-scaled fp8 instruction, ksteps of 16
            hipLaunchKernelGGL(mfma_probe_fp8_plain, dim3(nblocks), dim3(64), 0, s, a, b, ksteps, out);
            break;
        default: tj::set_error("probe: precision %d not probed", precision); return TJ_ERR_INVALID;
    }
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}
