// refine.hip -- FP64 refinement of the ambiguous band.
//
// Counterpart of refine_ambiguous_fp64_i64 (CEJ/src/run_g2a_tensorjoin.py:193-232):
// one queued pair id -> widen both FP32 rows to FP64, difference form,
// append if d2 <= T.  The summation order here is the sequential one of
// CEJ/src/audit_g2b_accepted_fp64.py:51-57 (k ascending, fl(s + fl(delta^2)),
// no FMA: this file is compiled with -ffp-contract=off), so the result is
// bit-identical to the CPU oracle.
//
// MI355X design: a fixed grid of persistent waves strides over the queue
// (its length is read from device memory, so the launch can sit in a
// hipGraph); one lane per pair keeps the sequential dependence chain while
// the wave covers 64 pairs; 16-byte row loads; accepted pairs compacted with
// a wave ballot and one atomic per wave.  Besides the decision, each pair
// reports |d2_screen - d2_fp64| against the screen's certified half-width
// (the screen-error statistics and a soundness check of the certificate).
#include "tj_internal.h"

namespace {

__device__ inline void atomic_max_pos_double(uint64_t* slot, double v) {
    // non-negative binary64 values order like their bit patterns
    if (v > 0.0) atomicMax((unsigned long long*)slot, (unsigned long long)__builtin_bit_cast(uint64_t, v));
}

__global__ void __launch_bounds__(256) refine_kernel(const float* __restrict__ R, const float* __restrict__ S, int d,
                                                     int jbits, const uint64_t* __restrict__ amb_keys,
                                                     const float2* __restrict__ amb_aux, int64_t amb_cap,
                                                     const JoinParams* __restrict__ pp, uint64_t* __restrict__ out_keys,
                                                     int64_t out_cap, uint64_t* __restrict__ stats) {
    const JoinParams P = *pp;
    int64_t count = (int64_t)stats[TJ_ST_AMB];
    if (count > amb_cap) count = amb_cap;
    const int lane = threadIdx.x & 63;
    const uint64_t jmask = (jbits >= 64) ? ~0ull : ((1ull << jbits) - 1);
    double max_err = 0.0, max_ratio = 0.0;
    uint64_t viol = 0, acc_cnt = 0;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (int64_t base = w * 64; base < count; base += nwaves * 64) {
        const int64_t t = base + lane;
        bool inside = false;
        uint64_t key = 0;
        if (t < count) {
            uint64_t raw = amb_keys[t];
            const bool certified = (raw >> 63) != 0;  // audit mode: certified inside by the screen
            key = raw & ~(1ull << 63);
            const uint64_t i = key >> jbits, j = key & jmask;
            const double d2 = d2_seq(R + i * (uint64_t)d, S + j * (uint64_t)d, d);
            inside = d2 <= P.T;
            const float2 aux = amb_aux[t];
            const double dj = d2 / P.scale_real;  // exact power-of-two scaling
            const double err = fabs(dj - (double)aux.x);
            max_err = fmax(max_err, err * P.scale_real);
            if (aux.y > 0.0f) max_ratio = fmax(max_ratio, err / (double)aux.y);
            // the certified half-width covers the exact d2; allow the binary64
            // gap g and the float rounding of the centre
            const bool bad = (certified && !inside) ||
                             err > (double)aux.y + 2.0 * P.g64 * dj + 0x1p-23 * fabs((double)aux.x);
            if (bad) {
                viol++;
                stats[TJ_ST_VIOL_KEY] = raw;
                stats[TJ_ST_VIOL_D2] = __builtin_bit_cast(uint64_t, dj);
                stats[TJ_ST_VIOL_EST] = (uint64_t)__builtin_bit_cast(uint32_t, aux.x) |
                                        ((uint64_t)__builtin_bit_cast(uint32_t, aux.y) << 32);
            }
        }
        const uint64_t m = __ballot(inside);
        if (m) {
            const int k = __popcll(m);
            uint64_t pos = 0;
            if (lane == 0) pos = atomicAdd((unsigned long long*)&stats[TJ_ST_OUT], (unsigned long long)k);
            pos = __shfl(pos, 0);
            if (inside) {
                const uint64_t at = pos + (uint64_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                if (at < (uint64_t)out_cap) out_keys[at] = key;
            }
            acc_cnt += (lane == 0) ? (uint64_t)k : 0;
        }
    }
    // wave reductions of the statistics
    for (int o = 32; o > 0; o >>= 1) {
        max_err = fmax(max_err, __shfl_xor(max_err, o));
        max_ratio = fmax(max_ratio, __shfl_xor(max_ratio, o));
        viol += __shfl_xor(viol, o);
    }
    if (lane == 0) {
        atomic_max_pos_double(&stats[TJ_ST_MAX_ERR], max_err);
        atomic_max_pos_double(&stats[TJ_ST_MAX_RATIO], max_ratio);
        if (viol) atomicAdd((unsigned long long*)&stats[TJ_ST_VIOLATIONS], (unsigned long long)viol);
        if (acc_cnt) atomicAdd((unsigned long long*)&stats[TJ_ST_FP64_ACCEPT], (unsigned long long)acc_cnt);
    }
}

}  // namespace

namespace tj {
int launch_refine(const float* R, const float* S, int d, int jbits, const uint64_t* amb_keys, const float* amb_aux,
                  int64_t amb_cap, const JoinParams* p, uint64_t* out_keys, int64_t out_cap, uint64_t* stats,
                  hipStream_t s) {
    if (amb_cap <= 0) return TJ_OK;
    int64_t blocks = (amb_cap + 255) / 256;
    int64_t maxb = (int64_t)num_cus() * 8;
    if (blocks > maxb) blocks = maxb;
    hipLaunchKernelGGL(refine_kernel, dim3((unsigned)blocks), dim3(256), 0, s, R, S, d, jbits, amb_keys,
                       (const float2*)amb_aux, amb_cap, p, out_keys, out_cap, stats);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}
}  // namespace tj
