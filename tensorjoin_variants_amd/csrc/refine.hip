// refine.hip -- FP64 refinement of a queue of pair keys.
//
// Counterpart of refine_ambiguous_fp64_i64 (CEJ/src/run_g2a_tensorjoin.py:193-232):
// queued pair ids in -> widen both FP32 rows to FP64, difference form,
// accepted ids out when d2 <= T.  The summation order here is the
// sequential one of CEJ/src/audit_g2b_accepted_fp64.py:51-57 (k ascending,
// fl(s + fl(delta^2)), no FMA: this file is compiled with -ffp-contract=off),
// so the decision is bit-identical to the CPU oracle's.
//
// Queue: keys i << jbits | j (bit 63, the screen's accept flag, is ignored),
// its length read from a device scalar (clamped to the capacity) so that the
// launch needs no host round trip and can sit in a hipGraph.  The join
// pipeline refines inside its certify pass (certify_impl.h); this stage
// serves callers that hold their own queue -- the screen's candidate queue
// (count at stats[TJ_ST_CANDIDATES]) among them.
//
// MI355X design: a fixed grid of waves strides over the queue; one lane per
// pair keeps the sequential dependence chain while the wave covers 64 pairs;
// 16-byte row loads; accepted keys compacted with a wave ballot and one
// output atomic per wave step.
#include "tj_internal.h"

namespace {

__global__ void __launch_bounds__(256) refine_kernel(const float* __restrict__ R, const float* __restrict__ S, int d,
                                                     int jbits, double T, const uint64_t* __restrict__ keys,
                                                     const uint64_t* __restrict__ count_p, int64_t cap,
                                                     uint64_t* __restrict__ out_keys, int64_t out_cap,
                                                     uint64_t* __restrict__ stats) {
    int64_t count = (int64_t)*count_p;
    if (count > cap) count = cap;
    const int lane = threadIdx.x & 63;
    const uint64_t jmask = (jbits >= 63) ? ((1ull << 63) - 1) : ((1ull << jbits) - 1);
    uint64_t acc_cnt = 0, done = 0;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (int64_t base = w * 64; base < count; base += nwaves * 64) {
        const int64_t t = base + lane;
        bool inside = false;
        uint64_t key = 0;
        if (t < count) {
            key = keys[t] & ~(1ull << 63);
            const uint64_t i = key >> jbits, j = key & jmask;
            inside = d2_seq(R + i * (uint64_t)d, S + j * (uint64_t)d, d) <= T;
            done++;
        }
        const uint64_t m = __ballot(inside);
        if (m) {
            const int k = __popcll(m);
            uint64_t pos = 0;
            if (lane == 0) pos = atomicAdd((unsigned long long*)&stats[TJ_ST_OUT], (unsigned long long)k);
            pos = __shfl(pos, 0);
            if (inside) {
                const uint64_t at =
                    pos + (uint64_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                if (at < (uint64_t)out_cap) out_keys[at] = key;
            }
            acc_cnt += (lane == 0) ? (uint64_t)k : 0;
        }
    }
    for (int o = 32; o > 0; o >>= 1) done += __shfl_xor(done, o);
    if (lane == 0) {
        if (acc_cnt) atomicAdd((unsigned long long*)&stats[TJ_ST_FP64_ACCEPT], (unsigned long long)acc_cnt);
        if (done) atomicAdd((unsigned long long*)&stats[TJ_ST_FP64_REFINED], (unsigned long long)done);
    }
}

}  // namespace

namespace tj {
int launch_refine(const float* R, const float* S, int d, int jbits, double T, const uint64_t* keys,
                  const uint64_t* count, int64_t cap, uint64_t* out_keys, int64_t out_cap, uint64_t* stats,
                  hipStream_t s) {
    if (cap <= 0) return TJ_OK;
    int64_t blocks = (cap + 255) / 256;
    const int64_t maxb = (int64_t)num_cus() * 8;
    if (blocks > maxb) blocks = maxb;
    hipLaunchKernelGGL(refine_kernel, dim3((unsigned)blocks), dim3(256), 0, s, R, S, d, jbits, T, keys, count, cap,
                       out_keys, out_cap, stats);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}
}  // namespace tj
