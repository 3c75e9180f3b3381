// certify_impl.h -- per-pair certificate for the screen's candidates, and the
// FP64 refinement of the ambiguous band.
//
// A candidate is a pair the screen's fast reject could not exclude (all
// pairs inside plus a thin band around the threshold).  For each one the
// exact binary64 interval of its squared distance is formed from the MFMA
// dot and the row certificate terms (classify, screen_impl.h): certified
// inside -> output; certified outside -> dropped; ambiguous -> the
// sequential FP64 difference form decides (the refinement of
// CEJ/src/run_g2a_tensorjoin.py:193-232, in the order of
// CEJ/src/audit_g2b_accepted_fp64.py:51-57).  This is the reference's
// cascade of accept / ambiguous queue / FP64 terminal
// (TG/src/g2_operator_r1.py:52-57) with the queue kept on the device.
//
// Each refined pair also reports |d2_screen - d2_fp64| and its ratio to the
// certified half-width; a pair outside its interval counts as a violation
// (always 0 for a sound certificate).  With the audit flag, certified-inside
// pairs are recomputed in FP64 as well.
//
// The kernel is a template over the key resolver (Src::fetch), shared by the
// single-join pipeline and the batch of joins.
#pragma once
#include "screen_impl.h"

namespace tjcert {

// Everything the certificate of one pair needs.
struct PairRef {
    const JoinParams* P;
    RowMeta mi, mj;
    const float* xi;
    const float* xj;
    int d;
};

// Certified FP32 middle filter (the a8 stage of the reference's cascade,
// certified_fp32_filter_i64, CEJ/src/run_g3c_b_r1_gpu_cascade.py:57-129),
// re-derived: S = sum_k fl32(fl32(x_k - y_k)^2) in any summation order.  Each
// difference and each square of binary32 values is rounded once (relative
// 2^-24, whatever the cancellation) and every term is non-negative, so
// |S - d2| <= gamma_{d+2} d2 plus an underflow term below (d+2) 2^-125
// (flushed subnormals included).  The oracle's own gap g is charged by
// deciding against T(1 -+ 2g).
// Returns 1 (certainly d2_seq <= T), 0 (certainly >), 2 (undecided: FP64).
__device__ inline int decide32(float S, int d, double T, double g64) {
    if (!(S < INFINITY)) return 2;  // overflow (or NaN): the FP64 form decides
    const double n = (double)d + 2.0;
    const double gam = n * 0x1p-24 / (1.0 - n * 0x1p-24);
    const double absb = n * 0x1p-125;
    const double Sd = (double)S;
    // S >= d2 (1 - gam) - absb  and  S <= d2 (1 + gam) + absb, inverted for
    // d2; a few binary64 roundings, absorbed by the outward 2^-50 steps
    double hi = (Sd + absb) / (1.0 - gam);
    hi += hi * 0x1p-50;
    double lo = (Sd - absb) / (1.0 + gam);
    lo -= fabs(lo) * 0x1p-50;
    const double t_acc = nextafter(T * (1.0 - 2.0 * g64), 0.0);
    const double t_rej = nextafter(T * (1.0 + 2.0 * g64), INFINITY);
    if (hi <= t_acc) return 1;
    if (lo > t_rej) return 0;
    return 2;
}

__device__ inline uint64_t shfl64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// Lanes per pair for the cooperative filter: each lane takes at most four
// float4 of each row per step (16 coordinates), so one load instruction of
// the wave covers G*16 contiguous bytes of each of 64/G row pairs.
__device__ inline int filter_lanes(int d) {
    const int q = (d + 15) >> 4;
    int g = 1;
    while (g < q && g < 64) g <<= 1;
    return g;
}

// The FP32 filter for every lane of the wave with amb set, the rows read by
// groups of G lanes (G from the width of the first pending pair; any G is
// correct, the width only sets how many lanes share a row).  sPos/sDec: 64
// bytes of LDS each, private to the wave.  Returns this lane's decision
// (2 for lanes without an ambiguous pair).
__device__ inline int wave_filter32(bool amb, const float* xi, const float* xj, int d, double T, double g64,
                                    uint8_t* sPos, uint8_t* sDec, int lane) {
    const uint64_t am = __ballot(amb);
    if (amb) sPos[tjscreen::lanes_below(am)] = (uint8_t)lane;
    sDec[lane] = 2;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int n = __popcll(am);
    for (int done = 0; done < n;) {
        const int first = __builtin_amdgcn_readfirstlane((int)sPos[done]);
        const int G = __builtin_amdgcn_readfirstlane(filter_lanes(__shfl(d, first)));
        const int lg = __builtin_ctz(G);
        const int q = lane >> lg, r = lane & (G - 1);
        const int idx = done + q;
        const bool act = idx < n;
        const int owner = sPos[act ? idx : done];
        const float* a = (const float*)shfl64((uint64_t)xi, owner);
        const float* b = (const float*)shfl64((uint64_t)xj, owner);
        const int dd = __shfl(d, owner);
        const double TT = __shfl(T, owner), gg = __shfl(g64, owner);
        float s = 0.0f;
        if (act) {
            if ((((uintptr_t)a | (uintptr_t)b) & 15) == 0 && (dd & 3) == 0) {
                const float4* a4 = (const float4*)a;
                const float4* b4 = (const float4*)b;
                const int nv = dd >> 2;
                for (int v = r; v < nv; v += 4 * G) {
                    float4 va[4], vb[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int vu = min(v + u * G, nv - 1);  // in bounds; masked below
                        va[u] = a4[vu];
                        vb[u] = b4[vu];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float t0 = va[u].x - vb[u].x, t1 = va[u].y - vb[u].y;
                        const float t2 = va[u].z - vb[u].z, t3 = va[u].w - vb[u].w;
                        const float e = (t0 * t0 + t1 * t1) + (t2 * t2 + t3 * t3);
                        s = s + (v + u * G < nv ? e : 0.0f);
                    }
                }
            } else {
                for (int k = r; k < dd; k += G) {
                    const float t = a[k] - b[k];
                    s = s + t * t;
                }
            }
        }
        for (int o = G >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (act && r == 0) sDec[owner] = (uint8_t)decide32(s, dd, TT, gg);
        done += 64 >> lg;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return sDec[lane];
}

__device__ inline void atomic_max_pos(uint64_t* slot, double v) {
    if (v > 0.0) atomicMax((unsigned long long*)slot, (unsigned long long)__builtin_bit_cast(uint64_t, v));
}

template <class Src>
__global__ void __launch_bounds__(256) certify_kernel(const Src src, const uint64_t* __restrict__ ckeys,
                                                      const float* __restrict__ cp, int64_t ccap,
                                                      uint64_t* __restrict__ out_keys, int64_t out_cap,
                                                      uint64_t* __restrict__ stats, int audit) {
    int64_t count = (int64_t)stats[TJ_ST_CANDIDATES];
    if (count > ccap) count = ccap;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double max_err = 0.0, max_ratio = 0.0;
    uint64_t viol = 0, n_cert = 0, n_amb = 0, n_amb_in = 0, n_f32_in = 0, n_f32_out = 0, n_f64 = 0;
    // accepted keys are staged per wave and appended with one atomic per
    // OCAP keys (most candidates are accepted: one atomic per iteration would
    // serialise on the output cursor)
    constexpr int OCAP = 512;
    __shared__ uint64_t sOut[4][OCAP];
    __shared__ uint8_t sPos[4][64], sDec[4][64];
    int nO = 0;
    auto flush = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        uint64_t pos = 0;
        if (lane == 0) pos = atomicAdd((unsigned long long*)&stats[TJ_ST_OUT], (unsigned long long)nO);
        pos = __shfl(pos, 0);
        for (int t = lane; t < nO; t += 64)
            if (pos + (uint64_t)t < (uint64_t)out_cap) out_keys[pos + t] = sOut[wave][t];
        __builtin_amdgcn_wave_barrier();
        nO = 0;
    };
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
    for (int64_t base = w * 64; base < count; base += nwaves * 64) {
        const int64_t t = base + lane;
        bool inside = false;
        uint64_t key = 0;
        PairRef pr;
        pr.P = nullptr;
        pr.xi = pr.xj = nullptr;
        pr.d = 0;
        float p = 0.0f, est = 0.0f, hw = 0.0f;
        int cls = 0;
        bool fast = false;
        if (t < count) {
            key = ckeys[t];
            p = cp[t];
            src.fetch(key, pr);
            cls = tjscreen::classify(p, pr.mi, pr.mj, *pr.P, &est, &hw);
            inside = cls == 1;
            if (cls == 1) n_cert++;
            if (cls == 2) n_amb++;
            // The screen's fast accept (s >= a_i + w_j, with the largest a_i
            // of its wave) never reaches this pass outside audit mode; in
            // audit mode every pair it would accept is checked here.
            fast = (double)p >= (double)pr.mi.m2.y + (double)pr.mj.m2.x;
        }
        // ambiguous: the certified FP32 filter first (rows read cooperatively
        // by the wave), the sequential FP64 form for what it leaves (audit
        // mode: FP64 for every ambiguous pair, checking the filter too)
        const int f32 = wave_filter32(cls == 2, pr.xi, pr.xj, pr.d, pr.P ? pr.P->T : 0.0, pr.P ? pr.P->g64 : 0.0,
                                      sPos[wave], sDec[wave], lane);
        if (t < count) {
            const JoinParams& P = *pr.P;
            if (cls == 2) {
                if (f32 == 1) { n_f32_in++; inside = true; }
                if (f32 == 0) n_f32_out++;
            }
            if ((cls == 2 && (f32 == 2 || audit)) || (audit && (cls == 1 || fast))) {
                n_f64++;
                const double d2 = d2_seq<2>(pr.xi, pr.xj, pr.d);
                const bool in64 = d2 <= P.T;
                const double dj = d2 / P.scale_real;  // exact power-of-two scaling
                const double err = fabs(dj - (double)est);
                max_err = fmax(max_err, err * P.scale_real);
                if (hw > 0.0f) max_ratio = fmax(max_ratio, err / (double)hw);
                // the interval encloses the exact d2; allow the binary64 gap g
                // and the binary32 rounding of the centre
                const bool outside_interval = err > (double)hw + 2.0 * P.g64 * dj + 0x1p-23 * fabs((double)est);
                const bool f32_wrong = cls == 2 && ((f32 == 1 && !in64) || (f32 == 0 && in64));
                if (outside_interval || ((cls == 1 || fast) && !in64) || f32_wrong) {
                    viol++;
                    stats[TJ_ST_VIOL_KEY] = key | (cls == 1 ? (1ull << 63) : 0ull);
                    stats[TJ_ST_VIOL_D2] = __builtin_bit_cast(uint64_t, dj);
                    stats[TJ_ST_VIOL_EST] = (uint64_t)__builtin_bit_cast(uint32_t, est) |
                                            ((uint64_t)__builtin_bit_cast(uint32_t, hw) << 32);
                }
                if (cls == 2 && f32 == 2) {
                    if (in64) n_amb_in++;
                    inside = in64;
                }
            }
        }
        const uint64_t m = __ballot(inside);
        if (m) {
            const int k = __popcll(m);
            if (nO + k > OCAP) flush();
            if (inside) sOut[wave][nO + tjscreen::lanes_below(m)] = key;
            nO += k;
        }
    }
    if (nO) flush();
    for (int o = 32; o > 0; o >>= 1) {
        max_err = fmax(max_err, __shfl_xor(max_err, o));
        max_ratio = fmax(max_ratio, __shfl_xor(max_ratio, o));
        viol += __shfl_xor(viol, o);
        n_cert += __shfl_xor(n_cert, o);
        n_amb += __shfl_xor(n_amb, o);
        n_amb_in += __shfl_xor(n_amb_in, o);
        n_f32_in += __shfl_xor(n_f32_in, o);
        n_f32_out += __shfl_xor(n_f32_out, o);
        n_f64 += __shfl_xor(n_f64, o);
    }
    if (lane == 0) {
        atomic_max_pos(&stats[TJ_ST_MAX_ERR], max_err);
        atomic_max_pos(&stats[TJ_ST_MAX_RATIO], max_ratio);
        if (viol) atomicAdd((unsigned long long*)&stats[TJ_ST_VIOLATIONS], (unsigned long long)viol);
        if (n_cert) atomicAdd((unsigned long long*)&stats[TJ_ST_CERT_ACCEPT], (unsigned long long)n_cert);
        if (n_amb) atomicAdd((unsigned long long*)&stats[TJ_ST_AMB], (unsigned long long)n_amb);
        if (n_amb_in) atomicAdd((unsigned long long*)&stats[TJ_ST_FP64_ACCEPT], (unsigned long long)n_amb_in);
        if (n_f32_in) atomicAdd((unsigned long long*)&stats[TJ_ST_FP32_ACCEPT], (unsigned long long)n_f32_in);
        if (n_f32_out) atomicAdd((unsigned long long*)&stats[TJ_ST_FP32_REJECT], (unsigned long long)n_f32_out);
        if (n_f64) atomicAdd((unsigned long long*)&stats[TJ_ST_FP64_REFINED], (unsigned long long)n_f64);
    }
}

}  // namespace tjcert
