// plain32.hip -- the plain FP32-first control: an FP32 GEMM-form tile over
// every pair, a rigorous binary64 interval per pair, the sequential FP64
// form for the pairs the interval cannot decide.
//
// Counterpart of the reference's pedantic FP32-first control, run_operator
// (G16/src/fp32_gpu_norm_operator.py:107-161: FP32 SGEMM panel ->
// classify_pedantic_panel with a gamma_{2D+2} interval, G15/src/
// fp32_first_kernels.py:7-62 -> FP64 terminal for the ambiguous pairs), the
// denominator the reference judged its certified join against.  Re-derived
// for this library's contract (the sequential FP64 predicate d2_seq <= T)
// and for any magnitudes, not only |x| <= 1:
//
//   p      = sum_k x_k y_k in binary32, one fused multiply-add per k, so
//            |p - x.y| <= gamma sum_k |x_k y_k| + 4 d 2^-126
//            <= gamma l_i l_j + 4 d 2^-126 =: E   (gamma = (d+1) u/(1 - (d+1) u),
//            u = 2^-24; l >= ||x|| from the row pass; the absolute term covers
//            the roundings of the subnormal range)
//   d2     = ||x||^2 + ||y||^2 - 2 x.y  in [nlo_i + nlo_j - 2p - 2E,
//                                           nhi_i + nhi_j - 2p + 2E]
//            evaluated in binary64 and widened by 2^-50 of the terms'
//            magnitudes (its own roundings);
//   accept if hi <= T (1 - 2g), reject if lo > T (1 + 2g) (g = (d + 4) 2^-52,
//   the gap between the exact d2 and the oracle's sequential binary64 sum,
//   plus an absolute (d + 2) 2^-1074 for its subnormal range), otherwise the
//   pair is evaluated in the sequential binary64 form, in place.
//
// A pair whose p overflows, or whose row norms are not finite, has an
// infinite or NaN interval and goes to the FP64 form.
//
// MI355X design: the GEMM runs on the vector ALU in binary32 (one v_fma_f32
// per product, the rounding the bound assumes; no packed FP32), 128 x 128
// pair tiles per 256-thread workgroup, 8 x 8 pairs per thread, K staged
// through LDS in slices of 32 as binary32 rows (16-byte LDS reads, 16 for
// 512 FMAs), tiles walked XCD-contiguously as the plain FP64 kernel.
#include <algorithm>
#include "tj_internal.h"

namespace {

constexpr int TM = 128, KC = 32;

// Row terms: [nlo, nhi] brackets ||x||^2 (binary64), l >= ||x|| (binary32).
struct RowNorm {
    double nlo, nhi;
    float l;
    float pad[3];
};

// One wave per row: binary64 sum of squares (each square exact in binary64,
// the sum sequential over the wave's partial sums; any order has relative
// error below gamma_{d}(2^-53)), bracketed outward.
__global__ void __launch_bounds__(256) norm_kernel(const float* __restrict__ x, int64_t n, int d,
                                                   RowNorm* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * 4;
    for (int64_t r = w; r < n; r += nw) {
        const float* row = x + r * (int64_t)d;
        double s = 0.0;
        for (int k = lane; k < d; k += 64) {
            const double v = (double)row[k];
            s = s + v * v;
        }
        for (int o = 32; o > 0; o >>= 1) s = s + __shfl_xor(s, o);
        if (lane == 0) {
            const double gam = (double)(d + 2) * 0x1p-53 / (1.0 - (double)(d + 2) * 0x1p-53);
            RowNorm rn;
            rn.nlo = nextafter(s - s * gam, 0.0);
            rn.nhi = nextafter(s + s * gam, INFINITY);
            if (rn.nlo < 0.0) rn.nlo = 0.0;
            rn.l = f32_up(sqrt(rn.nhi) * (1.0 + 0x1p-50));  // above the rounding of sqrt and *
            rn.pad[0] = rn.pad[1] = rn.pad[2] = 0.0f;
            out[r] = rn;
        }
    }
}

struct Plain32Args {
    const float* R; int64_t nR;
    const float* S; int64_t nS;
    const RowNorm* nr; const RowNorm* ns;
    int d, jbits, self_join;
    double T;
    int64_t tiles_r, tiles_s, per_xcd;
    uint64_t* out_keys; int64_t out_cap;
    uint64_t* stats;
};

__device__ void plain32_tile(const Plain32Args& A, int64_t bi, int64_t bj, float (*sR)[TM], float (*sS)[TM],
                             uint64_t& n_cert, uint64_t& n_amb, uint64_t& n_amb_in) {
    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    const int64_t r0 = bi * TM, c0 = bj * TM;
    const int d = A.d;

    float acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = 0.0f;

    // loader: thread -> (row = tid / 2, k half = tid % 2), 16 coordinates each
    const int lrow = tid >> 1, lk = (tid & 1) * 16;
    const int64_t gr = r0 + lrow, gs = c0 + lrow;
    const bool vec = (d & 3) == 0 && ((((uintptr_t)A.R) | ((uintptr_t)A.S)) & 15) == 0;
    for (int k0 = 0; k0 < d; k0 += KC) {
        __syncthreads();
        if (vec && k0 + lk + 16 <= d) {
            // 16-byte loads of the row's 16 coordinates
#pragma unroll
            for (int u = 0; u < 16; u += 4) {
                const float4 a = gr < A.nR ? *(const float4*)(A.R + gr * d + k0 + lk + u) : make_float4(0.f, 0.f, 0.f, 0.f);
                const float4 b = gs < A.nS ? *(const float4*)(A.S + gs * d + k0 + lk + u) : make_float4(0.f, 0.f, 0.f, 0.f);
                sR[lk + u][lrow] = a.x; sR[lk + u + 1][lrow] = a.y; sR[lk + u + 2][lrow] = a.z; sR[lk + u + 3][lrow] = a.w;
                sS[lk + u][lrow] = b.x; sS[lk + u + 1][lrow] = b.y; sS[lk + u + 2][lrow] = b.z; sS[lk + u + 3][lrow] = b.w;
            }
        } else {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int k = k0 + lk + u;
                sR[lk + u][lrow] = (gr < A.nR && k < d) ? A.R[gr * d + k] : 0.0f;
                sS[lk + u][lrow] = (gs < A.nS && k < d) ? A.S[gs * d + k] : 0.0f;
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < KC; ++k) {
            const float4 xa = *(const float4*)&sR[k][ty * 8], xb = *(const float4*)&sR[k][ty * 8 + 4];
            const float4 ya = *(const float4*)&sS[k][tx * 8], yb = *(const float4*)&sS[k][tx * 8 + 4];
            const float x[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
            const float y[8] = {ya.x, ya.y, ya.z, ya.w, yb.x, yb.y, yb.z, yb.w};
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) acc[a][b] = __builtin_fmaf(x[a], y[b], acc[a][b]);
        }
    }

    const double gam = (double)(d + 1) * 0x1p-24 / (1.0 - (double)(d + 1) * 0x1p-24);
    const double absd = 4.0 * (double)d * 0x1p-126;
    const double g64 = (double)(d + 4) * 0x1p-52;
    const double tabs = (double)(d + 2) * 0x1p-1074;
    const double t_acc = nextafter(A.T * (1.0 - 2.0 * g64) - tabs, 0.0);
    const double t_rej = nextafter(A.T * (1.0 + 2.0 * g64) + tabs, INFINITY);
    // the column rows' terms once per tile
    double nlo_j[8], nhi_j[8];
    float l_j[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int64_t j = c0 + tx * 8 + b;
        const RowNorm r = A.ns[j < A.nS ? j : 0];
        nlo_j[b] = r.nlo;
        nhi_j[b] = r.nhi;
        l_j[b] = r.l;
    }
    // classify: bit 8 a + b of `in` / `amb` = pair (a, b) certified inside /
    // undecided
    uint64_t in = 0, amb = 0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int64_t i = r0 + ty * 8 + a;
        const RowNorm ri = A.nr[i < A.nR ? i : 0];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int64_t j = c0 + tx * 8 + b;
            const bool valid = i < A.nR && j < A.nS && (!A.self_join || i < j);
            const double p = (double)acc[a][b];
            const double E = gam * (double)ri.l * (double)l_j[b] + absd;
            const double mag = ri.nhi + nhi_j[b] + 2.0 * fabs(p) + 2.0 * E;
            const double lo = (ri.nlo + nlo_j[b] - 2.0 * p - 2.0 * E) - mag * 0x1p-50;
            const double hi = (ri.nhi + nhi_j[b] - 2.0 * p + 2.0 * E) + mag * 0x1p-50;
            const bool inside = valid && hi <= t_acc;
            const bool undecided = valid && !inside && !(lo > t_rej);
            in |= (uint64_t)inside << (8 * a + b);
            amb |= (uint64_t)undecided << (8 * a + b);
        }
    }
    n_cert += (uint64_t)__popcll(in);
    n_amb += (uint64_t)__popcll(amb);
    // the undecided pairs (or pairs with an infinite or NaN interval): the
    // sequential binary64 form, as the oracle
    for (uint64_t m = amb; m; m &= m - 1) {
        const int q = __builtin_ctzll(m);
        const int64_t i = r0 + ty * 8 + (q >> 3), j = c0 + tx * 8 + (q & 7);
        if (d2_seq<2>(A.R + i * d, A.S + j * d, d) <= A.T) {
            in |= 1ull << q;
            n_amb_in++;
        }
    }
    const int lane = tid & 63;
    for (int q = 0; q < 64; ++q) {
        const bool inside = (in >> q) & 1;
        const uint64_t m = __ballot(inside);
        if (m) {
            const int k = __popcll(m);
            uint64_t pos = 0;
            if (lane == 0) pos = atomicAdd((unsigned long long*)&A.stats[TJ_ST_OUT], (unsigned long long)k);
            pos = __shfl(pos, 0);
            if (inside) {
                const int64_t i = r0 + ty * 8 + (q >> 3), j = c0 + tx * 8 + (q & 7);
                const uint64_t at = pos + (uint64_t)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
                if (at < (uint64_t)A.out_cap) A.out_keys[at] = key_pack((uint64_t)i, (uint64_t)j, A.jbits);
            }
        }
    }
}

__global__ void __launch_bounds__(256) plain32_kernel(const Plain32Args A) {
    __shared__ __attribute__((aligned(16))) float sR[KC][TM];
    __shared__ __attribute__((aligned(16))) float sS[KC][TM];
    uint64_t n_cert = 0, n_amb = 0, n_amb_in = 0;
    // persistent: workgroup g walks tiles g, g + grid, ...; the walk order is
    // XCD-contiguous (dispatch places workgroup g on XCD g % 8)
    for (int64_t vt = blockIdx.x; vt < A.per_xcd * 8; vt += gridDim.x) {
        const int64_t t = (vt % 8) * A.per_xcd + vt / 8;
        if (t >= A.tiles_r * A.tiles_s) continue;
        const int64_t bi = t / A.tiles_s, bj = t % A.tiles_s;
        if (A.self_join && bj < bi) continue;
        plain32_tile(A, bi, bj, sR, sS, n_cert, n_amb, n_amb_in);
    }
    for (int o = 32; o > 0; o >>= 1) {
        n_cert += __shfl_xor(n_cert, o);
        n_amb += __shfl_xor(n_amb, o);
        n_amb_in += __shfl_xor(n_amb_in, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (n_cert) atomicAdd((unsigned long long*)&A.stats[TJ_ST_CERT_ACCEPT], (unsigned long long)n_cert);
        if (n_amb) {
            atomicAdd((unsigned long long*)&A.stats[TJ_ST_AMB], (unsigned long long)n_amb);
            atomicAdd((unsigned long long*)&A.stats[TJ_ST_FP64_REFINED], (unsigned long long)n_amb);
        }
        if (n_amb_in) atomicAdd((unsigned long long*)&A.stats[TJ_ST_FP64_ACCEPT], (unsigned long long)n_amb_in);
    }
}

}  // namespace

namespace tj {
size_t plain32_norm_bytes(int64_t n) { return sizeof(RowNorm) * (size_t)(n > 0 ? n : 1); }

int launch_plain32(const float* R, int64_t nR, const float* S, int64_t nS, int d, int jbits, double T, bool self_join,
                   void* normR, void* normS, uint64_t* out_keys, int64_t out_cap, uint64_t* stats, hipStream_t s) {
    if (nR == 0 || nS == 0) return TJ_OK;
    const int64_t nb_r = std::min<int64_t>((nR + 3) / 4, (int64_t)num_cus() * 16);
    hipLaunchKernelGGL(norm_kernel, dim3((unsigned)nb_r), dim3(256), 0, s, R, nR, d, (RowNorm*)normR);
    if (!self_join) {
        const int64_t nb_s = std::min<int64_t>((nS + 3) / 4, (int64_t)num_cus() * 16);
        hipLaunchKernelGGL(norm_kernel, dim3((unsigned)nb_s), dim3(256), 0, s, S, nS, d, (RowNorm*)normS);
    }
    Plain32Args a{};
    a.R = R; a.nR = nR; a.S = S; a.nS = nS;
    a.nr = (const RowNorm*)normR;
    a.ns = (const RowNorm*)(self_join ? normR : normS);
    a.d = d; a.jbits = jbits; a.self_join = self_join ? 1 : 0;
    a.T = T;
    a.tiles_r = (nR + TM - 1) / TM;
    a.tiles_s = (nS + TM - 1) / TM;
    a.per_xcd = (a.tiles_r * a.tiles_s + 7) / 8;
    a.out_keys = out_keys; a.out_cap = out_cap;
    a.stats = stats;
    int64_t grid = a.per_xcd * 8;
    const int64_t maxg = (int64_t)num_cus() * 4;
    if (grid > maxg) grid = maxg;
    hipLaunchKernelGGL(plain32_kernel, dim3((unsigned)grid), dim3(256), 0, s, a);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}
}  // namespace tj
