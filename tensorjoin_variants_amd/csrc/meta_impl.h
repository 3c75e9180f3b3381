// meta_impl.h -- device code shared by the single-join and batched row
// metadata passes: the parameter block of one join and the quantized copy,
// norm bounds and screen seeds of one row.
#pragma once
#include "tj_internal.h"

namespace tjmeta {

// -------------------------------------------------------------- parameters
__device__ inline JoinParams make_params(uint32_t maxabs_bits, double T, int d, int dp, int prec) {
    JoinParams q;
    int G = 0;
    if (maxabs_bits != 0) G = ilogb((double)__builtin_bit_cast(float, maxabs_bits)) - prec_top_exp(prec);
    q.G = G;
    q.T = T;
    q.Tj = ldexp(T, -2 * G);
    q.d = d;
    q.dp = dp;
    q.prec = prec;
    q.terms = prec_terms(prec);
    q.esize = prec_esize(prec);
    // Relative gap between the exact squared distance and its sequential
    // binary64 evaluation: every term suffers one rounding in the difference,
    // one in the square and at most d-1 in the running sum, so
    // |d2_seq - d2| <= gamma_{d+2} d2 with unit roundoff 2^-53; (d+4)*2^-52
    // dominates gamma_{d+2} with room for the products below.
    q.g64 = (double)(d + 4) * 0x1p-52;
    const double two_g = 2.0 * q.g64;
    // Tj*(1-2g) rounded down / Tj*(1+2g) rounded up (one binary64 rounding
    // each, absorbed by stepping one more ulp outward).
    q.Tj_acc = nextafter(q.Tj * (1.0 - two_g), 0.0);
    q.Tj_rej = nextafter(q.Tj * (1.0 + two_g), INFINITY);
    // MFMA FP32 accumulation model (order- and rounding-mode-agnostic): the
    // k products of one accumulator chain are exact in binary32 (fp16/bf16/
    // fp8 significands multiply exactly); each of at most 2k+2 additions or
    // internal alignments loses at most 2^-23 relative, so
    // |p - sum a b| <= gamma * sum |a b| + alpha, gamma = m u / (1 - m u),
    // m = 2k+2, u = 2^-23; alpha covers flush-to-zero of subnormal partials.
    // Same form as the reference's FP16 premise (F16C/NUMERICAL_SCOPE.md:11-43,
    // gamma = 1026/(2^23-1026) at k = 512).
    // (k counts the products of every term of the hi/lo split; +1 for the
    // accumulator seed t_i of the fast path)
    const double k = (double)prec_terms(prec) * dp + 1.0;
    const double mu = (2.0 * k + 2.0) * 0x1p-23;
    q.gamma_acc = mu / (1.0 - mu) * (1.0 + 0x1p-40);
    // The FP8 MFMA (scaled f8f6f4 and plain fp8 alike) sums the products of
    // one instruction in a fixed-point window aligned to the largest product,
    // about 18 bits wide: measured on the device (tj_mfma_probe), a dominant
    // product plus 63 products 2^-k smaller loses up to 2^-12.4 of the
    // largest product, random operands up to 2^-11.6 of sum |a b|.  The
    // premise adds 2^-9 sum |a b| (a 6x margin) to the binary32 chain term.
    if (prec_is_fp8(prec)) q.gamma_acc += 0x1p-9;
    q.alpha = (2.0 * k + 2.0) * 0x1p-126;
    q.rho = prec_rho(prec);
    q.scale_real = ldexp(1.0, 2 * G);
    q.pad[0] = q.pad[1] = 0;
    return q;
}

// ------------------------------------------------------------------ rowmeta
// One group of GW lanes (16, 32 or 64, aligned within the wave) per row.
// Lane t of the group handles coordinates t, t+GW, ... of the padded row.
// Sums of non-negative binary64 terms, any order: relative error
// <= (n+1) 2^-53.
// (butterflies: DPP moves for partners within 16 lanes)
template <int GW>
__device__ inline double group_sum(double x) {
    const int lane = (int)__lane_id();
#pragma unroll
    for (int o = GW / 2; o > 0; o >>= 1) x += tj::xor_lane_f64(x, o, lane);
    return x;
}
template <int GW>
__device__ inline double group_max(double x) {
    const int lane = (int)__lane_id();
#pragma unroll
    for (int o = GW / 2; o > 0; o >>= 1) x = fmax(x, tj::xor_lane_f64(x, o, lane));
    return x;
}

// Lanes per row for a padded width: no idle lanes for narrow rows.
__host__ __device__ inline int group_width(int dp) { return dp <= 16 ? 16 : (dp <= 32 ? 32 : 64); }
// Lanes per row of the row-metadata pass: at least four coordinates per
// lane, so a 64-row wave takes 64 / width rows per step (more rows in
// flight for mid widths).
__host__ __device__ inline int meta_width(int dp) { return dp <= 64 ? 16 : (dp <= 128 ? 32 : 64); }

// The per-row sums of one row (binary64), held by every lane of its group.
struct RowSums {
    double nx, nh, nl, nr;  // |x|^2, |hi|^2, |lo|^2, |residual|^2 (join units)
    int e_hi, e_lo;         // FP8 per-row scale exponents
};

__device__ inline RowSums shfl_sums(const RowSums& s, int src) {
    RowSums t;
    t.nx = __shfl(s.nx, src);
    t.nh = __shfl(s.nh, src);
    t.nl = __shfl(s.nl, src);
    t.nr = __shfl(s.nr, src);
    t.e_hi = __shfl(s.e_hi, src);
    t.e_lo = __shfl(s.e_lo, src);
    return t;
}

// Quantize one row (its GW lanes; lane t of the group handles coordinates
// t, t+GW, ... of the padded row -- or, for rows of whole 16-byte quads
// (d % 4 == 0, the row 16-byte aligned, dp a multiple of 4 GW), the quads
// t, t+GW, ...: one 16-byte load and one 8- or 4-byte code store per four
// coordinates), write its codes at index `row`, and return its sums (binary64
// sums of non-negative terms: their bounds hold in any order).  act == false:
// the group takes part in the shuffles but stores nothing (a group past the
// last row).
template <int PREC, int GW = 64>
__device__ inline RowSums rowmeta_sums(const float* __restrict__ xr, int64_t row, int d, int dp, const JoinParams& P,
                                       int lane, uint8_t* __restrict__ qhi, uint8_t* __restrict__ qlo, bool act) {
    constexpr int TERMS = (PREC == P_FP16X3 || PREC == P_BF16X3 || PREC == P_FP8X3) ? 3 : 1;
    constexpr bool FP8 = (PREC == P_FP8 || PREC == P_FP8X3);
    constexpr bool BF16 = (PREC == P_BF16 || PREC == P_BF16X3);
    const double sc = ldexp(1.0, -P.G);

    // FP8: the per-row scale needs the row max first.
    int e_hi = 0, e_lo = 0;
    const bool quads = (d & 3) == 0 && dp % (4 * GW) == 0 && ((uintptr_t)xr & 15) == 0;
    if (FP8) {
        double m = 0.0;
        if (quads) {
            for (int k = 4 * lane; k < d; k += 4 * GW) {
                const float4 q = *(const float4*)(xr + k);
                m = fmax(m, fmax(fmax(fabs((double)q.x * sc), fabs((double)q.y * sc)),
                                 fmax(fabs((double)q.z * sc), fabs((double)q.w * sc))));
            }
        } else {
            for (int k = lane; k < d; k += GW) m = fmax(m, fabs((double)xr[k] * sc));
        }
        m = group_max<GW>(m);
        e_hi = fp8_row_exp(m);
        if (TERMS == 3) {
            // lo carries the residual of the hi rounding: at most half an ulp
            // of e4m3 (2^-4 relative) of the largest element.
            double mr = 0.0;
            for (int k = lane; k < d; k += GW) {
                double v = (double)xr[k] * sc, b;
                enc_e4m3(ldexp(v, -e_hi), &b);
                mr = fmax(mr, fabs(v - ldexp(b, e_hi)));
            }
            mr = group_max<GW>(mr);
            e_lo = fp8_row_exp(mr);
        }
    }

    double nx = 0.0, nh = 0.0, nl = 0.0, nr = 0.0;
    // one coordinate: its codes (hi, lo) and its terms of the four sums
    auto one = [&](double v, uint32_t& ch, uint32_t& cl) {
        double bh = 0.0, bl = 0.0;
        cl = 0;
        if (FP8) {
            ch = enc_e4m3(ldexp(v, -e_hi), &bh);
            bh = ldexp(bh, e_hi);
            if (TERMS == 3) {
                cl = enc_e4m3(ldexp(v - bh, -e_lo), &bl);
                bl = ldexp(bl, e_lo);
            }
        } else {
            ch = BF16 ? enc_bf16(v, &bh) : enc_fp16(v, &bh);
            if (TERMS == 3) cl = BF16 ? enc_bf16(v - bh, &bl) : enc_fp16(v - bh, &bl);
        }
        const double r = v - bh - bl;  // exact: the codes are roundings of v and of v - bh
        nx += v * v;
        nh += bh * bh;
        nl += bl * bl;
        nr += r * r;
    };
    if (quads) {
        for (int k = 4 * lane; k < dp; k += 4 * GW) {
            // a quad lies wholly inside the row or wholly in its padding
            const float4 q = k < d ? *(const float4*)(xr + k) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            uint32_t h[4], l[4];
            one((double)q.x * sc, h[0], l[0]);
            one((double)q.y * sc, h[1], l[1]);
            one((double)q.z * sc, h[2], l[2]);
            one((double)q.w * sc, h[3], l[3]);
            if (!act) continue;
            if (FP8) {
                *(uint32_t*)(qhi + row * dp + k) = h[0] | (h[1] << 8) | (h[2] << 16) | (h[3] << 24);
                if (TERMS == 3) *(uint32_t*)(qlo + row * dp + k) = l[0] | (l[1] << 8) | (l[2] << 16) | (l[3] << 24);
            } else {
                *(uint2*)((uint16_t*)qhi + row * dp + k) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
                if (TERMS == 3)
                    *(uint2*)((uint16_t*)qlo + row * dp + k) = make_uint2(l[0] | (l[1] << 16), l[2] | (l[3] << 16));
            }
        }
    } else {
        for (int k = lane; k < dp; k += GW) {
            uint32_t h, l;
            one(k < d ? (double)xr[k] * sc : 0.0, h, l);
            if (!act) continue;
            if (FP8) {
                qhi[row * dp + k] = (uint8_t)h;
                if (TERMS == 3) qlo[row * dp + k] = (uint8_t)l;
            } else {
                ((uint16_t*)qhi)[row * dp + k] = (uint16_t)h;
                if (TERMS == 3) ((uint16_t*)qlo)[row * dp + k] = (uint16_t)l;
            }
        }
    }
    RowSums out;
    out.nx = group_sum<GW>(nx);
    out.nh = group_sum<GW>(nh);
    out.nl = group_sum<GW>(nl);
    out.nr = group_sum<GW>(nr);
    out.e_hi = e_hi;
    out.e_lo = e_lo;
    return out;
}

// A row's certificate terms from its sums (one lane per row).
template <int PREC>
__device__ inline RowMeta rowmeta_terms(const RowSums& S, int dp, const JoinParams& P) {
    constexpr int TERMS = (PREC == P_FP16X3 || PREC == P_BF16X3 || PREC == P_FP8X3) ? 3 : 1;
    const double nx = S.nx, nh = S.nh, nl = S.nl, nr = S.nr;
    const double rel = (double)(dp + 4) * 0x1p-52;  // sum of dp non-negative terms, each squared once
    const double root_pad = 1.0 + 0x1p-48;           // sqrt rounding
    const float nlo = f32_down(nx * (1.0 - rel));
    const float nhi = f32_up(nx * (1.0 + rel));
    const float H = f32_up(sqrt(nh * (1.0 + rel)) * root_pad);
    const float L = TERMS == 3 ? f32_up(sqrt(nl * (1.0 + rel)) * root_pad) : 0.0f;
    const float E = f32_up(sqrt(nr * (1.0 + rel)) * root_pad);

    // Fast-path terms.  The screen seeds the MFMA accumulator of row i with
    // t_i, so one chain yields s_ij = t_i + p_ij (p the dot of the quantized
    // rows) with |s - (t_i + sum a b)| <= gamma (|t_i| + sum |a b|) + alpha.
    // (Its unseeded bodies compare fl(t_i + p) instead, p accumulated from
    // zero: by the monotonicity of rounding, fl(t_i + p) < v_j implies
    // t_i + p < v_j, which the bounds below reject without the gamma |t_i|
    // slack; the accept side is covered by the screen's margin on u_j.)
    // With Z = H + L the residual part of the exact dot is bounded by
    //   L_i L_j + Z_i E_j + E_i Z_j + E_i E_j,  sum |a b| <= H_i H_j [+ H_i L_j + L_i H_j],
    // and 2ab <= rho a^2 + b^2/rho splits everything except gamma |t_i| into
    //   beta_i + beta_j,  beta = ((gamma + rho) Z^2 + L^2 + (1/rho + 1) E^2) / 2.
    // The exact squared distance then satisfies
    //   d2 >= nlo_i + nlo_j - 2 s + 2 t_i - 2 gamma |t_i| - 2 beta_i - 2 beta_j - 2 alpha.
    // Choosing 2 t_i - 2 gamma |t_i| >= D_i = Tj_rej + 2 beta_i + 2 alpha - nlo_i
    // leaves  d2 > Tj_rej  whenever  s < v_j = (nlo_j - 2 beta_j) / 2:
    // the whole fast reject is one binary32 compare of the MFMA output.
    const double Z = (double)H + (double)L;
    const double rho = P.rho;
    const double pad = 1.0 + 0x1p-44;
    const double beta = ((P.gamma_acc + rho) * Z * Z + (double)L * L + (1.0 / rho + 1.0) * (double)E * E) * 0.5 * pad;
    double D = P.Tj_rej + 2.0 * beta + 2.0 * P.alpha - (double)nlo;
    D += fabs(D) * 0x1p-50 + 0x1p-1000;  // binary64 rounding of D, outward
    const double tseed = D >= 0.0 ? D / (2.0 * (1.0 - P.gamma_acc)) : D / (2.0 * (1.0 + P.gamma_acc));
    const float t32 = f32_up(tseed + fabs(tseed) * 0x1p-50 + 0x1p-1000);  // rounded upward
    double v = 0.5 * ((double)nlo - 2.0 * beta);
    v -= fabs(v) * 0x1p-50 + 0x1p-1000;
    const float v32 = f32_down(v);
    // Fast accept, from the same model in the other direction:
    //   d2 <= nhi_i + nhi_j - 2 s + 2 t_i + 2 gamma |t_i| + 2 alpha + 2 beta_i + 2 beta_j,
    // so s >= a_i + w_j with
    //   a_i >= (nhi_i + 2 t_i + 2 gamma |t_i| + 2 beta_i + 2 alpha - Tj_acc) / 2,
    //   w_j >= (nhi_j + 2 beta_j) / 2
    // certifies d2 <= Tj_acc (t_i is the binary32 seed actually used).
    const double tt = (double)t32;
    double a = 0.5 * ((double)nhi + 2.0 * tt + 2.0 * P.gamma_acc * fabs(tt) + 2.0 * beta + 2.0 * P.alpha - P.Tj_acc);
    a += fabs(a) * 0x1p-50 + 0x1p-1000;
    double w = 0.5 * ((double)nhi + 2.0 * beta);
    w += fabs(w) * 0x1p-50 + 0x1p-1000;
    RowMeta m;
    m.m0 = make_float4(nlo, nhi, H, E);
    uint32_t packed = (uint32_t)(S.e_hi + 32768) | ((uint32_t)(S.e_lo + 32768) << 16);
    m.m1 = make_float4(L, t32, v32, __builtin_bit_cast(float, packed));
    m.m2 = make_float4(f32_up(w), f32_up(a), 0.0f, 0.0f);
    return m;
}

// One wave, ROWS consecutive rows [r0, r0 + ROWS) of a range of n rows: the
// rows are quantized GW lanes each (64 / GW at a time), lane L collects row
// r0 + L's sums, and then every lane derives its own row's terms -- the
// binary64 tail runs on 64 lanes at once instead of one lane per group.
// src(r): the input row r of the range; row r's codes and metadata go to
// index base + r.
template <int PREC, int GW, class Src, int ROWS = 64>
__device__ inline void rowmeta_wave(Src src, int64_t n, int64_t r0, int d, int dp, const JoinParams& P, int lane,
                                    uint8_t* __restrict__ qhi, uint8_t* __restrict__ qlo, RowMeta* __restrict__ meta,
                                    int64_t base) {
    constexpr int RPW = 64 / GW;
    static_assert(ROWS % RPW == 0 && ROWS <= 64, "whole steps, one lane per row");
    const int g = lane / GW, lg = lane % GW;
    RowSums mine{};
    for (int step = 0; step < ROWS / RPW; ++step) {
        const int64_t first = r0 + (int64_t)step * RPW;
        if (first >= n) break;  // uniform
        const int64_t r = first + g;
        const bool act = r < n;
        const int64_t rc = act ? r : n - 1;
        const RowSums s = rowmeta_sums<PREC, GW>(src(rc), base + rc, d, dp, P, lg, qhi, qlo, act);
        // lane L <- the group of row r0 + L (one group per step: the
        // butterfly sums already sit in every lane)
        const RowSums t = RPW == 1 ? s : shfl_sums(s, (lane % RPW) * GW);
        if (lane / RPW == step) mine = t;
    }
    const int64_t my = r0 + lane;
    if (lane < ROWS && my < n) meta[base + my] = rowmeta_terms<PREC>(mine, dp, P);
}

}  // namespace tjmeta
