// screen_impl.h -- the certified MFMA distance screen (gfx950).
//
// Counterpart of the reference's fused screen tiles scan16 / scan8
// (TG/src/g2_kernels.py:18-82) and analytic_certificate_ragged_safe_i64
// (CEJ/src/g4b_r1_ragged_safe_kernel.py:7-123): a low-precision dot-product
// tile, a rigorous per-pair interval for the squared distance, a three-way
// class (certified in / certified out / ambiguous), and compaction of the
// certified-in pairs into the output and of the ambiguous pairs into a queue
// for the FP64 refinement.
//
// MI355X design:
//  * One persistent work-queue launch.  A work item is a panel of
//    BM = 4 waves x RB x 32 stationary rows (R) against a range of streamed
//    columns (S).  Items are taken from an atomic cursor, largest first
//    (triangular panels of a self-join shrink with the panel index); for an
//    R x S join consecutive items share the same column range, so the
//    workgroups resident on one XCD sweep the same S blocks through its L2.
//  * The stationary rows' quantized fragments stay in VGPRs for the whole
//    sweep (d <= 512 bytes-per-lane budget); S arrives in 32-column blocks,
//    double-buffered in LDS (row stride padded by 16 B against bank
//    conflicts), shared by the 4 waves.
//  * v_mfma_f32_32x32x16_{f16,bf16} / v_mfma_scale_f32_32x32x64_f8f6f4 with
//    FP32 accumulation; the k order inside a fragment is any fixed
//    permutation shared by A and B (a dot product does not depend on it).
//  * Fast path per pair: w = fma(-2, p, nprime_j) and one compare against the
//    row threshold thr_i (both precomputed by the row-metadata pass, see
//    meta.hip).  w > thr_i is a rigorous reject.  The rest (accepted pairs
//    and the ambiguous band) take the exact per-pair interval in binary64.
//  * Compaction: wave ballot + mbcnt prefix into a per-wave LDS staging
//    buffer, flushed to HBM with one atomic per 256 pairs.
#pragma once
#include "tj_internal.h"

namespace tjscreen {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr int ACAP = 256;  // per-wave staged accepted keys
constexpr int MCAP = 128;  // per-wave staged ambiguous entries

template <int PREC>
struct Cfg {
    static constexpr bool FP8 = (PREC == P_FP8 || PREC == P_FP8X3);
    static constexpr bool BF16 = (PREC == P_BF16 || PREC == P_BF16X3);
    static constexpr int TERMS = (PREC == P_FP16X3 || PREC == P_BF16X3 || PREC == P_FP8X3) ? 3 : 1;
    static constexpr int NT = TERMS == 3 ? 2 : 1;   // arrays per row (hi, lo)
    static constexpr int ES = FP8 ? 1 : 2;          // bytes per element
    static constexpr int KSTEP = FP8 ? 64 : 16;     // k per MFMA
    static constexpr int CH = FP8 ? 32 : 16;        // bytes per lane per MFMA operand
};

// A fragment: CH bytes per lane.
template <int CH> struct Frag;
template <> struct Frag<16> { uint4 v; };
template <> struct Frag<32> { uint4 v0, v1; };

__device__ inline void ld_frag(Frag<16>& f, const uint8_t* p) { f.v = *(const uint4*)p; }
__device__ inline void ld_frag(Frag<32>& f, const uint8_t* p) { f.v0 = *(const uint4*)p; f.v1 = *(const uint4*)(p + 16); }

template <int PREC>
__device__ inline void mma(f32x16& acc, const Frag<Cfg<PREC>::CH>& a, const Frag<Cfg<PREC>::CH>& b, int sa, int sb) {
    if constexpr (Cfg<PREC>::FP8) {
        i32x8 av = {(int)a.v0.x, (int)a.v0.y, (int)a.v0.z, (int)a.v0.w, (int)a.v1.x, (int)a.v1.y, (int)a.v1.z, (int)a.v1.w};
        i32x8 bv = {(int)b.v0.x, (int)b.v0.y, (int)b.v0.z, (int)b.v0.w, (int)b.v1.x, (int)b.v1.y, (int)b.v1.z, (int)b.v1.w};
        // format 0 = OCP e4m3 for both operands; E8M0 scales (127 = 2^0) per row / column
        acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc, 0, 0, 0, sa, 0, sb);
    } else if constexpr (Cfg<PREC>::BF16) {
        (void)sa; (void)sb;
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a.v), __builtin_bit_cast(bf16x8, b.v), acc, 0, 0, 0);
    } else {
        (void)sa; (void)sb;
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8, a.v), __builtin_bit_cast(half8, b.v), acc, 0, 0, 0);
    }
}

__device__ inline int lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}
__device__ inline uint64_t uni64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Exact per-pair interval in binary64 (join units) and its class:
// 1 = certified inside, 0 = certified outside, 2 = ambiguous.
template <int TERMS>
__device__ inline int classify(float p, const RowMeta& mi, const RowMeta& mj, const JoinParams& P,
                               float* est, float* hw) {
    const double pd = (double)p;
    const double Hi = mi.m0.z, Ei = mi.m0.w, Hj = mj.m0.z, Ej = mj.m0.w;
    double B;
    if (TERMS == 1) {
        B = P.gamma_acc * Hi * Hj + Hi * Ej + Ei * Hj + Ei * Ej + P.alpha;
    } else {
        const double Li = mi.m1.x, Lj = mj.m1.x;
        const double Zi = Hi + Li, Zj = Hj + Lj;
        B = P.gamma_acc * (Hi * Hj + Hi * Lj + Li * Hj) + Li * Lj + Zi * Ej + Ei * Zj + Ei * Ej + P.alpha;
    }
    B *= 1.0 + 0x1p-44;  // a dozen binary64 roundings, each <= 2^-53 relative
    const double slo = (double)mi.m0.x + (double)mj.m0.x;  // exact
    const double shi = (double)mi.m0.y + (double)mj.m0.y;  // exact
    double lo = slo - 2.0 * pd - 2.0 * B;
    double hi = shi - 2.0 * pd + 2.0 * B;
    const double sl = (shi + 2.0 * fabs(pd) + 2.0 * B) * 0x1p-48;
    lo -= sl;
    hi += sl;
    *est = (float)(0.5 * (slo + shi) - 2.0 * pd);
    *hw = f32_up(0.5 * (hi - lo) * (1.0 + 0x1p-40));
    if (hi <= P.Tj_acc) return 1;
    if (lo > P.Tj_rej) return 0;
    return 2;
}

template <int PREC, int DP, int RB, bool SELF>
__global__ void __launch_bounds__(256, 2) screen_kernel(const tj::ScreenArgs A) {
    using C = Cfg<PREC>;
    constexpr int ROWB = DP * C::ES;             // bytes of one quantized row
    constexpr int LROW = C::NT * ROWB + 16;      // LDS row stride (padded)
    constexpr int KS = DP / C::KSTEP;            // MFMA steps per row
    constexpr int R16 = ROWB / 16;               // 16-byte units per row
    constexpr int UNITS = C::NT * 32 * R16;      // 16-byte units per S block
    constexpr int UPT = (UNITS + 255) / 256;     // units per thread
    constexpr int BM = 4 * RB * 32;

    __shared__ __attribute__((aligned(16))) uint8_t sB[2][32 * LROW];
    __shared__ uint64_t sAcc[4][ACAP];
    __shared__ uint64_t sAmbK[4][MCAP];
    __shared__ float2 sAmbA[4][MCAP];
    __shared__ float sP[4][16 * 64];
    __shared__ float sThr[4][RB * 32];
    __shared__ int sItem;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int hh = lane >> 5, c32 = lane & 31;
    const JoinParams P = *A.params;
    const int64_t na = A.na, nb = A.nb;
    const int total_items = A.panels * A.col_splits;
    const uint8_t* qa[2] = {(const uint8_t*)A.qa_hi, (const uint8_t*)A.qa_lo};
    const uint8_t* qb[2] = {(const uint8_t*)A.qb_hi, (const uint8_t*)A.qb_lo};

    int nAcc = 0, nAmb = 0;
    uint64_t cntCert = 0, cntCand = 0;

    auto flush_acc = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd((unsigned long long*)&A.stats[TJ_ST_OUT], (unsigned long long)nAcc);
        base = uni64(base);
        for (int t = lane; t < nAcc; t += 64) {
            uint64_t idx = base + (uint64_t)t;
            if (idx < (uint64_t)A.out_cap) A.out_keys[idx] = sAcc[wave][t];
        }
        __builtin_amdgcn_wave_barrier();
        nAcc = 0;
    };
    auto flush_amb = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd((unsigned long long*)&A.stats[TJ_ST_AMB], (unsigned long long)nAmb);
        base = uni64(base);
        for (int t = lane; t < nAmb; t += 64) {
            uint64_t idx = base + (uint64_t)t;
            if (idx < (uint64_t)A.amb_cap) {
                A.amb_keys[idx] = sAmbK[wave][t];
                ((float2*)A.amb_aux)[idx] = sAmbA[wave][t];
            }
        }
        __builtin_amdgcn_wave_barrier();
        nAmb = 0;
    };

    for (;;) {
        if (tid == 0) sItem = atomicAdd((unsigned long long*)&A.stats[TJ_ST_WORK], 1ull);
        __syncthreads();
        const int item = sItem;
        __syncthreads();
        if (item >= total_items) break;
        const int panel = item % A.panels;
        const int split = item / A.panels;
        const int64_t r0 = (int64_t)panel * BM;
        const int64_t rw0 = r0 + (int64_t)wave * RB * 32;
        int64_t cb, ce;
        if (SELF) { cb = r0; ce = nb; }
        else { cb = (int64_t)split * A.cols_per_split; ce = cb + A.cols_per_split; if (ce > nb) ce = nb; }
        if (cb >= ce) continue;
        const int nblk = (int)((ce - cb + 31) / 32);

        // ---- stationary fragments, thresholds and scales
        Frag<C::CH> fa[C::NT][RB][KS];
        float thr[RB][16];
        int sa[C::NT][RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            int64_t row = rw0 + rb * 32 + c32;
            int64_t rowc = row < na ? row : na - 1;
#pragma unroll
            for (int a = 0; a < C::NT; ++a)
#pragma unroll
                for (int t = 0; t < KS; ++t) ld_frag(fa[a][rb][t], qa[a] + rowc * ROWB + (2 * t + hh) * C::CH);
            if constexpr (C::FP8) {
                uint32_t pk = __builtin_bit_cast(uint32_t, A.ma[rowc].m1.w);
                sa[0][rb] = (int)(pk & 0xFFFFu) - 32768 + 127;
                if (C::NT == 2) sa[C::NT - 1][rb] = (int)(pk >> 16) - 32768 + 127;
            } else {
#pragma unroll
                for (int a = 0; a < C::NT; ++a) sa[a][rb] = 127;
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                int64_t ri = rw0 + rb * 32 + 8 * (r >> 2) + 4 * hh + (r & 3);
                thr[rb][r] = ri < na ? A.ma[ri].m1.y : -INFINITY;
            }
            {
                int64_t ri = rw0 + rb * 32 + c32;
                if (hh == 0) sThr[wave][rb * 32 + c32] = ri < na ? A.ma[ri].m1.y : -INFINITY;
            }
        }

        // ---- S block staging (global -> VGPR -> LDS), double buffered
        uint4 pre[UPT];
        auto load_block = [&](int blk) {
            const int64_t c0 = cb + (int64_t)blk * 32;
#pragma unroll
            for (int u = 0; u < UPT; ++u) {
                int unit = tid + u * 256;
                if (UNITS % 256 == 0 || unit < UNITS) {
                    int a = unit / (32 * R16);
                    int rem = unit % (32 * R16);
                    int col = rem / R16, q = rem % R16;
                    int64_t j = c0 + col;
                    if (j >= nb) j = nb - 1;
                    pre[u] = *(const uint4*)(qb[a] + j * ROWB + q * 16);
                }
            }
        };
        auto store_block = [&](int buf) {
#pragma unroll
            for (int u = 0; u < UPT; ++u) {
                int unit = tid + u * 256;
                if (UNITS % 256 == 0 || unit < UNITS) {
                    int a = unit / (32 * R16);
                    int rem = unit % (32 * R16);
                    int col = rem / R16, q = rem % R16;
                    *(uint4*)(&sB[buf][col * LROW + a * ROWB + q * 16]) = pre[u];
                }
            }
        };
        load_block(0);
        store_block(0);
        __syncthreads();

        for (int blk = 0; blk < nblk; ++blk) {
            const int buf = blk & 1;
            const int64_t c0 = cb + (int64_t)blk * 32;
            if (blk + 1 < nblk) load_block(blk + 1);
            const int64_t jcol = c0 + c32;
            float nj = INFINITY;
            int sbh = 127, sbl = 127;
            if (jcol < nb) {
                float4 m1 = A.mb[jcol].m1;
                nj = m1.z;
                if constexpr (C::FP8) {
                    uint32_t pk = __builtin_bit_cast(uint32_t, m1.w);
                    sbh = (int)(pk & 0xFFFFu) - 32768 + 127;
                    sbl = (int)(pk >> 16) - 32768 + 127;
                }
            }
            const bool skip = SELF && (c0 + 31 <= rw0);
            if (!skip) {
                f32x16 acc[RB];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) acc[rb] = (f32x16){};
                const uint8_t* bb = &sB[buf][c32 * LROW + hh * C::CH];
#pragma unroll
                for (int t = 0; t < KS; ++t) {
                    Frag<C::CH> fb;
                    ld_frag(fb, bb + 2 * t * C::CH);
                    if constexpr (C::NT == 1) {
#pragma unroll
                        for (int rb = 0; rb < RB; ++rb) mma<PREC>(acc[rb], fa[0][rb][t], fb, sa[0][rb], sbh);
                    } else {
                        Frag<C::CH> fl;
                        ld_frag(fl, bb + ROWB + 2 * t * C::CH);
#pragma unroll
                        for (int rb = 0; rb < RB; ++rb) {
                            mma<PREC>(acc[rb], fa[0][rb][t], fb, sa[0][rb], sbh);
                            mma<PREC>(acc[rb], fa[0][rb][t], fl, sa[0][rb], sbl);
                            mma<PREC>(acc[rb], fa[1][rb][t], fb, sa[1][rb], sbh);
                        }
                    }
                }
                // ---- fast reject: w = fma(-2, p, nprime_j) > thr_i rejects rigorously
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    uint64_t any = 0;
#pragma unroll
                    for (int r = 0; r < 16; ++r) any |= __ballot(fmaf(-2.0f, acc[rb][r], nj) <= thr[rb][r]);
                    if (any == 0) continue;
                    // ---- per-pair certificate on the candidates: the block goes
                    // to LDS and is walked by one rolled loop.
#pragma unroll
                    for (int r = 0; r < 16; ++r) sP[wave][r * 64 + lane] = acc[rb][r];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
#pragma unroll 1
                    for (int r = 0; r < 16; ++r) {
                        const float p = sP[wave][r * 64 + lane];
                        const int rloc = rb * 32 + 8 * (r >> 2) + 4 * hh + (r & 3);
                        const bool cand = fmaf(-2.0f, p, nj) <= sThr[wave][rloc];
                        const uint64_t m = __ballot(cand);
                        if (m == 0) continue;
                        cntCand += (uint64_t)__popcll(m);
                        int cls = 0;
                        float est = 0.f, hw = 0.f;
                        uint64_t key = 0;
                        const int64_t i = rw0 + rloc;
                        if (cand && i < na && jcol < nb && (!SELF || i < jcol)) {
                            const RowMeta mi = A.ma[i];
                            const RowMeta mj = A.mb[jcol];
                            cls = classify<C::TERMS>(p, mi, mj, P, &est, &hw);
                            key = key_pack((uint64_t)i, (uint64_t)jcol, A.jbits);
                        }
                        if (A.audit && cls == 1) { cls = 3; key |= 1ull << 63; }
                        const uint64_t mA = __ballot(cls == 1);
                        const uint64_t mM = __ballot(cls >= 2);
                        if (mA) {
                            const int k = uni(__popcll(mA));
                            if (nAcc + k > ACAP) flush_acc();
                            if (cls == 1) sAcc[wave][nAcc + lanes_below(mA)] = key;
                            nAcc += k;
                            cntCert += (uint64_t)k;
                        }
                        if (mM) {
                            const int k = uni(__popcll(mM));
                            if (nAmb + k > MCAP) flush_amb();
                            if (cls >= 2) {
                                const int pos = nAmb + lanes_below(mM);
                                sAmbK[wave][pos] = key;
                                sAmbA[wave][pos] = make_float2(est, hw);
                            }
                            nAmb += k;
                            if (A.audit) cntCert += (uint64_t)__popcll(__ballot(cls == 3));
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            if (blk + 1 < nblk) store_block(buf ^ 1);
            __syncthreads();
        }
    }
    if (nAcc) flush_acc();
    if (nAmb) flush_amb();
    if (lane == 0) {
        if (cntCert) atomicAdd((unsigned long long*)&A.stats[TJ_ST_CERT_ACCEPT], (unsigned long long)cntCert);
        if (cntCand) atomicAdd((unsigned long long*)&A.stats[TJ_ST_CANDIDATES], (unsigned long long)cntCand);
    }
}

// Rows per wave: as many 32-row blocks as keep the stationary fragments
// within 128 VGPRs (512 bytes per lane), at most 2.
template <int PREC, int DP>
constexpr int rb_for() {
    using C = Cfg<PREC>;
    constexpr int bytes_per_block = C::NT * DP * C::ES / 2;  // per lane, one 32-row block (64 lanes x 2 halves)
    return bytes_per_block * 2 <= 512 ? 2 : 1;
}

template <int PREC, int DP, bool SELF>
int launch_one(const tj::ScreenArgs& a, int grid, hipStream_t s) {
    constexpr int RB = rb_for<PREC, DP>();
    hipLaunchKernelGGL((screen_kernel<PREC, DP, RB, SELF>), dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError() == hipSuccess ? TJ_OK : TJ_ERR_HIP;
}

template <int PREC, int DP, bool SELF>
int occupancy_one() {
    constexpr int RB = rb_for<PREC, DP>();
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, screen_kernel<PREC, DP, RB, SELF>, 256, 0) != hipSuccess) return 1;
    return blocks > 0 ? blocks : 1;
}

}  // namespace tjscreen

// Instantiate the dispatch for one precision over the supported widths.
#define TJ_SCREEN_DISPATCH(NAME, PREC, ...)                                                       \
    namespace tj {                                                                                \
    int NAME(int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int query) {  \
        switch (dp) {                                                                             \
            __VA_ARGS__                                                                           \
            default: set_error("screen: unsupported padded width %d", dp); return TJ_ERR_UNSUPPORTED; \
        }                                                                                         \
    }                                                                                             \
    }
#define TJ_CASE(PREC, DPV)                                                                                    \
    case DPV:                                                                                                 \
        if (query == 1) return self_join ? tjscreen::occupancy_one<PREC, DPV, true>()                        \
                                         : tjscreen::occupancy_one<PREC, DPV, false>();                      \
        if (query == 2) return 4 * tjscreen::rb_for<PREC, DPV>() * 32;                                        \
        return self_join ? tjscreen::launch_one<PREC, DPV, true>(a, grid, s) : tjscreen::launch_one<PREC, DPV, false>(a, grid, s);
