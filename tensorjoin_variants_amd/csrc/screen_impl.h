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

template <int PREC>
__device__ inline f32x16 mma_c(const Frag<Cfg<PREC>::CH>& a, const Frag<Cfg<PREC>::CH>& b, const f32x16& c, int sa, int sb) {
    f32x16 d = c;
    mma<PREC>(d, a, b, sa, sb);
    return d;
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
// s is the screen's MFMA output t_i + p (t_i the accumulator seed of row i):
//   p = s - t_i,  |p - x_i.x_j| <= B_ij,
//   B_ij = gamma (|t_i| + H_i H_j [+ H_i L_j + L_i H_j]) + L_i L_j + Z_i E_j + E_i Z_j + E_i E_j + alpha
//   d2 in [nlo_i + nlo_j - 2p - 2B, nhi_i + nhi_j - 2p + 2B]
// widened by the binary64 rounding of this evaluation; the oracle's own
// gap g is charged in Tj_acc / Tj_rej.
__device__ inline int classify(float s, const RowMeta& mi, const RowMeta& mj, const JoinParams& P,
                               float* est, float* hw) {
    const double ti = (double)mi.m1.y;
    const double pd = (double)s - ti;
    const double Hi = mi.m0.z, Ei = mi.m0.w, Hj = mj.m0.z, Ej = mj.m0.w;
    double B;
    if (P.terms == 1) {
        B = P.gamma_acc * (fabs(ti) + Hi * Hj) + Hi * Ej + Ei * Hj + Ei * Ej + P.alpha;
    } else {
        const double Li = mi.m1.x, Lj = mj.m1.x;
        const double Zi = Hi + Li, Zj = Hj + Lj;
        B = P.gamma_acc * (fabs(ti) + Hi * Hj + Hi * Lj + Li * Hj) + Li * Lj + Zi * Ej + Ei * Zj + Ei * Ej + P.alpha;
    }
    B *= 1.0 + 0x1p-44;  // a dozen binary64 roundings, each <= 2^-53 relative
    const double slo = (double)mi.m0.x + (double)mj.m0.x;  // exact
    const double shi = (double)mi.m0.y + (double)mj.m0.y;  // exact
    double lo = slo - 2.0 * pd - 2.0 * B;
    double hi = shi - 2.0 * pd + 2.0 * B;
    const double sl = (shi + 2.0 * fabs(pd) + 2.0 * fabs(ti) + 2.0 * B) * 0x1p-48;
    lo -= sl;
    hi += sl;
    *est = (float)(0.5 * (slo + shi) - 2.0 * pd);
    *hw = f32_up(0.5 * (hi - lo) * (1.0 + 0x1p-40));
    if (hi <= P.Tj_acc) return 1;
    if (lo > P.Tj_rej) return 0;
    return 2;
}

// Wait until at most n vector-memory operations of this wave are in flight
// (n is a small compile-time bound of the prefetch pipeline).
__device__ inline void wait_vm(int n) {
    switch (n) {
        case 0: __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
        case 1: __asm__ volatile("s_waitcnt vmcnt(1)" ::: "memory"); break;
        case 2: __asm__ volatile("s_waitcnt vmcnt(2)" ::: "memory"); break;
        case 3: __asm__ volatile("s_waitcnt vmcnt(3)" ::: "memory"); break;
        case 4: __asm__ volatile("s_waitcnt vmcnt(4)" ::: "memory"); break;
        case 5: __asm__ volatile("s_waitcnt vmcnt(5)" ::: "memory"); break;
        case 6: __asm__ volatile("s_waitcnt vmcnt(6)" ::: "memory"); break;
        case 7: __asm__ volatile("s_waitcnt vmcnt(7)" ::: "memory"); break;
        case 8: __asm__ volatile("s_waitcnt vmcnt(8)" ::: "memory"); break;
        case 9: __asm__ volatile("s_waitcnt vmcnt(9)" ::: "memory"); break;
        case 10: __asm__ volatile("s_waitcnt vmcnt(10)" ::: "memory"); break;
        case 11: __asm__ volatile("s_waitcnt vmcnt(11)" ::: "memory"); break;
        case 12: __asm__ volatile("s_waitcnt vmcnt(12)" ::: "memory"); break;
        case 13: __asm__ volatile("s_waitcnt vmcnt(13)" ::: "memory"); break;
        case 14: __asm__ volatile("s_waitcnt vmcnt(14)" ::: "memory"); break;
        default: __asm__ volatile("s_waitcnt vmcnt(15)" ::: "memory"); break;
    }
}
// Workgroup barrier without the implicit vmcnt(0) of __syncthreads (which
// would drain the prefetch pipeline); LDS stores are drained explicitly.
__device__ inline void lds_barrier() {
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// Direct global -> LDS loads (global_load_lds_dword[x4]: lane l of the wave
// writes LDS[m0 + size * l]).  Issued as inline asm so that the compiler does
// not treat every later LDS read as aliasing the DMA target and drain the
// whole prefetch pipeline; completion is tracked by the explicit vmcnt waits
// of the pipeline (wait_vm).  No VGPR is written by these instructions.
typedef __attribute__((address_space(3))) uint8_t lds_u8;
__device__ inline uint32_t lds_addr(const void* p) {
    return (uint32_t)(uintptr_t)(const lds_u8*)p;
}
__device__ inline void dma16(const void* src, uint32_t lds_wave_base) {
    __asm__ volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(lds_wave_base), "v"(src) : "memory", "m0");
}
__device__ inline void dma4(const void* src, uint32_t lds_wave_base) {
    __asm__ volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dword %1, off" ::"s"(lds_wave_base), "v"(src) : "memory", "m0");
}

template <int PREC, int DP>
struct Geo {
    using C = Cfg<PREC>;
    static constexpr int ROWB = DP * C::ES;                 // bytes of one quantized row
    static constexpr int R16 = ROWB / 16;                   // 16-byte units per row
    static constexpr int UNITS = C::NT * 32 * R16;          // units per 32-column block
    static constexpr int LPB = UNITS / 256;                 // full direct loads per thread per block
    static constexpr int BLKB = UNITS * 16;                 // bytes per stage
    static constexpr int NSTAGE = (32768 / BLKB) < 2 ? 2 : ((32768 / BLKB) > 6 ? 6 : (32768 / BLKB));
    static constexpr int FSH = R16 >= 16 ? 0 : (R16 == 8 ? 1 : (R16 == 4 ? 2 : 3));
    static constexpr int FMASK = (R16 >= 16 ? 16 : R16) - 1;
    static constexpr int KS = DP / C::KSTEP;
    // LDS unit holding 16-byte chunk q of column col (XOR swizzle: the 16
    // columns read by one 16-lane phase land in 16 distinct bank groups)
    __device__ static inline int unit(int col, int q) { return col * R16 + (q ^ ((col >> FSH) & FMASK)); }
};

constexpr int CCAP = 256;  // per-wave staged candidates

// Operands of one panel sweep.  Row indices are positions in the quantized
// arrays; the emitted key of pair (i, j) is key_pack(i + koff_a, j + koff_b),
// which lets a batch of joins share one row space and one key space.
struct PanelIO {
    const uint8_t* qa[2]; const RowMeta* ma; int64_t na;  // rows [.., na) are valid
    const uint8_t* qb[2]; const RowMeta* mb; int64_t nb;  // columns [.., nb) are valid
    int64_t koff_a, koff_b;
    int jbits;
};

// The global candidate queue (pair key, MFMA output s) and its cursor.
struct CandOut {
    uint64_t* keys; float* aux; int64_t cap; uint64_t* counter;
};

// Append the nC candidates staged in this wave's LDS buffers to the queue
// (one atomic per flush; entries beyond the capacity are counted, not kept).
__device__ __forceinline__ void cand_flush(const CandOut& co, const uint64_t* sCK, const float* sCP, int& nC, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint64_t base = 0;
    if (lane == 0) base = atomicAdd((unsigned long long*)co.counter, (unsigned long long)nC);
    base = uni64(base);
    for (int t = lane; t < nC; t += 64) {
        const uint64_t idx = base + (uint64_t)t;
        if (idx < (uint64_t)co.cap) {
            co.keys[idx] = sCK[t];
            co.aux[idx] = sCP[t];
        }
    }
    // keep the in-order vmcnt accounting of the prefetch pipeline exact
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    nC = 0;
}

// LDS bytes of the S ring of one (precision, width) instantiation.
template <int PREC, int DP>
constexpr int ring_bytes() { return Geo<PREC, DP>::NSTAGE * Geo<PREC, DP>::BLKB; }

// One work item: the panel of BM = 4 x RB x 32 stationary rows starting at r0
// against columns [cb, ce), all 4 waves of the workgroup.  sB/sNp/sSc are the
// workgroup's S ring (NSTAGE slots), sP/sCK/sCP this wave's staging buffers.
template <int PREC, int DP, int RB, bool SELF>
__device__ __forceinline__ void screen_panel(const PanelIO& io, const CandOut& co, int64_t r0, int64_t cb, int64_t ce,
                                             uint8_t* sB, float* sNp, uint32_t* sSc, float* sP, uint64_t* sCK,
                                             float* sCP, int& nC) {
    using C = Cfg<PREC>;
    using Gm = Geo<PREC, DP>;
    constexpr int NSTAGE = Gm::NSTAGE;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int hh = lane >> 5, c32 = lane & 31;
    const int64_t na = io.na, nb = io.nb;
    const int64_t rw0 = r0 + (int64_t)wave * RB * 32;
    const int nblk = (int)((ce - cb + 31) / 32);

    // ---- stationary fragments, thresholds and scales (VGPR resident)
    Frag<C::CH> fa[C::NT][RB][Gm::KS];
    f32x16 seed[RB];
    int sa[C::NT][RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int64_t row = rw0 + rb * 32 + c32;
        const int64_t rowc = row < na ? row : na - 1;
#pragma unroll
        for (int a = 0; a < C::NT; ++a)
#pragma unroll
            for (int t = 0; t < Gm::KS; ++t) ld_frag(fa[a][rb][t], io.qa[a] + rowc * Gm::ROWB + (2 * t + hh) * C::CH);
        if constexpr (C::FP8) {
            const uint32_t pk = __builtin_bit_cast(uint32_t, io.ma[rowc].m1.w);
            sa[0][rb] = (int)(pk & 0xFFFFu) - 32768 + 127;
            if (C::NT == 2) sa[C::NT - 1][rb] = (int)(pk >> 16) - 32768 + 127;
        } else {
#pragma unroll
            for (int a = 0; a < C::NT; ++a) sa[a][rb] = 127;
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t ri = rw0 + rb * 32 + 8 * (r >> 2) + 4 * hh + (r & 3);
            seed[rb][r] = ri < na ? io.ma[ri].m1.y : -INFINITY;
        }
    }

    // Consume the stationary registers here, so that the compiler's wait
    // counts (blind to the asm DMA below) settle before the block loop
    // instead of inserting vmcnt(0) inside it, which would drain the ring.
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
        for (int a = 0; a < C::NT; ++a)
#pragma unroll
            for (int t = 0; t < Gm::KS; ++t) {
                if constexpr (C::CH == 16) {
                    const uint4 q = fa[a][rb][t].v;
                    __asm__ volatile("" ::"v"(q.x), "v"(q.y), "v"(q.z), "v"(q.w));
                } else {
                    const uint4 q0 = fa[a][rb][t].v0, q1 = fa[a][rb][t].v1;
                    __asm__ volatile("" ::"v"(q0.x), "v"(q0.y), "v"(q0.z), "v"(q0.w));
                    __asm__ volatile("" ::"v"(q1.x), "v"(q1.y), "v"(q1.z), "v"(q1.w));
                }
            }
#pragma unroll
        for (int r = 0; r < 16; ++r) __asm__ volatile("" ::"v"(seed[rb][r]));
    }

    // ---- S blocks: direct global->LDS loads into an NSTAGE ring
    auto issue = [&](int blk) {
        const int slot = blk % NSTAGE;
        const int64_t c0 = cb + (int64_t)blk * 32;
        const uint32_t dst = __builtin_amdgcn_readfirstlane(lds_addr(sB + slot * Gm::BLKB));
#pragma unroll
        for (int u = 0; u * 256 < Gm::UNITS; ++u) {
            const int pbase = wave * 64 + u * 256;  // this wave's 64 consecutive units
            if (pbase < Gm::UNITS) {
                const int p = pbase + lane;
                const int a = p / (32 * Gm::R16);
                const int rem = p % (32 * Gm::R16);
                const int col = rem / Gm::R16;
                const int qs = rem % Gm::R16;
                const int q = qs ^ ((col >> Gm::FSH) & Gm::FMASK);
                int64_t j = c0 + col;
                if (j >= nb) j = nb - 1;
                // a select, not io.qb[a]: a runtime index would put io on the stack
                const uint8_t* src = (C::NT == 2 && a != 0) ? io.qb[1] : io.qb[0];
                dma16(src + j * Gm::ROWB + q * 16, __builtin_amdgcn_readfirstlane(dst + (uint32_t)pbase * 16u));
            }
        }
        if (wave == 0 && lane < 32) {
            int64_t j = c0 + lane;
            if (j >= nb) j = nb - 1;
            dma4(&io.mb[j].m1.z, __builtin_amdgcn_readfirstlane(lds_addr(sNp + slot * 64)));
            if constexpr (C::FP8) dma4(&io.mb[j].m1.w, __builtin_amdgcn_readfirstlane(lds_addr(sSc + slot * 64)));
        }
    };
    const int pre = nblk < NSTAGE - 1 ? nblk : NSTAGE - 1;
    for (int b = 0; b < pre; ++b) issue(b);

    for (int blk = 0; blk < nblk; ++blk) {
        const int younger = (nblk - 1 - blk) < (NSTAGE - 2) ? (nblk - 1 - blk) : (NSTAGE - 2);
        wait_vm(younger * Gm::LPB);
        lds_barrier();
        if (blk + NSTAGE - 1 < nblk) issue(blk + NSTAGE - 1);
        const int slot = blk % NSTAGE;
        const int64_t c0 = cb + (int64_t)blk * 32;
        const int64_t jcol = c0 + c32;
        const bool skip = SELF && (c0 + 31 <= rw0);
        if (skip) continue;
        float nj = INFINITY;
        int sbh = 127, sbl = 127;
        if (jcol < nb) {
            nj = sNp[slot * 64 + c32];
            if constexpr (C::FP8) {
                const uint32_t pk = sSc[slot * 64 + c32];
                sbh = (int)(pk & 0xFFFFu) - 32768 + 127;
                sbl = (int)(pk >> 16) - 32768 + 127;
            }
        }
        f32x16 acc[RB];
        const uint8_t* bbase = sB + slot * Gm::BLKB;
#pragma unroll
        for (int t = 0; t < Gm::KS; ++t) {
            Frag<C::CH> fb;
            if constexpr (C::CH == 16) {
                fb.v = *(const uint4*)(bbase + Gm::unit(c32, 2 * t + hh) * 16);
            } else {
                fb.v0 = *(const uint4*)(bbase + Gm::unit(c32, 2 * (2 * t + hh)) * 16);
                fb.v1 = *(const uint4*)(bbase + Gm::unit(c32, 2 * (2 * t + hh) + 1) * 16);
            }
            if constexpr (C::NT == 1) {
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    if (t == 0) acc[rb] = mma_c<PREC>(fa[0][rb][t], fb, seed[rb], sa[0][rb], sbh);
                    else mma<PREC>(acc[rb], fa[0][rb][t], fb, sa[0][rb], sbh);
                }
            } else {
                Frag<C::CH> fl;
                const uint8_t* lbase = bbase + 32 * Gm::R16 * 16;
                if constexpr (C::CH == 16) {
                    fl.v = *(const uint4*)(lbase + Gm::unit(c32, 2 * t + hh) * 16);
                } else {
                    fl.v0 = *(const uint4*)(lbase + Gm::unit(c32, 2 * (2 * t + hh)) * 16);
                    fl.v1 = *(const uint4*)(lbase + Gm::unit(c32, 2 * (2 * t + hh) + 1) * 16);
                }
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    if (t == 0) acc[rb] = mma_c<PREC>(fa[0][rb][t], fb, seed[rb], sa[0][rb], sbh);
                    else mma<PREC>(acc[rb], fa[0][rb][t], fb, sa[0][rb], sbh);
                    mma<PREC>(acc[rb], fa[0][rb][t], fl, sa[0][rb], sbl);
                    mma<PREC>(acc[rb], fa[1][rb][t], fb, sa[1][rb], sbh);
                }
            }
        }
        // ---- fast reject: s = t_i + p_ij < v_j rejects rigorously; one
        // max over the 16 rows of a lane, one compare, one ballot per block
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            float mx = acc[rb][0];
#pragma unroll
            for (int r = 1; r < 16; ++r) mx = fmaxf(mx, acc[rb][r]);
            if (__ballot(mx >= nj) == 0) continue;
            // ---- rare: stage the candidates (pair key, s) for the certify
            // pass.  A wave-uniform 16-bit mask names the registers that
            // hold one; the block goes to LDS and only those are visited.
            uint32_t rm = 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) rm |= (__ballot(acc[rb][r] >= nj) != 0 ? 1u : 0u) << r;
#pragma unroll
            for (int r = 0; r < 16; ++r) sP[r * 64 + lane] = acc[rb][r];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            rm = uni((int)rm);
            while (rm) {
                const int r = __builtin_ctz(rm);
                rm &= rm - 1;
                const float sv = sP[r * 64 + lane];
                const int rloc = rb * 32 + 8 * (r >> 2) + 4 * hh + (r & 3);
                const int64_t i = rw0 + rloc;
                const bool cand = sv >= nj && i < na && jcol < nb && (!SELF || i < jcol);
                const uint64_t m = __ballot(cand);
                if (m == 0) continue;
                const int k = uni(__popcll(m));
                if (nC + k > CCAP) cand_flush(co, sCK, sCP, nC, lane);
                if (cand) {
                    const int pos = nC + lanes_below(m);
                    sCK[pos] = key_pack((uint64_t)(i + io.koff_a), (uint64_t)(jcol + io.koff_b), io.jbits);
                    sCP[pos] = sv;
                }
                nC += k;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// The single-join screen: a persistent grid takes (panel, column split)
// items from an atomic cursor until the item space is exhausted.
template <int PREC, int DP, int RB, bool SELF>
__global__ void __launch_bounds__(256, 2) screen_kernel(const tj::ScreenArgs A) {
    using Gm = Geo<PREC, DP>;
    constexpr int BM = 4 * RB * 32;
    __shared__ __attribute__((aligned(16))) uint8_t sB[Gm::NSTAGE * Gm::BLKB];
    __shared__ __attribute__((aligned(16))) float sNp[Gm::NSTAGE * 64];
    __shared__ __attribute__((aligned(16))) uint32_t sSc[Gm::NSTAGE * 64];
    __shared__ float sP[4][16 * 64];
    __shared__ uint64_t sCK[4][CCAP];
    __shared__ float sCP[4][CCAP];
    __shared__ int sItem;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int total_items = A.panels * A.col_splits;
    PanelIO io;
    io.qa[0] = (const uint8_t*)A.qa_hi; io.qa[1] = (const uint8_t*)A.qa_lo; io.ma = A.ma; io.na = A.na;
    io.qb[0] = (const uint8_t*)A.qb_hi; io.qb[1] = (const uint8_t*)A.qb_lo; io.mb = A.mb; io.nb = A.nb;
    io.koff_a = 0; io.koff_b = 0;
    io.jbits = A.jbits;
    const CandOut co{A.amb_keys, A.amb_aux, A.amb_cap, &A.stats[TJ_ST_CANDIDATES]};
    int nC = 0;

    for (;;) {
        if (tid == 0) sItem = atomicAdd((unsigned long long*)&A.stats[TJ_ST_WORK], 1ull);
        __syncthreads();
        const int item = sItem;
        __syncthreads();
        if (item >= total_items) break;
        const int panel = item % A.panels;
        const int split = item / A.panels;
        const int64_t r0 = (int64_t)panel * BM;
        int64_t cb, ce;
        if (SELF) { cb = r0; ce = A.nb; }
        else { cb = (int64_t)split * A.cols_per_split; ce = cb + A.cols_per_split; if (ce > A.nb) ce = A.nb; }
        if (cb >= ce) continue;
        screen_panel<PREC, DP, RB, SELF>(io, co, r0, cb, ce, sB, sNp, sSc, sP[wave], sCK[wave], sCP[wave], nC);
    }
    if (nC) cand_flush(co, sCK[wave], sCP[wave], nC, lane);
}

// Rows per wave: two 32-row blocks while their stationary fragments stay
// within 64 VGPRs (256 bytes per lane) together, else one.
template <int PREC, int DP>
constexpr int rb_for() {
    using C = Cfg<PREC>;
    constexpr int bytes_per_block = C::NT * DP * C::ES / 2;  // per lane, one 32-row block (64 lanes x 2 halves)
    return bytes_per_block * 2 <= 256 ? 2 : 1;
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
