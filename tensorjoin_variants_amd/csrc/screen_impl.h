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
//    BM = 4 waves x RB x 32 stationary rows (R; RB = 1, 2 or 4) against a range of streamed
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
//  * Fast path per pair: s = t_i + p against the column level v_j (row and
//    column terms from the row-metadata pass, meta_impl.h); s < v_j is a
//    rigorous reject, s >= a_i + w_j a rigorous accept.  Bodies of up to two
//    32-row blocks per wave seed the MFMA accumulator with t_i; four-block
//    bodies (no VGPRs left for seeds) accumulate from zero and add t_i in
//    binary32 afterwards.  The band between takes the exact per-pair
//    interval in binary64.
//  * Compaction: wave ballot + mbcnt prefix into a per-wave LDS staging
//    buffer, flushed to HBM with one atomic per 256 pairs.
#pragma once
#include <type_traits>
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
// s is the screen's t_i + p, either the MFMA output of a chain seeded with
// t_i or fl(t_i + p) after an unseeded chain; both satisfy
//   |s - t_i - x_i.x_j| <= B_ij,
//   B_ij = gamma (|t_i| + H_i H_j [+ H_i L_j + L_i H_j]) + L_i L_j + Z_i E_j + E_i Z_j + E_i E_j + alpha
//          + 2^-24 |s| + 2^-149   (the rounding of the binary32 sum, if any)
//   d2 in [nlo_i + nlo_j - 2p - 2B, nhi_i + nhi_j - 2p + 2B],  p = s - t_i,
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
    B += fabs((double)s) * 0x1p-24 + 0x1p-149;
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

// Wait until at most N vector-memory operations of this wave are in flight.
template <int N>
__device__ inline void wait_vm_c() {
    static_assert(N >= 0 && N <= 63, "vmcnt field is 6 bits");
    __asm__ volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}
// ... at most younger * P operations: the loads of the `younger` blocks
// issued after the one about to be consumed (younger <= 4).
template <int P>
__device__ inline void wait_blocks(int younger) {
    switch (younger) {
        case 0: wait_vm_c<0>(); break;
        case 1: wait_vm_c<P>(); break;
        case 2: wait_vm_c<2 * P>(); break;
        case 3: wait_vm_c<3 * P>(); break;
        default: wait_vm_c<4 * P>(); break;
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
// Scalar-base forms: a wave-uniform 64-bit base in SGPRs plus a per-lane
// 32-bit byte offset, so the per-block address math is scalar.
__device__ inline void dma16s(uint64_t sbase, uint32_t voff, uint32_t lds_wave_base) {
    __asm__ volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, %2" ::"s"(lds_wave_base), "v"(voff), "s"(sbase)
                     : "memory", "m0");
}
__device__ inline void dma4s(uint64_t sbase, uint32_t voff, uint32_t lds_wave_base) {
    __asm__ volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dword %1, %2" ::"s"(lds_wave_base), "v"(voff), "s"(sbase)
                     : "memory", "m0");
}

// max of the 16 accumulator registers of a lane as a 3-level tree of max3
// (a chain would be 8 dependent instructions long)
__device__ inline float max16(const f32x16& a) {
    const float m0 = fmaxf(fmaxf(a[0], a[1]), a[2]), m1 = fmaxf(fmaxf(a[3], a[4]), a[5]);
    const float m2 = fmaxf(fmaxf(a[6], a[7]), a[8]), m3 = fmaxf(fmaxf(a[9], a[10]), a[11]);
    const float m4 = fmaxf(fmaxf(a[12], a[13]), a[14]);
    const float n0 = fmaxf(fmaxf(m0, m1), m2), n1 = fmaxf(fmaxf(m3, m4), a[15]);
    return fmaxf(n0, n1);
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

// Per-wave staged candidates: 256 for the single-join screens (a larger
// buffer slows the FP8 d = 512 body by a fifth), 896 for the batched
// screens, whose dense-survivor joins then flush (two atomics on the two
// shared cursors, a wait for their return) 1.75 times less often; 896 is
// what the batched kernel's LDS leaves for two workgroups per CU.
constexpr int CCAP = 256;
constexpr int CCAP_BATCH = 896;

// Operands of one panel sweep.  Row indices are positions in the quantized
// arrays; the emitted key of pair (i, j) is key_pack(i + koff_a, j + koff_b),
// which lets a batch of joins share one row space and one key space.
struct PanelIO {
    const uint8_t* qa[2]; const RowMeta* ma; int64_t na;  // rows [.., na) are valid
    const uint8_t* qb[2]; const RowMeta* mb; int64_t nb;  // columns [.., nb) are valid
    int64_t koff_a, koff_b;
    int jbits;
};

// Where the screen's survivors go: certified-inside pairs (fast accept) as
// keys straight to the output, the band as (key, s) to the candidate queue
// of the certificate pass.  Entries beyond a capacity are counted, not kept.
struct CandOut {
    uint64_t* keys; float* aux; int64_t cap; uint64_t* counter;  // band queue and its cursor
    uint64_t* out; int64_t out_cap; uint64_t* out_counter;       // output keys and cursor
    uint64_t* cert_counter;                                      // count of fast accepts
};

constexpr uint64_t ACCEPT_BIT = 1ull << 63;  // staged-key flag: certified inside

// Route the nC entries staged in this wave's LDS buffers: flagged keys to
// the output, the others with their s to the band queue.  At most two
// atomics per flush (the fast accepts are counted in nAcc and added to
// cert_counter once, by cand_finish).  Keys use at most 62 bits, so bit 63
// is free for the flag.
__device__ __forceinline__ void cand_flush(const CandOut& co, const uint64_t* sCK, const float* sCP, int& nC,
                                           uint32_t& nAcc, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    nC = uni(nC);
    int n_acc = 0, n_band = 0;
    for (int t = lane; t - lane < nC; t += 64) {
        const bool valid = t < nC;
        const bool acc = valid && (sCK[t] & ACCEPT_BIT) != 0;
        n_acc += __popcll(__ballot(acc));
        n_band += __popcll(__ballot(valid && !acc));
    }
    // wave-uniform by construction; say so, or the atomics below become
    // per-lane loops
    n_acc = uni(n_acc);
    n_band = uni(n_band);
    uint64_t base_o = 0, base_b = 0;
    nAcc += (uint32_t)n_acc;
    if (lane == 0) {
        if (n_acc) base_o = atomicAdd((unsigned long long*)co.out_counter, (unsigned long long)n_acc);
        if (n_band) base_b = atomicAdd((unsigned long long*)co.counter, (unsigned long long)n_band);
    }
    base_o = uni64(base_o);
    base_b = uni64(base_b);
    for (int t = lane; t - lane < nC; t += 64) {
        const bool valid = t < nC;
        const uint64_t key = valid ? sCK[t] : 0;
        const bool acc = valid && (key & ACCEPT_BIT) != 0;
        const bool band = valid && !acc;
        const uint64_t ma = __ballot(acc), mb = __ballot(band);
        if (acc) {
            const uint64_t idx = base_o + (uint64_t)lanes_below(ma);
            if (idx < (uint64_t)co.out_cap) co.out[idx] = key & ~ACCEPT_BIT;
        }
        if (band) {
            const uint64_t idx = base_b + (uint64_t)lanes_below(mb);
            if (idx < (uint64_t)co.cap) {
                co.keys[idx] = key;
                co.aux[idx] = sCP[t];
            }
        }
        base_o += (uint64_t)__popcll(ma);
        base_b += (uint64_t)__popcll(mb);
    }
    // keep the in-order vmcnt accounting of the prefetch pipeline exact
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    nC = 0;
}

// End of a wave's work: flush what is staged and add its fast accepts.
__device__ __forceinline__ void cand_finish(const CandOut& co, const uint64_t* sCK, const float* sCP, int& nC,
                                            uint32_t& nAcc, int lane) {
    if (nC) cand_flush(co, sCK, sCP, nC, nAcc, lane);
    if (lane == 0 && nAcc) atomicAdd((unsigned long long*)co.cert_counter, (unsigned long long)nAcc);
}

// LDS bytes of the S ring of one (precision, width) instantiation.
template <int PREC, int DP>
constexpr int ring_bytes() { return Geo<PREC, DP>::NSTAGE * Geo<PREC, DP>::BLKB; }

// One work item: the panel of BM = 4 x RB x 32 stationary rows starting at r0
// against columns [cb, ce), all 4 waves of the workgroup.  sB/sNp/sW/sSc are
// the workgroup's S ring (NSTAGE slots: quantized columns, reject levels v,
// accept terms w, FP8 scales), sCK/sCP this wave's staging buffers, sT this
// wave's RB x 32 row levels t (16-byte aligned; used by unseeded bodies).
// With fast_accept == 0 every survivor goes to the band queue (audit mode).
//
// Unseeded bodies (RB = 4): p is the MFMA dot accumulated from zero and the
// pair's level is s = fl(p + t_i).  Rounding to nearest is monotone and v_j
// is a binary32 number, so p + t_i >= v_j implies s >= v_j: s < v_j rejects
// as rigorously as the seeded chain's output does.  Per 32 x 32 tile a lane
// first tests fl(max_r p_r + max_r t_r) >= v_j (p_r + t_r is at most
// max p + max t), and only tiles that pass form the 16 sums.
template <int PREC, int DP, int RB, bool SELF, int CC = CCAP, bool GROUPED = false>
__device__ __forceinline__ void screen_panel(const PanelIO& io, const CandOut& co, int64_t r0, int64_t cb, int64_t ce,
                                             uint8_t* sB, float* sNp, float* sW, uint32_t* sSc, uint64_t* sCK,
                                             float* sCP, float* sT, int& nC, uint32_t& nAcc, int fast_accept) {
    using C = Cfg<PREC>;
    using Gm = Geo<PREC, DP>;
    constexpr int NSTAGE = Gm::NSTAGE;
    constexpr bool SEEDED = RB <= 2;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int hh = lane >> 5, c32 = lane & 31;
    const int64_t na = io.na, nb = io.nb;
    const int64_t rw0 = r0 + (int64_t)uni(wave) * RB * 32;  // uniform: tile tests stay scalar
    const int nblk = (int)((ce - cb + 31) / 32);

    // ---- stationary fragments, row levels and scales (VGPR resident)
    Frag<C::CH> fa[C::NT][RB][Gm::KS];
    f32x16 seed[SEEDED ? RB : 1];
    float tmax[RB];  // unseeded: largest t over the lane's 16 rows of each block (-inf: none valid)
    int sa[C::NT][RB];
    if constexpr (!SEEDED) {
        // t of the wave's rows; -inf past na, so such rows never survive
        for (int k = lane; k < RB * 32; k += 64) {
            const int64_t row = rw0 + k;
            sT[k] = row < na ? io.ma[row].m1.y : -INFINITY;
        }
    }
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
        if constexpr (SEEDED) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t ri = rw0 + rb * 32 + 8 * (r >> 2) + 4 * hh + (r & 3);
                seed[rb][r] = ri < na ? io.ma[ri].m1.y : -INFINITY;
            }
        } else {
            // register r of a lane holds row 8 (r / 4) + 4 hh + r % 4 of the block
            float tm = -INFINITY;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float4 tv = *(const float4*)(sT + rb * 32 + 8 * g + 4 * hh);
                tm = fmaxf(tm, fmaxf(fmaxf(tv.x, tv.y), fmaxf(tv.z, tv.w)));
            }
            tmax[rb] = tm;
        }
    }
    // largest accept term a_i over this wave's valid rows: s >= w_j + a_w
    // then certifies every row of the wave against column j
    float a_w = -INFINITY;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const int64_t row = rw0 + rb * 32 + c32;
        if (row < na) a_w = fmaxf(a_w, io.ma[row].m2.y);
    }
    for (int o = 32; o > 0; o >>= 1) a_w = fmaxf(a_w, __shfl_xor(a_w, o));
    if (!fast_accept) a_w = INFINITY;

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
        if constexpr (SEEDED) {
#pragma unroll
            for (int r = 0; r < 16; ++r) __asm__ volatile("" ::"v"(seed[rb][r]));
        } else {
            __asm__ volatile("" ::"v"(tmax[rb]));
        }
    }
    __asm__ volatile("" ::"v"(a_w));

    // ---- S blocks: direct global->LDS loads into an NSTAGE ring.  The
    // block's UNITS 16-byte units are spread evenly over the 4 waves (64
    // lanes per load, or UNITS/4 lanes when UNITS < 256), and every wave
    // also loads 8 columns of v, w (and the FP8 scales), so every wave
    // issues the same P loads per block and waits on one uniform count.
    // Unit p is 16-byte chunk q of column col; its byte offset from the
    // block's first row, col * ROWB + q * 16, is fixed per lane, so each
    // block only adds the scalar c0 * ROWB.  The tail block (columns past
    // nb) recomputes offsets clamped to the last row.
    constexpr bool WIDE = Gm::UNITS >= 256;
    constexpr int NU = WIDE ? Gm::UNITS / 256 : 1;  // 16-byte loads per wave per block
    constexpr int UPW = WIDE ? 64 : Gm::UNITS / 4;  // active lanes per load
    static_assert(!WIDE || Gm::UNITS % 256 == 0, "unit count is a multiple of 256 or below it");
    constexpr int NM = C::FP8 ? 3 : 2;              // per-column metadata loads per wave per block
    constexpr int P = NU + NM;
    const int wv = uni(wave);
    uint32_t voff[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int p = WIDE ? wv * 64 + u * 256 + lane : wv * UPW + (lane < UPW ? lane : 0);
        const int rem = p % (32 * Gm::R16);
        const int col = rem / Gm::R16;
        const int q = (rem % Gm::R16) ^ ((col >> Gm::FSH) & Gm::FMASK);
        voff[u] = (uint32_t)(col * Gm::ROWB + q * 16);
    }
    const uint64_t qb0 = uni64((uint64_t)(uintptr_t)io.qb[0]);
    const uint64_t qb1 = uni64((uint64_t)(uintptr_t)io.qb[1]);
    const uint64_t mbb = uni64((uint64_t)(uintptr_t)io.mb);
    const uint32_t ldsB = uni((int)lds_addr(sB)), ldsNp = uni((int)lds_addr(sNp)), ldsW = uni((int)lds_addr(sW)),
                   ldsSc = uni((int)lds_addr(sSc));
    // this lane's metadata column within a block (lanes 0-7), formed anew
    // per block: kept live across the loop it is spilled when registers are
    // tight, and the reload's vmcnt(0) would drain the prefetch ring
    auto meta_col = [&]() {
        int l = lane;
        __asm__ volatile("" : "+v"(l));
        return 8 * wv + (l & 7);
    };
    auto issue = [&](int blk, int slot) {
        const int64_t c0 = cb + (int64_t)blk * 32;
        const uint32_t dst = ldsB + (uint32_t)(slot * Gm::BLKB);
        const uint64_t mbase = mbb + (uint64_t)c0 * sizeof(RowMeta);
        const uint32_t mo = (uint32_t)(slot * 64 + 8 * wv) * 4u;
        if (c0 + 32 <= nb) {
            // every column valid: the offsets are the per-lane constants
            if (lane < UPW) {
#pragma unroll
                for (int u = 0; u < NU; ++u) {
                    const int pbase = WIDE ? wv * 64 + u * 256 : wv * UPW;
                    const int a = pbase / (32 * Gm::R16);  // hi or lo array: uniform
                    const uint64_t base = ((C::NT == 2 && a != 0) ? qb1 : qb0) + (uint64_t)c0 * Gm::ROWB;
                    dma16s(base, voff[u], dst + (uint32_t)pbase * 16u);
                }
            }
            if (lane < 8) {
                const uint32_t off = (uint32_t)meta_col() * (uint32_t)sizeof(RowMeta);
                dma4s(mbase, off + (uint32_t)offsetof(RowMeta, m1) + 8, ldsNp + mo);
                dma4s(mbase, off + (uint32_t)offsetof(RowMeta, m2), ldsW + mo);
                if constexpr (C::FP8) dma4s(mbase, off + (uint32_t)offsetof(RowMeta, m1) + 12, ldsSc + mo);
            }
        } else {
            // tail block: columns past nb read the last row instead
            const int last = (int)(nb - 1 - c0);
            if (lane < UPW) {
#pragma unroll
                for (int u = 0; u < NU; ++u) {
                    const int pbase = WIDE ? wv * 64 + u * 256 : wv * UPW;
                    const int a = pbase / (32 * Gm::R16);
                    const uint64_t base = ((C::NT == 2 && a != 0) ? qb1 : qb0) + (uint64_t)c0 * Gm::ROWB;
                    const int col = ((pbase + lane) % (32 * Gm::R16)) / Gm::R16;
                    const uint32_t off = col > last ? (uint32_t)(last * Gm::ROWB) + (voff[u] - (uint32_t)(col * Gm::ROWB))
                                                    : voff[u];
                    dma16s(base, off, dst + (uint32_t)pbase * 16u);
                }
            }
            if (lane < 8) {
                const int mcol = meta_col();
                const uint32_t off = (uint32_t)(mcol > last ? last : mcol) * (uint32_t)sizeof(RowMeta);
                dma4s(mbase, off + (uint32_t)offsetof(RowMeta, m1) + 8, ldsNp + mo);
                dma4s(mbase, off + (uint32_t)offsetof(RowMeta, m2), ldsW + mo);
                if constexpr (C::FP8) dma4s(mbase, off + (uint32_t)offsetof(RowMeta, m1) + 12, ldsSc + mo);
            }
        }
    };
    // ---- block body: the MFMA tile of block blk (ring slot cslot) for the
    // wave's RB row blocks, the fast reject and the survivors
    // (always inlined: an outlined copy would take its uniform arguments in
    // VGPRs, which the scalar asm operands below cannot accept)
    auto body = [&](int blk, int cslot) __attribute__((always_inline)) {
        const int64_t c0 = cb + (int64_t)blk * 32;
        const int64_t jcol = c0 + c32;
        const bool skip = SELF && (c0 + 31 <= rw0);
        if (skip) return;
        // v_j (+inf past the last column: never a survivor) and FP8 scales
        float nj = sNp[cslot * 64 + c32];
        int sbh = 127, sbl = 127;
        if constexpr (C::FP8) {
            const uint32_t pk = sSc[cslot * 64 + c32];
            sbh = (int)(pk & 0xFFFFu) - 32768 + 127;
            sbl = (int)(pk >> 16) - 32768 + 127;
        }
        if (c0 + 32 > nb && c32 > (int)(nb - 1 - c0)) nj = INFINITY;
        const float wj = sW[cslot * 64 + c32];  // read with v_j: the survivor path then has it in a register
        f32x16 acc[RB];
        const uint8_t* bbase = sB + cslot * Gm::BLKB;
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
                    if (t == 0) acc[rb] = mma_c<PREC>(fa[0][rb][t], fb, SEEDED ? seed[rb] : f32x16{}, sa[0][rb], sbh);
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
                    if (t == 0) acc[rb] = mma_c<PREC>(fa[0][rb][t], fb, SEEDED ? seed[rb] : f32x16{}, sa[0][rb], sbh);
                    else mma<PREC>(acc[rb], fa[0][rb][t], fb, sa[0][rb], sbh);
                    mma<PREC>(acc[rb], fa[0][rb][t], fl, sa[0][rb], sbl);
                    mma<PREC>(acc[rb], fa[1][rb][t], fb, sa[1][rb], sbh);
                }
            }
        }
        // ---- fast reject: s = t_i + p_ij < v_j rejects rigorously; one
        // max over the 16 rows of a lane (unseeded: plus the largest t of
        // those rows), one compare, one ballot per block
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            float mx = max16(acc[rb]);
            if constexpr (!SEEDED) mx += tmax[rb];
            if (__ballot(mx >= nj) == 0) continue;
            // ---- survivors: register r of this lane is row
            // ibase + off(r), off(r) = 8 (r / 4) + r % 4, column jcol.  Each
            // register with a survivor in any lane costs one ballot and one
            // staged (key, s) per survivor; the key's bit 63 marks a fast
            // accept, s >= u_j = w_j + a_w (rounded up).
            const int64_t ibase = rw0 + rb * 32 + 4 * hh;
            const uint64_t kb = ((uint64_t)(ibase + io.koff_a) << io.jbits) | (uint64_t)(jcol + io.koff_b);
            // (columns past nb have nj = +inf: no survivor reads uj there)
            const float u = wj + a_w;
            // >= w_j + a_w despite the rounding of the sum (and, unseeded, s
            // >= uj still gives t_i + p >= w_j + a_w despite that of s)
            const float uj = u + (fabsf(u) * 0x1p-22f + 0x1p-126f);
            // unseeded: a per-lane level thr <= v_j - max_r t_r (the binary32
            // difference, lowered by 2^-22 of itself: more than its rounding
            // and that of the margin) screens the 16 registers with one
            // compare each; p_r + t_r >= v_j implies p_r >= thr, so only the
            // registers that pass form s_r = fl(p_r + t_r) with their own t
            float thr = 0.0f;
            if constexpr (!SEEDED) {
                const float dl = nj - tmax[rb];
                thr = dl - (fabsf(dl) * 0x1p-22f + 0x1p-126f);
            }
            // Stage register r's survivors; DIAG adds the self-join's i < j:
            // off(r) < e <=> ibase + off(r) < jcol.  Ballots of single
            // compares are the compare masks themselves.
            auto stage = [&](auto diag) {
                constexpr bool DIAG = decltype(diag)::value;
                int e = 64;
                if constexpr (DIAG) {
                    // jcol - ibase = (c0 - first row of the tile) + c32 - 4 hh, small here
                    const int df = (int)(c0 - (rw0 + rb * 32)) + c32 - 4 * hh;
                    e = df < 0 ? 0 : (df > 64 ? 64 : df);
                }
                // all 16 compare masks first (independent compares pipeline),
                // then the scalar tests and the writes
                uint64_t ms[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    ms[r] = __builtin_amdgcn_ballot_w64(acc[rb][r] >= (SEEDED ? nj : thr));
                    if constexpr (DIAG) ms[r] &= __builtin_amdgcn_ballot_w64(8 * (r >> 2) + (r & 3) < e);
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int off = 8 * (r >> 2) + (r & 3);
                    float sv = acc[rb][r];
                    uint64_t m = ms[r];
                    if (m == 0) continue;
                    if constexpr (!SEEDED) {
                        sv += sT[rb * 32 + off + 4 * hh];
                        m &= __builtin_amdgcn_ballot_w64(sv >= nj);
                        if (m == 0) continue;
                    }
                    // the lane's own survivor test: the compare itself for a
                    // seeded body off the diagonal (one VALU op), else its
                    // bit of the mask (two tests made it)
                    const bool mine = (DIAG || !SEEDED) ? ((m >> lane) & 1) != 0 : sv >= nj;
                    const int k = __popcll(m);
                    nC = uni(nC);  // wave-uniform: keeps the capacity test a scalar branch
                    if (__builtin_expect(nC + k > CC, 0)) cand_flush(co, sCK, sCP, nC, nAcc, lane);  // out of line
                    if (mine) {
                        const int pos = nC + lanes_below(m);
                        // the row offset's key increment, formed here by one
                        // scalar shift: hoisted out of the column loop, the 16
                        // increments would be spilled to VGPR lanes
                        int js = io.jbits;
                        __asm__ volatile("" : "+s"(js));
                        uint64_t key = kb + ((uint64_t)off << js);
                        if (sv >= uj) key |= ACCEPT_BIT;
                        sCK[pos] = key;
                        sCP[pos] = sv;
                    }
                    nC = uni(nC + k);
                }
            };
            // Grouped staging (the batched screens, whose self-joins leave a
            // few percent of a tile's pairs as survivors, spread over most
            // of a lane's 16 registers): the registers go in groups of
            // GR = CC / 64 (at most CC survivors a group); a group's masks,
            // their counts and running offsets are formed first (compares,
            // then scalar bit counts and adds), one capacity test covers the
            // group, and every register's survivors are then written at
            // positions that depend on no other register.  The single-join
            // bodies keep the per-register form above: their survivors are
            // rare, and the grouped form's code changes the schedule of
            // their hot loop (the d = 128 body ran 8 % slower with it).
            auto gstage = [&](auto diag) {
                constexpr bool DIAG = decltype(diag)::value;
                int e = 64;
                if constexpr (DIAG) {
                    const int df = (int)(c0 - (rw0 + rb * 32)) + c32 - 4 * hh;
                    e = df < 0 ? 0 : (df > 64 ? 64 : df);
                }
                int js = io.jbits;
                __asm__ volatile("" : "+s"(js));
                constexpr int GR = CC >= 1024 ? 16 : (CC >= 512 ? 8 : 4);
                static_assert(CC >= 256, "a group of GR registers fits the staging buffer");
#pragma unroll
                for (int g0 = 0; g0 < 16; g0 += GR) {
                    bool pr[GR];
                    float sv[GR];
                    uint64_t ms[GR];
                    int base[GR];
                    int tot = 0;
                    // unseeded: the row levels t of the group's registers
                    // (register r holds row 8 (r / 4) + 4 hh + r % 4 of the
                    // block: four registers are one 16-byte read)
                    float tr[SEEDED ? 1 : GR];
                    if constexpr (!SEEDED) {
#pragma unroll
                        for (int q = 0; q < GR; q += 4) {
                            const float4 tv = *(const float4*)(sT + rb * 32 + 8 * ((g0 + q) >> 2) + 4 * hh);
                            tr[q] = tv.x; tr[q + 1] = tv.y; tr[q + 2] = tv.z; tr[q + 3] = tv.w;
                        }
                    }
#pragma unroll
                    for (int q = 0; q < GR; ++q) {
                        const int r = g0 + q;
                        const int off = 8 * (r >> 2) + (r & 3);
                        sv[q] = acc[rb][r];
                        bool p = acc[rb][r] >= (SEEDED ? nj : thr);
                        if constexpr (!SEEDED) {
                            sv[q] += tr[q];
                            p = p && sv[q] >= nj;
                        }
                        if constexpr (DIAG) p = p && off < e;
                        pr[q] = p;
                        ms[q] = __builtin_amdgcn_ballot_w64(p);
                        base[q] = tot;
                        tot += __popcll(ms[q]);
                    }
                    tot = uni(tot);
                    if (tot == 0) continue;
                    nC = uni(nC);  // wave-uniform: keeps the capacity test a scalar branch
                    if (__builtin_expect(nC + tot > CC, 0)) cand_flush(co, sCK, sCP, nC, nAcc, lane);  // out of line
#pragma unroll
                    for (int q = 0; q < GR; ++q) {
                        if (ms[q] == 0) continue;
                        if (pr[q]) {
                            const int off = 8 * ((g0 + q) >> 2) + ((g0 + q) & 3);
                            const int pos = nC + base[q] + lanes_below(ms[q]);
                            uint64_t key = kb + ((uint64_t)off << js);
                            if (sv[q] >= uj) key |= ACCEPT_BIT;
                            sCK[pos] = key;
                            sCP[pos] = sv[q];
                        }
                    }
                    nC = uni(nC + tot);
                }
            };
            if constexpr (GROUPED) {
                if (SELF && rw0 + rb * 32 + 31 >= c0) gstage(std::true_type{});
                else gstage(std::false_type{});
            } else {
                if (SELF && rw0 + rb * 32 + 31 >= c0) stage(std::true_type{});
                else stage(std::false_type{});
            }
        }
    };

    // Waves of this workgroup may still be reading the ring from a previous
    // sweep (a batched work item sweeps up to four column ranges back to
    // back, with no barrier between them): every wave passes here before a
    // slot is refilled.
    lds_barrier();
    if constexpr (NSTAGE >= 4) {
        // Rings of four slots or more: one barrier per two blocks.  After the
        // barrier both blocks have landed for every wave and the two
        // slots consumed before it are free, so the next two blocks go
        // there (two blocks of compute ahead of their use); the waves of
        // a workgroup, whose survivor paths differ, wait for each other
        // half as often.
        const int pre = nblk < 2 ? nblk : 2;
        for (int b = 0; b < pre; ++b) issue(b, b);
        for (int blk0 = 0; blk0 < nblk; blk0 += 2) {
            wait_vm_c<0>();
            lds_barrier();
            if (blk0 + 2 < nblk) issue(blk0 + 2, (blk0 + 2) % NSTAGE);
            if (blk0 + 3 < nblk) issue(blk0 + 3, (blk0 + 3) % NSTAGE);
            body(blk0, blk0 % NSTAGE);
            if (blk0 + 1 < nblk) body(blk0 + 1, (blk0 + 1) % NSTAGE);
        }
    } else {
        const int pre = nblk < NSTAGE - 1 ? nblk : NSTAGE - 1;
        for (int b = 0; b < pre; ++b) issue(b, b);
        int slot = 0, islot = pre == NSTAGE - 1 ? NSTAGE - 1 : 0;  // consumed / next issued ring slots
        for (int blk = 0; blk < nblk; ++blk) {
            static_assert(NSTAGE - 2 <= 4, "wait_blocks covers up to 4 younger blocks");
            if (blk + NSTAGE - 2 < nblk) wait_vm_c<P * (NSTAGE - 2)>();  // steady state
            else wait_blocks<P>(nblk - 1 - blk);
            lds_barrier();
            if (blk + NSTAGE - 1 < nblk) {
                issue(blk + NSTAGE - 1, islot);
                islot = islot + 1 == NSTAGE ? 0 : islot + 1;
            }
            const int cslot = slot;
            slot = slot + 1 == NSTAGE ? 0 : slot + 1;
            body(blk, cslot);
        }
    }
}

// Workgroups per CU the compiler should plan registers for: two (up to 256
// VGPRs each; bodies with 128 fragment VGPRs fit too, without spills, and
// their 77 KiB of LDS twice fits the CU's 160 KiB); three for single-tile
// waves with at most 64 fragment VGPRs (up to 168 each).
template <int PREC, int DP, int RB>
constexpr int min_blocks() {
    using C = Cfg<PREC>;
    constexpr int frag = C::NT * RB * (DP / C::KSTEP) * (C::CH / 4);
    return (RB == 1 && frag <= 64) ? 3 : 2;
}

template <int PREC, int DP, int RB, bool SELF>
__global__ void __launch_bounds__(256, (min_blocks<PREC, DP, RB>())) screen_kernel(const tj::ScreenArgs A) {
    using Gm = Geo<PREC, DP>;
    constexpr int BM = 4 * RB * 32;
    __shared__ __attribute__((aligned(16))) uint8_t sB[Gm::NSTAGE * Gm::BLKB];
    __shared__ __attribute__((aligned(16))) float sNp[Gm::NSTAGE * 64];
    __shared__ __attribute__((aligned(16))) float sW[Gm::NSTAGE * 64];
    __shared__ __attribute__((aligned(16))) uint32_t sSc[Gm::NSTAGE * 64];
    __shared__ uint64_t sCK[4][CCAP];
    __shared__ float sCP[4][CCAP];
    __shared__ __attribute__((aligned(16))) float sT[4][RB <= 2 ? 4 : RB * 32];
    __shared__ int sItem;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int total_items = A.panels * A.col_splits;
    PanelIO io;
    io.qa[0] = (const uint8_t*)A.qa_hi; io.qa[1] = (const uint8_t*)A.qa_lo; io.ma = A.ma; io.na = A.na;
    io.qb[0] = (const uint8_t*)A.qb_hi; io.qb[1] = (const uint8_t*)A.qb_lo; io.mb = A.mb; io.nb = A.nb;
    io.koff_a = 0; io.koff_b = 0;
    io.jbits = A.jbits;
    const CandOut co{A.amb_keys, A.amb_aux, A.amb_cap, &A.stats[TJ_ST_CANDIDATES],
                     A.out_keys, A.out_cap, &A.stats[TJ_ST_OUT], &A.stats[TJ_ST_CERT_ACCEPT]};
    int nC = 0;
    uint32_t nAcc = 0;  // fast accepts of this wave (wave-uniform)

    for (;;) {
        if (tid == 0) sItem = atomicAdd((unsigned long long*)&A.stats[TJ_ST_WORK], 1ull);
        __syncthreads();
        const int item = uni(sItem);
        __syncthreads();
        if (item >= total_items) break;
        const int panel = item % A.panels;
        const int split = item / A.panels;
        const int64_t r0 = (int64_t)panel * BM;
        int64_t cb, ce;
        cb = (SELF ? r0 : 0) + (int64_t)split * A.cols_per_split;  // a self-join's panel starts at its diagonal
        ce = cb + A.cols_per_split;
        if (ce > A.nb) ce = A.nb;
        if (cb >= ce) continue;
        screen_panel<PREC, DP, RB, SELF>(io, co, r0, cb, ce, sB, sNp, sW, sSc, sCK[wave], sCP[wave], sT[wave], nC,
                                         nAcc, !A.audit);
    }
    cand_finish(co, sCK[wave], sCP[wave], nC, nAcc, lane);
}

// Rows per wave for batched joins: two 32-row blocks while their stationary
// fragments stay within 64 VGPRs (256 bytes per lane) together, else one.
template <int PREC, int DP>
constexpr int rb_batch() {
    using C = Cfg<PREC>;
    constexpr int bytes_per_block = C::NT * DP * C::ES / 2;  // per lane, one 32-row block (64 lanes x 2 halves)
    return bytes_per_block * 2 <= 256 ? 2 : 1;
}
// Single joins: R x S with single-term FP16 / BF16 codes of width 64 or 128
// take four blocks per wave (each LDS fragment feeds four MFMAs, the block
// overheads are spread over twice the MFMA work; unseeded, as the 64 seed
// registers would not fit); everything else as batched joins.
template <int PREC, int DP, bool SELF>
constexpr int rb_for() {
    using C = Cfg<PREC>;
    return (!SELF && C::NT == 1 && !C::FP8 && DP >= 64 && DP <= 128) ? 4 : rb_batch<PREC, DP>();
}

// The alternative row blocking a caller may select for a body (0: none):
// the four-block bodies can also run as two seeded blocks per wave.
template <int PREC, int DP, bool SELF>
constexpr int alt_rb() {
    return rb_for<PREC, DP, SELF>() == 4 ? 2 : 0;
}

// query 0: launch; 1: workgroups per CU; 2: rows per workgroup.
template <int PREC, int DP, bool SELF, int RB>
int run_rb(int query, const tj::ScreenArgs& a, int grid, hipStream_t s) {
    if (query == 1) {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, screen_kernel<PREC, DP, RB, SELF>, 256, 0) != hipSuccess)
            return 1;
        return blocks > 0 ? blocks : 1;
    }
    if (query == 2) return 4 * RB * 32;
    hipLaunchKernelGGL((screen_kernel<PREC, DP, RB, SELF>), dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError() == hipSuccess ? TJ_OK : TJ_ERR_HIP;
}

// rb: 32-row blocks per wave, 0 for the body's default.
template <int PREC, int DP, bool SELF>
int dispatch_rb(int rb, int query, const tj::ScreenArgs& a, int grid, hipStream_t s) {
    constexpr int R0 = rb_for<PREC, DP, SELF>();
    constexpr int R1 = alt_rb<PREC, DP, SELF>();
    if (rb == 0 || rb == R0) return run_rb<PREC, DP, SELF, R0>(query, a, grid, s);
    if constexpr (R1 > 0) {
        if (rb == R1) return run_rb<PREC, DP, SELF, R1>(query, a, grid, s);
    }
    tj::set_error("screen: %d row blocks per wave are not built for this precision and width", rb);
    return TJ_ERR_UNSUPPORTED;
}

}  // namespace tjscreen

// Instantiate the dispatch for one precision over the supported widths.
#define TJ_SCREEN_DISPATCH(NAME, PREC, ...)                                                              \
    namespace tj {                                                                                       \
    int NAME(int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int query, int rb) { \
        switch (dp) {                                                                                    \
            __VA_ARGS__                                                                                  \
            default: set_error("screen: unsupported padded width %d", dp); return TJ_ERR_UNSUPPORTED;   \
        }                                                                                                \
    }                                                                                                    \
    }
#define TJ_CASE(PREC, DPV)                                                             \
    case DPV:                                                                          \
        return self_join ? tjscreen::dispatch_rb<PREC, DPV, true>(rb, query, a, grid, s) \
                         : tjscreen::dispatch_rb<PREC, DPV, false>(rb, query, a, grid, s);
