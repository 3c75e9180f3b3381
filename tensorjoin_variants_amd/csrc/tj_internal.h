// tj_internal.h -- shared definitions of libtjoin (gfx950 only).
//
// Units.  Every certificate quantity lives in "join units": x_join = x * 2^-G,
// with the power of two G chosen per join from the largest |x| (exact
// scaling, so the binary64 predicate scales exactly too:
// d2_seq(x_join) = d2_seq(x) * 2^-2G).  T_join = T * 2^-2G.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "../../include/tjoin.h"

#define TJ_VERSION 10000

// ----------------------------------------------------------------- errors
namespace tj {
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
}  // namespace tj

#define TJ_HIP(call)                                          \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return tj::hip_fail(e_, #call); \
    } while (0)
#define TJ_CHECK(cond, code, ...)        \
    do {                                 \
        if (!(cond)) {                   \
            tj::set_error(__VA_ARGS__);  \
            return (code);               \
        }                                \
    } while (0)

// ------------------------------------------------------- parameter block
// Written once per join on the device (tj_params), read by every stage so
// that no host round trip sits between stages.
struct JoinParams {
    double T;          // squared threshold, real units
    double Tj;         // T in join units
    double Tj_acc;     // accept iff interval hi <= Tj_acc  (Tj * (1 - 2g), rounded down)
    double Tj_rej;     // reject iff interval lo >  Tj_rej  (Tj * (1 + 2g), rounded up)
    double g64;        // relative gap between exact d2 and the binary64 sequential sum
    double gamma_acc;  // MFMA accumulation model: |p - sum a b| <= gamma * sum |a b| + alpha
    double alpha;      // absolute term of the accumulation model (join units)
    double rho;        // ratio used to split H_i E_j products into per-row terms
    double scale_real; // 2^(2G): join-unit d2 -> real d2
    int G;             // x_join = x * 2^-G
    int d, dp, prec, terms, esize;
    int pad[2];
};

// -------------------------------------------------------- precision traits
enum { P_NONE = 0, P_FP16 = 1, P_FP16X3 = 2, P_BF16 = 3, P_BF16X3 = 4, P_FP8 = 5, P_FP8X3 = 6 };

__host__ __device__ inline int prec_terms(int p) { return (p == P_FP16X3 || p == P_BF16X3 || p == P_FP8X3) ? 3 : 1; }
__host__ __device__ inline int prec_esize(int p) { return (p == P_FP8 || p == P_FP8X3) ? 1 : 2; }
__host__ __device__ inline bool prec_is_fp8(int p) { return p == P_FP8 || p == P_FP8X3; }
__host__ __device__ inline bool prec_is_bf16(int p) { return p == P_BF16 || p == P_BF16X3; }
// K elements per MFMA step of one lane pair (16-bit: 32x32x16; fp8: 32x32x64)
__host__ __device__ inline int prec_kstep(int p) { return prec_is_fp8(p) ? 64 : 16; }
// Exponent of the largest |x_join|: fp16 puts it in [2^14, 2^15) to use the
// whole normal range; bf16 and fp8 (per-row scales) put it in [1, 2).
__host__ __device__ inline int prec_top_exp(int p) { return (p == P_FP16 || p == P_FP16X3) ? 14 : 0; }
// Typical ||residual|| / ||quantized|| ratio, the weight of the per-row split
// of the H_i E_j cross terms (any positive value is sound; this one is tight).
__host__ __device__ inline double prec_rho(int p) {
    switch (p) {
        case P_FP16: return 0x1p-12;
        case P_FP16X3: return 0x1p-23;
        case P_BF16: return 0x1p-9;
        case P_BF16X3: return 0x1p-17;
        case P_FP8: return 0x1p-5;
        case P_FP8X3: return 0x1p-9;
        default: return 1.0;
    }
}
inline int round_up(int x, int m) { return (x + m - 1) / m * m; }
inline int padded_dim(int d, int prec) {
    // The screen instantiations cover these padded widths.
    const int k = prec_kstep(prec);
    int dp = round_up(d, k);
    if (!prec_is_fp8(prec)) {
        if (dp <= 16) return 16;
        if (dp <= 32) return 32;
        if (dp <= 64) return 64;
        if (dp <= 128) return 128;
        if (dp <= 256) return 256;
        if (dp <= 512) return 512;
        return -1;
    }
    if (dp <= 64) return 64;
    if (dp <= 128) return 128;
    if (dp <= 256) return 256;
    if (dp <= 512) return 512;
    return -1;
}

// ------------------------------------------------------ directed rounding
// binary64 -> binary32 rounded toward -inf / +inf, written from the
// definition: round to nearest, then step one ulp if the result lies on the
// wrong side of the exact value.
__host__ __device__ inline float f32_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return f;
}
__host__ __device__ inline float f32_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}

// ---------------------------------------------------------- format codecs
// IEEE binary16, round to nearest even, subnormals flushed to zero.
__device__ inline uint16_t enc_fp16(double v, double* back) {
    if (fabs(v) < 0x1p-14) { *back = 0.0; return (uint16_t)(signbit(v) ? 0x8000u : 0u); }
    _Float16 h = (_Float16)(float)v;  // v is a binary32 value here: one rounding
    *back = (double)(float)h;
    return __builtin_bit_cast(uint16_t, h);
}
// bfloat16 (upper half of binary32), round to nearest even, |v| < 2^-60 flushed.
__device__ inline uint16_t enc_bf16(double v, double* back) {
    if (fabs(v) < 0x1p-60) { *back = 0.0; return 0; }
    uint32_t u = __builtin_bit_cast(uint32_t, (float)v);
    uint32_t r = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
    *back = (double)__builtin_bit_cast(float, r << 16);
    return (uint16_t)r;
}
// OCP float8 e4m3 (bias 7, max 448, no infinities), round to nearest even,
// subnormals flushed.  Caller guarantees |v| <= 448.
__device__ inline uint8_t enc_e4m3(double v, double* back) {
    double a = fabs(v);
    if (a < 0x1p-6) { *back = 0.0; return 0; }
    int e = ilogb(a);               // a in [2^e, 2^(e+1))
    double step = ldexp(1.0, e - 3);
    double q = rint(a / step);      // 8..16, exact division by a power of two
    if (q >= 16.0) { e += 1; q = 8.0; step *= 2.0; }
    if (e > 8 || (e == 8 && q > 14.0)) { e = 8; q = 14.0; step = ldexp(1.0, 5); }
    double mag = q * step;
    *back = signbit(v) ? -mag : mag;
    return (uint8_t)((signbit(v) ? 0x80u : 0u) | ((uint32_t)(e + 7) << 3) | ((uint32_t)q - 8u));
}

// Per-row FP8 scale exponent: x_join * 2^-e has max magnitude in (112, 448].
__device__ inline int fp8_row_exp(double maxabs) {
    if (!(maxabs > 0.0)) return 0;
    int f = ilogb(maxabs);
    int e = f - 8;
    if (ldexp(maxabs, -e) > 448.0) e += 1;
    if (e < -100) e = -100;
    return e;
}

// ------------------------------------------------------------- row meta
// Three float4 per row:
//   m0 = {nlo, nhi, H, E}     ||x_join||^2 bracket, ||hi||, ||residual|| upper bounds
//   m1 = {L, t, v, e}         ||lo|| bound (hi/lo split), accumulator seed t of the
//                             row as the stationary side, reject level v of the row
//                             as the streamed side (s = t_i + p_ij < v_j rejects),
//                             FP8 scale exponents packed as int bits
//                             (e_hi in bits 0-15, e_lo in 16-31, biased by 32768).
//   m2 = {w, a, 0, 0}         accept terms: s = t_i + p_ij >= a_i + w_j certifies
//                             the pair inside (a of the row as the stationary
//                             side, w as the streamed side).
struct RowMeta { float4 m0, m1, m2; };

// Sequential binary64 difference form, k ascending, no FMA (every file is
// compiled with -ffp-contract=off): the predicate of
// CEJ/src/audit_g2b_accepted_fp64.py:51-57.
// CH float4 of each row are loaded before any is used (CH = 8: one 128-byte
// line, so a lane walking its own rows takes each line from L2 once instead
// of once per 16-byte piece; smaller CH where registers are scarcer).
template <int CH = 8>
__device__ inline double d2_seq(const float* __restrict__ x, const float* __restrict__ y, int d) {
    double s = 0.0;
    int k = 0;
    if (((((uintptr_t)x) | ((uintptr_t)y)) & 15) == 0) {
        for (; k + 4 * CH <= d; k += 4 * CH) {
            float4 a[CH], b[CH];
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                a[q] = *(const float4*)(x + k + 4 * q);
                b[q] = *(const float4*)(y + k + 4 * q);
            }
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                double t;
                t = (double)a[q].x - (double)b[q].x; s = s + t * t;
                t = (double)a[q].y - (double)b[q].y; s = s + t * t;
                t = (double)a[q].z - (double)b[q].z; s = s + t * t;
                t = (double)a[q].w - (double)b[q].w; s = s + t * t;
            }
        }
        for (; k + 4 <= d; k += 4) {
            const float4 a = *(const float4*)(x + k);
            const float4 b = *(const float4*)(y + k);
            double t;
            t = (double)a.x - (double)b.x; s = s + t * t;
            t = (double)a.y - (double)b.y; s = s + t * t;
            t = (double)a.z - (double)b.z; s = s + t * t;
            t = (double)a.w - (double)b.w; s = s + t * t;
        }
    }
    for (; k < d; ++k) {
        const double t = (double)x[k] - (double)y[k];
        s = s + t * t;
    }
    return s;
}

__host__ __device__ inline uint64_t key_pack(uint64_t i, uint64_t j, int jbits) { return (i << jbits) | j; }
inline int bits_for(int64_t n) { int b = 0; while (((int64_t)1 << b) < n) ++b; return b; }

namespace tj {

// ------------------------------------------------------ cross-lane moves
// v of lane (lane ^ jj): DPP moves within a 16-lane row for jj < 16 (quad
// permutations for 1 and 2, row rotations for 4 and 8: no LDS round trip),
// a bpermute for 16 and 32 (the gfx950 row / half swaps
// v_permlane16/32_swap measured slower there: midsort<64, 8> 0.39 -> 0.44 ms).
// jj is a compile-time constant after unrolling, so the switch folds away.
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, int jj, int lane) {
    const int x = (int)v;
    switch (jj) {
        case 1: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
        case 2: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
        case 4: {
            // row_ror:4 brings lane l - 4, row_ror:12 lane l + 4 (mod 16)
            const int a = __builtin_amdgcn_mov_dpp(x, 0x124, 0xF, 0xF, false);
            const int b = __builtin_amdgcn_mov_dpp(x, 0x12C, 0xF, 0xF, false);
            return (uint32_t)((lane & 4) ? a : b);
        }
        case 8: return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);  // row_ror:8
        default: return (uint32_t)__shfl_xor(x, jj);
    }
}
// The same for a binary64 value (its two halves).
__device__ __forceinline__ double xor_lane_f64(double v, int jj, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = xor_lane((uint32_t)u, jj, lane), hi = xor_lane((uint32_t)(u >> 32), jj, lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// ------------------------------------------------------------- launchers

// Workspace layout of the grid join (grid.hip), kept in the plan (api.hip).
struct GridLayout {
    size_t bounds, bpart, params, count_nr, count_ns, rkeys, skeys, sort_tmp, srow, rrow, hkey, hval, bitmap, bmpre,
        bmsum, runs, cursor, cnt, src, off, scan_part, outj, outj2, longrows, total;
    int64_t cand_cap;  // ids the probe's candidate buffer holds
    uint64_t hslots;
    uint64_t bm_bits;  // capacity of the cell-occupancy bitmap
    int idxbits;
};
// Output of the grid join: (i, j) rows, or the neighbour table of
// GDS-Join (counts per point + neighbour ids grouped by point).
enum { GRID_EMIT_I32 = TJ_OUT_I32, GRID_EMIT_I64 = TJ_OUT_I64, GRID_EMIT_TABLE = 16 };
GridLayout grid_layout(int64_t nR, int64_t nS, int64_t out_cap);
// Candidate buffer of a grid join with nR queries and out_cap output slots.
int64_t grid_cand_capacity(int64_t nR, int64_t out_cap);
// Row-grouped output (rowsort.hip).  launch_row_offsets: off[r] = counts of
// rows before r, off[n] = the total (also to *total when not NULL); part is
// row_scan_part_bytes(n) of scratch.  launch_row_sort: the ids of row r sit
// in outj at src[r] (or at off[r] when src is NULL), off[r + 1] - off[r] of
// them (clamped to out_cap); each row is sorted and emitted to out at
// positions [off[r], off[r + 1]), as (ilocal ? ilocal[r] : r, id) rows in
// emit's format or bare ids (GRID_EMIT_TABLE); outj2 (out_cap ids) is the
// long rows' merge buffer, longlist row_sort_scratch_bytes(n) of scratch
// (256-byte aligned).
size_t row_scan_part_bytes(int64_t n);
size_t row_sort_scratch_bytes(int64_t n);
int launch_row_offsets(const uint32_t* cnt, int64_t n, uint64_t* part, uint64_t* off, uint64_t* total,
                       hipStream_t s);
int launch_row_sort(const uint64_t* off, int64_t n, uint32_t* outj, uint32_t* outj2, int64_t out_cap, int emit,
                    void* out, const int32_t* ilocal, uint32_t* longlist, const uint64_t* src, hipStream_t s);
// self_join: S is R (one index); upper: keep j > i only (the self-join's
// i < j output).  Table output: counts (uint32 [nR]) and table (uint32,
// out_cap entries); pairs output: pairs ([out_cap, 2] in emit's format).
int launch_grid_join(const float* R, int64_t nR, const float* S, int64_t nS, int d, double T, bool self_join,
                     bool upper, uint8_t* ws, const GridLayout& L, int64_t out_cap, int emit, void* pairs,
                     uint32_t* counts, uint64_t* stats, hipStream_t s);
struct ScreenArgs {
    const void* qa_hi; const void* qa_lo; const RowMeta* ma; int64_t na;
    const void* qb_hi; const void* qb_lo; const RowMeta* mb; int64_t nb;
    const JoinParams* params;
    uint64_t* out_keys; int64_t out_cap;
    uint64_t* amb_keys; float* amb_aux; int64_t amb_cap;
    uint64_t* stats;
    int jbits;
    int panels, col_splits, cols_per_split; // work decomposition
    int audit;
};
// rb: 32-row blocks per wave of the body, 0 for its default.
int launch_screen(int prec, int dp, bool self_join, const ScreenArgs& a, int grid, hipStream_t s, int rb);
int screen_rows_per_wg(int prec, int dp, bool self_join, int rb);
int screen_max_grid(int prec, int dp, bool self_join, int rb);
// Work decomposition of the persistent screen over an nA x nB join: row
// panels of bm rows, column splits, and the persistent grid.  rb and
// split_cols (column split width) are tile choices, 0 for the defaults.
struct ScreenLayout { int bm, panels, col_splits, cols_per_split, grid, rb; };
int screen_layout(int64_t nA, int64_t nB, bool self_join, int dp, int prec, ScreenLayout* L, int rb,
                  int64_t split_cols);

int launch_scale_scan(const float* x, int64_t n, int d, uint32_t* maxabs, uint64_t* stats, hipStream_t s);
int launch_params(const uint32_t* maxabs, double T, int d, int dp, int prec, JoinParams* p, uint64_t* stats, hipStream_t s);
int launch_rowmeta(const float* x, int64_t n, int d, int dp, int prec, const JoinParams* p,
                   void* qhi, void* qlo, RowMeta* meta, uint64_t* stats, hipStream_t s);
int launch_refine(const float* R, const float* S, int d, int jbits, double T, const uint64_t* keys,
                  const uint64_t* count, int64_t cap, uint64_t* out_keys, int64_t out_cap, uint64_t* stats,
                  hipStream_t s);
int launch_certify(const float* R, const float* S, int d, int jbits, const uint64_t* cand_keys,
                   const float* cand_p, int64_t cand_cap, const RowMeta* ma, const RowMeta* mb,
                   const JoinParams* p, uint64_t* out_keys, int64_t out_cap, uint64_t* stats, int audit,
                   hipStream_t s);
int launch_plain(const float* R, int64_t nR, const float* S, int64_t nS, int d, int jbits, double T,
                 bool self_join, uint64_t* out_keys, int64_t out_cap, uint64_t* stats, hipStream_t s);
size_t plain32_norm_bytes(int64_t n);
int launch_plain32(const float* R, int64_t nR, const float* S, int64_t nS, int d, int jbits, double T, bool self_join,
                   void* normR, void* normS, uint64_t* out_keys, int64_t out_cap, uint64_t* stats, hipStream_t s);
size_t sort_temp_bytes(int64_t capacity);
// result == NULL: the sorted keys end in `keys`; otherwise *result names the
// buffer holding them (keys or the ping-pong half of temp), saving the copy
// back after an odd number of passes.  lo_bit (a multiple of 8) skips the
// digits below it: keys already in order of those bits among equal higher
// bits (e.g. written in index order with the index in the low bits) come
// out as from the full sort.
int launch_sort(void* temp, uint64_t* keys, const uint64_t* count, int64_t capacity, int bits, hipStream_t s,
                uint64_t** result = nullptr, int lo_bit = 0);
int launch_unpack(const uint64_t* keys, const uint64_t* count, int64_t cap, int jbits, void* pairs,
                  int out_format, hipStream_t s);
int launch_fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t s);
int num_cus();
}  // namespace tj
