/*
 * tjoracle.c -- CPU FP64 oracle for the exact epsilon-join.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against; nothing in the product (tensorjoin_variants_amd)
 * links or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * Predicate (one pair):   d2_seq(x, y) <= T
 *   d2_seq is the sequential binary64 difference form
 *       s = 0;  for k = 0..d-1:  delta = (double)x[k] - (double)y[k];
 *                                s = fl(s + fl(delta * delta))
 *   with k ascending and no fused multiply-add.  This is the statement at
 *   CEJ/src/audit_g2b_accepted_fp64.py:51-57 (the reference's independent
 *   FP64 audit of every accepted pair) and GDS/main.cu:964 (GDS-Join's CPU
 *   brute force, runningDist += (a-b)*(a-b)).  The inclusive "<=" compare
 *   against T = fl64(eps^2) follows CEJ/src/run_g2a_tensorjoin.py:222
 *   (inside = distance_d2 <= threshold_d2) and CEJ/src/prepare_g2_cifar.py:104-108.
 *
 * Output order: pairs (i, j) lexicographically ascending; a self-join emits
 * only i < j (the upper triangle of CEJ/src/prepare_g2_cifar.py:130-153,
 * without the diagonal and without the mirrored lower half).
 *
 * Build: gcc -O3 -ffp-contract=off -fno-fast-math -fopenmp (oracle/Makefile).
 * SIMD is used across pairs only (each lane keeps its own sequential sum);
 * the k order inside one pair is never changed.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TJO_BLOCK 64 /* rows per work unit */
#define TJO_LANES 8  /* pairs evaluated side by side */

double tjo_d2(const float* x, const float* y, int64_t d) {
    double s = 0.0;
    for (int64_t k = 0; k < d; ++k) {
        double delta = (double)x[k] - (double)y[k];
        double sq = delta * delta;
        s = s + sq;
    }
    return s;
}

/* d2 of one x against TJO_LANES consecutive rows of a k-major block yt[k][lane]. */
__attribute__((target_clones("avx2", "default")))
static void d2_lanes(const float* x, const double* yt, int64_t d, double* out) {
    double s[TJO_LANES];
    for (int l = 0; l < TJO_LANES; ++l) s[l] = 0.0;
    for (int64_t k = 0; k < d; ++k) {
        double xv = (double)x[k];
        const double* yk = yt + k * TJO_LANES;
        for (int l = 0; l < TJO_LANES; ++l) {
            double delta = xv - yk[l];
            double sq = delta * delta;
            s[l] = s[l] + sq;
        }
    }
    for (int l = 0; l < TJO_LANES; ++l) out[l] = s[l];
}

typedef struct {
    int64_t* data; /* pairs, 2 per entry */
    int64_t n, cap;
} pairbuf;

static int pb_push(pairbuf* b, int64_t i, int64_t j) {
    if (b->n == b->cap) {
        int64_t nc = b->cap ? b->cap * 2 : 1024;
        int64_t* nd = (int64_t*)realloc(b->data, (size_t)nc * 2 * sizeof(int64_t));
        if (!nd) return -1;
        b->data = nd;
        b->cap = nc;
    }
    b->data[2 * b->n] = i;
    b->data[2 * b->n + 1] = j;
    b->n++;
    return 0;
}

/* Transpose S rows [j0, j0+TJO_LANES) into yt (k-major, doubles); rows past nS are +inf-safe zeros. */
static void pack_lanes(const float* S, int64_t nS, int64_t d, int64_t j0, double* yt) {
    for (int l = 0; l < TJO_LANES; ++l) {
        int64_t j = j0 + l;
        if (j < nS) {
            const float* y = S + j * d;
            for (int64_t k = 0; k < d; ++k) yt[k * TJO_LANES + l] = (double)y[k];
        } else {
            for (int64_t k = 0; k < d; ++k) yt[k * TJO_LANES + l] = 0.0;
        }
    }
}

/*
 * Brute-force join.  Writes up to `cap` pairs to out[2*p], out[2*p+1] and
 * returns the total number of pairs found (which may exceed cap: the caller
 * then retries with a larger buffer).  Returns -1 on allocation failure.
 * count_only != 0 skips pair storage (used for the CPU throughput baseline).
 */
int64_t tjo_join(const float* R, int64_t nR, const float* S, int64_t nS, int64_t d,
                 double T, int self, int64_t* out, int64_t cap, int nthreads,
                 int count_only) {
    if (self) { S = R; nS = nR; }
    int64_t nblocks = (nR + TJO_BLOCK - 1) / TJO_BLOCK;
    pairbuf* bufs = (pairbuf*)calloc((size_t)(nblocks > 0 ? nblocks : 1), sizeof(pairbuf));
    int64_t* counts = (int64_t*)calloc((size_t)(nblocks > 0 ? nblocks : 1), sizeof(int64_t));
    int failed = 0;
    if (!bufs || !counts) { free(bufs); free(counts); return -1; }
    /* pack S once, k-major per lane group */
    int64_t ngroups = (nS + TJO_LANES - 1) / TJO_LANES;
    double* St = (double*)malloc((size_t)(ngroups > 0 ? ngroups : 1) * d * TJO_LANES * sizeof(double));
    if (!St) { free(bufs); free(counts); return -1; }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t g = 0; g < ngroups; ++g) pack_lanes(S, nS, d, g * TJO_LANES, St + g * d * TJO_LANES);

#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int64_t b = 0; b < nblocks; ++b) {
        int64_t i0 = b * TJO_BLOCK, i1 = i0 + TJO_BLOCK < nR ? i0 + TJO_BLOCK : nR;
        double d2[TJO_LANES];
        for (int64_t i = i0; i < i1; ++i) {
            const float* x = R + i * d;
            int64_t jstart = self ? i + 1 : 0;
            for (int64_t g = jstart / TJO_LANES; g < ngroups; ++g) {
                d2_lanes(x, St + g * d * TJO_LANES, d, d2);
                for (int l = 0; l < TJO_LANES; ++l) {
                    int64_t j = g * TJO_LANES + l;
                    if (j < jstart || j >= nS) continue;
                    if (d2[l] <= T) {
                        counts[b]++;
                        if (!count_only && pb_push(&bufs[b], i, j)) failed = 1;
                    }
                }
            }
        }
    }
    free(St);
    int64_t total = 0;
    for (int64_t b = 0; b < nblocks; ++b) {
        if (!count_only && !failed) {
            int64_t n = bufs[b].n;
            int64_t room = cap - total;
            if (room > 0 && out) {
                int64_t m = n < room ? n : room;
                memcpy(out + 2 * total, bufs[b].data, (size_t)m * 2 * sizeof(int64_t));
            }
        }
        total += counts[b];
        free(bufs[b].data);
    }
    free(bufs);
    free(counts);
    return failed ? -1 : total;
}

/* Full d2 matrix (small inputs only): out[i*nS + j] = d2_seq(R_i, S_j). */
void tjo_d2_matrix(const float* R, int64_t nR, const float* S, int64_t nS, int64_t d,
                   double* out, int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < nR; ++i)
        for (int64_t j = 0; j < nS; ++j) out[i * nS + j] = tjo_d2(R + i * d, S + j * d, d);
}

/* d2 of an explicit list of pairs (used to audit ambiguous queues). */
void tjo_d2_pairs(const float* R, const float* S, int64_t d, const int64_t* pairs,
                  int64_t n, double* out, int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t p = 0; p < n; ++p)
        out[p] = tjo_d2(R + pairs[2 * p] * d, S + pairs[2 * p + 1] * d, d);
}

/*
 * Uniform-grid join on the CPU (the full-size checker for low-dimensional
 * R x S joins).  Grid layout follows GDS/main.cu:891-950: per-dimension
 * bounds padded outward, nCells = ceil(extent / side), row-major linear cell
 * id with dimension 0 fastest (GDS/main.cu:867-883), and a 3^d neighbour
 * loop (GDS/kernel.cu:132-312).  The cell side is sqrt(T) * (1 + 2^-10) so
 * that every pair with d2_seq <= T sits in adjacent cells despite rounding
 * of the cell coordinates; the predicate itself is the sequential d2 above.
 * Supports d <= 3.  Output is sorted by (i, j).
 */
typedef struct { int64_t key, idx; } kv;
static int kv_cmp(const void* a, const void* b) {
    const kv* x = (const kv*)a; const kv* y = (const kv*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}
static int i64_cmp(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : (x > y);
}

int64_t tjo_grid_join(const float* R, int64_t nR, const float* S, int64_t nS, int64_t d,
                      double T, int self, int64_t* out, int64_t cap, int nthreads) {
    if (d < 1 || d > 3) return -2;
    if (self) { S = R; nS = nR; }
    double side = sqrt(T) * (1.0 + 0x1p-10);
    if (!(side > 0.0)) return -3;
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    int64_t ncell[3] = {1, 1, 1};
    for (int64_t k = 0; k < d; ++k) {
        mn[k] = INFINITY; mx[k] = -INFINITY;
        for (int64_t i = 0; i < nR; ++i) { double v = R[i * d + k]; if (v < mn[k]) mn[k] = v; if (v > mx[k]) mx[k] = v; }
        for (int64_t j = 0; j < nS; ++j) { double v = S[j * d + k]; if (v < mn[k]) mn[k] = v; if (v > mx[k]) mx[k] = v; }
        if (nR + nS == 0) { mn[k] = 0; mx[k] = 0; }
    }
    /* Widen the side until every axis has at most 2^38 cells and the linear
     * id fits 62 bits: (x - lo) / side is then exact to far less than the
     * 2^-10 slack of the side (relative error < 2^-51 of at most 2^38 cells),
     * so a pair closer than sqrt(T) on an axis sits in adjacent cells.  One
     * far outlier would otherwise stretch an axis to ~2^55 cells, where
     * rounding moves points by several cells. */
    for (;;) {
        double prod = 1.0;
        int wide = 0;
        for (int64_t k = 0; k < d; ++k) {
            lo[k] = mn[k] - side; hi[k] = mx[k] + side;
            double nc = ceil((hi[k] - lo[k]) / side) + 1.0;
            if (nc > 0x1p38) wide = 1;
            prod *= nc;
        }
        if (!wide && prod <= 0x1p62) break;
        side *= 2.0;
    }
    for (int64_t k = 0; k < d; ++k) ncell[k] = (int64_t)ceil((hi[k] - lo[k]) / side) + 1;
    kv* cells = (kv*)malloc((size_t)(nS > 0 ? nS : 1) * sizeof(kv));
    if (!cells) return -1;
    for (int64_t j = 0; j < nS; ++j) {
        int64_t key = 0, mul = 1;
        for (int64_t k = 0; k < d; ++k) {
            int64_t c = (int64_t)floor(((double)S[j * d + k] - lo[k]) / side);
            key += c * mul; mul *= ncell[k];
        }
        cells[j].key = key; cells[j].idx = j;
    }
    qsort(cells, (size_t)nS, sizeof(kv), kv_cmp);
    int64_t* cnt = (int64_t*)calloc((size_t)(nR > 0 ? nR : 1), sizeof(int64_t));
    int64_t** rows = (int64_t**)calloc((size_t)(nR > 0 ? nR : 1), sizeof(int64_t*));
    int failed = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < nR; ++i) {
        int64_t c[3] = {0, 0, 0};
        for (int64_t k = 0; k < d; ++k) c[k] = (int64_t)floor(((double)R[i * d + k] - lo[k]) / side);
        int64_t n = 0, rcap = 0;
        int64_t* buf = NULL;
        int64_t lo2 = d > 2 ? -1 : 0, hi2 = d > 2 ? 1 : 0;
        int64_t lo1 = d > 1 ? -1 : 0, hi1 = d > 1 ? 1 : 0;
        for (int64_t o2 = lo2; o2 <= hi2; ++o2)
            for (int64_t o1 = lo1; o1 <= hi1; ++o1)
                for (int64_t o0 = -1; o0 <= 1; ++o0) {
                    int64_t key = (c[0] + o0) + (c[1] + o1) * ncell[0] + (c[2] + o2) * ncell[0] * ncell[1];
                    /* lower_bound on key */
                    int64_t a = 0, b = nS;
                    while (a < b) { int64_t m = (a + b) / 2; if (cells[m].key < key) a = m + 1; else b = m; }
                    for (int64_t p = a; p < nS && cells[p].key == key; ++p) {
                        int64_t j = cells[p].idx;
                        if (self && j <= i) continue;
                        if (tjo_d2(R + i * d, S + j * d, d) <= T) {
                            if (n == rcap) {
                                rcap = rcap ? 2 * rcap : 16;
                                int64_t* nb = (int64_t*)realloc(buf, (size_t)rcap * sizeof(int64_t));
                                if (!nb) { failed = 1; break; }
                                buf = nb;
                            }
                            buf[n++] = j;
                        }
                    }
                }
        if (n > 1) qsort(buf, (size_t)n, sizeof(int64_t), i64_cmp);
        cnt[i] = n; rows[i] = buf;
    }
    int64_t total = 0;
    for (int64_t i = 0; i < nR; ++i) {
        for (int64_t p = 0; p < cnt[i]; ++p) {
            if (out && total < cap) { out[2 * total] = i; out[2 * total + 1] = rows[i][p]; }
            total++;
        }
        free(rows[i]);
    }
    free(rows); free(cnt); free(cells);
    return failed ? -1 : total;
}

int tjo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
