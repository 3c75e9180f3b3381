// certify.hip -- the certificate pass of a single join (certify_impl.h).
#include "certify_impl.h"

namespace tjcert {

// Resolves a pair key of one join: both rows' metadata and FP32 rows, and the
// join's parameter block.
struct SingleSrc {
    const float* R; const float* S; int d; int jbits;
    const RowMeta* ma; const RowMeta* mb; const JoinParams* pp;
    __device__ inline void fetch(uint64_t key, PairRef& r) const {
        const uint64_t i = key >> jbits, j = key & ((1ull << jbits) - 1);
        r.P = pp;
        r.mi = ma[i];
        r.mj = mb[j];
        r.xi = R + i * (uint64_t)d;
        r.xj = S + j * (uint64_t)d;
        r.d = d;
    }
};

}  // namespace tjcert

namespace tj {
int launch_certify(const float* R, const float* S, int d, int jbits, const uint64_t* cand_keys, const float* cand_p,
                   int64_t cand_cap, const RowMeta* ma, const RowMeta* mb, const JoinParams* p, uint64_t* out_keys,
                   int64_t out_cap, uint64_t* stats, int audit, hipStream_t s) {
    if (cand_cap <= 0) return TJ_OK;
    int64_t blocks = (cand_cap + 255) / 256;
    const int64_t maxb = (int64_t)num_cus() * 8;
    if (blocks > maxb) blocks = maxb;
    const tjcert::SingleSrc src{R, S, d, jbits, ma, mb, p};
    hipLaunchKernelGGL(tjcert::certify_kernel<tjcert::SingleSrc>, dim3((unsigned)blocks), dim3(256), 0, s, src,
                       cand_keys, cand_p, cand_cap, out_keys, out_cap, stats, audit);
    TJ_HIP(hipGetLastError());
    return TJ_OK;
}
}  // namespace tj
