// noalloc.hip -- tj_join without any allocation: for callers that cannot
// allocate device memory (code under hipGraph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// The join's scratch (row metadata, quantized codes, the candidate queue,
// output keys and the sort's ping-pong buffer -- everything a tj_join
// workspace holds) lives in one file-scope device array of
// TJ_NOALLOC_SCRATCH_BYTES.  Each GPU the library runs on holds its own copy
// of the array until the process ends; its device address is looked up once
// per device (hipGetSymbolAddress, with that device current) and kept, and
// only that address ever reaches a kernel or a copy.  The plan is laid out in
// a stack struct, so the call allocates nothing, host or device.
//
// Every call on one GPU shares that GPU's array and it keeps the last call's
// contents: the pipeline clears what must start at zero (stats, the scale
// maximum, the grid's tables) with fill kernels, as tj_join does.  Calls that
// overlap on one GPU (two streams, two threads) race on the array.
#include <mutex>
#include "plan_impl.h"

namespace {

static __device__ __align__(256) uint8_t scratch[TJ_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEVICES];  // device address of `scratch` per device

// The current device's copy of the scratch array (looked up on the first
// call on that device).
int device_scratch(void** out) {
    int dev = 0;
    TJ_HIP(hipGetDevice(&dev));
    TJ_CHECK(dev >= 0 && dev < MAX_DEVICES, TJ_ERR_UNSUPPORTED, "device %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[dev] == nullptr) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return tj::hip_fail(e, "hipGetSymbolAddress(scratch)");
        TJ_CHECK(ptr != nullptr, TJ_ERR_HIP, "hipGetSymbolAddress(scratch) gave NULL");
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return TJ_OK;
}

}  // namespace

extern "C" int tj_join_noalloc(const float* R, int64_t nR, const float* S, int64_t nS, int d, int method,
                               int precision, double T, int64_t out_capacity, int64_t amb_capacity, unsigned flags,
                               void* pairs, int out_format, uint64_t* stats, void* stream) {
    TJ_CHECK((flags & (TJ_FLAG_TIMING | TJ_FLAG_PRUNE)) == 0, TJ_ERR_INVALID,
             "tj_join_noalloc takes TJ_FLAG_AUDIT only (timing needs events, pruning is a batch option)");
    tj_plan p{};
    int rc = tj::plan_init(&p, nR, nS, d, method, precision, out_capacity, amb_capacity, flags);
    if (rc) return rc;
    TJ_CHECK(p.total <= (size_t)TJ_NOALLOC_SCRATCH_BYTES, TJ_ERR_WORKSPACE,
             "this join needs %zu bytes of scratch; tj_join_noalloc holds %zu (smaller inputs or capacities)",
             p.total, (size_t)TJ_NOALLOC_SCRATCH_BYTES);
    void* ws = nullptr;
    if ((rc = device_scratch(&ws))) return rc;
    if ((rc = tj::join_check(&p, R, S, T, ws, pairs, out_format, stats))) return rc;
    const tj::JoinCall c{&p, R, S, T, ws, pairs, out_format, stats, false};
    return tj::join_enqueue(c, (hipStream_t)stream);
}
