// ipt_noalloc.hip — ipt_transpose_213_noalloc: the 213 transpose without any
// allocation, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid it).
//
// The per-call path keeps its barrier state and scratch staging in arenas it
// allocates and grows.  This entry point uses one file-scope device array
// instead, `scratch` below: its first kNoAllocSyncBytes hold the cooperative
// kernel's barrier state (control block + barrier lines, zero at rest), the
// rest the staging of the decomposition's scratch tiers.  Host code never
// passes the array's name to a kernel (in host code it is only a stand-in
// for the symbol): the first call on a device looks up the device address
// with hipGetSymbolAddress, with that device current, and every later use on
// that device goes through the address kept for it.  Each device holds its
// own copy of the array for the life of the process.
//
// Every call on a device shares that device's array and leaves the barrier
// state at zero; calls that overlap on one device (several streams or host
// threads at once) race on it.
#include "ipt_internal.h"
#include "ipt_kernels.h"
#include "ipt_select.h"

#include <map>
#include <mutex>

namespace ipt {

constexpr size_t kNoAllocBytes = (size_t)32 << 20;      // the whole array
constexpr size_t kNoAllocSyncBytes = (size_t)1 << 20;   // of which barrier state

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Zero n words (the barrier state before its first use on a device): a
// kernel, not a memset, so that a capture records an ordinary kernel node.
__global__ void __launch_bounds__(256) k_noalloc_zero(uint32_t* __restrict__ w, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        w[i] = 0u;
}

struct NoAlloc {
    void* ptr = nullptr;  // device address of `scratch` on this device
    bool zeroed = false;  // the barrier state was cleared by an eager launch
    Arena sync, staging;  // fixed arenas over the array
};
static std::mutex g_na_mu;
static std::map<int, NoAlloc> g_na;  // per device

// The device's state, set up on its first call (hipGetSymbolAddress with the
// device current; the caller made it current).
static ipt_status noalloc_state(const DeviceCtx* ctx, NoAlloc** out) {
    std::lock_guard<std::mutex> g(g_na_mu);
    NoAlloc& na = g_na[ctx->dev];
    if (!na.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch)", e);
            return IPT_ERR_HIP;
        }
        na.ptr = p;
        na.sync.ptr = p;
        na.sync.bytes = kNoAllocSyncBytes;
        na.sync.fixed = true;
        na.staging.ptr = static_cast<unsigned char*>(p) + kNoAllocSyncBytes;
        na.staging.bytes = kNoAllocBytes - kNoAllocSyncBytes;
        na.staging.fixed = true;
    }
    *out = &na;
    return IPT_OK;
}

// Reports (and, after the device went idle, clears) a tripped barrier of the
// no-alloc path on this device; called by ipt_device_check.
ipt_status noalloc_check(const DeviceCtx* ctx, bool* tripped) {
    *tripped = false;
    NoAlloc* na = nullptr;
    {
        std::lock_guard<std::mutex> g(g_na_mu);
        auto it = g_na.find(ctx->dev);
        if (it == g_na.end() || !it->second.ptr) return IPT_OK;
        na = &it->second;
    }
    uint32_t flag = 0;
    hipError_t e = hipMemcpy(&flag, static_cast<uint32_t*>(na->ptr) + 2, sizeof(flag),
                             hipMemcpyDeviceToHost);
    if (e == hipSuccess && flag) {
        *tripped = true;
        e = hipMemset(na->ptr, 0, kNoAllocSyncBytes);
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (e != hipSuccess) {
        set_error("no-alloc barrier state", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

}  // namespace ipt

using namespace ipt;

extern "C" {

ipt_status ipt_transpose_213_noalloc(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                                     void* stream) {
    ipt_status st = validate(data, elem_bytes, d1, d2, d3);
    if (st != IPT_OK) return st;
    DeviceCtx* ctx;
    if ((st = get_ctx(&ctx)) != IPT_OK) return st;
    const hipStream_t hs = (hipStream_t)stream;
    const size_t bytes = (size_t)d1 * (size_t)d2 * (size_t)d3 * (size_t)elem_bytes;
    if ((st = prepare_memory(ctx, data, bytes, hs, true)) != IPT_OK) return st;
    NoAlloc* na;
    if ((st = noalloc_state(ctx, &na)) != IPT_OK) return st;
    const Shape s = make_shape(elem_bytes, d1, d2, d3);
    if (s.d1 == 1 || s.d2 == 1) return IPT_OK;

    // The dispatcher's choice (tuning table, selector or heuristic), with the
    // routes that need allocations replaced: the block-cycle path (move lists)
    // and cooperative groups whose barrier state exceeds the array's share
    // take the decomposition.
    int v;
    Tune t;
    if (selector_loaded()) {
        const Candidate& c = kCandidates[select_candidate(ctx, s)];
        v = c.variant;
        t = resolve_tune(ctx, &c.tune);
    } else {
        v = choose_variant(ctx, s);
        t = resolve_tune(ctx, nullptr);
    }
    const int decomp = s.d1 >= s.d2 ? IPT_VARIANT_C2R : IPT_VARIANT_R2C;
    if (v == IPT_VARIANT_BLOCK_CYCLE) {
        v = decomp;
        t = resolve_tune(ctx, nullptr);
    }
    if (v == IPT_VARIANT_GCS) {
        const GcsGeom gg = gcs_geom(ctx, s, t.lds_cap, true, 3, true);
        const uint64_t need = (32 + bar_words(gg.g) * (uint64_t)s.d3) * sizeof(uint32_t);
        if (!gg.ok || (gg.g > 1 && need > na->sync.bytes)) {
            v = decomp;
            t = resolve_tune(ctx, nullptr);
        }
    }
    if (v == IPT_VARIANT_C2R || v == IPT_VARIANT_R2C) {
        // Refuse before anything is enqueued when even one strip or row of
        // the scratch tiers does not fit the staging; a narrow LDS strip
        // replaces the optional scratch tier when that one would not fit.
        const bool c2r = v == IPT_VARIANT_C2R;
        if (decomp_scratch_need(s, c2r, t) > (int64_t)na->staging.bytes) t.opt_scratch = false;
        const int64_t need = decomp_scratch_need(s, c2r, t);
        if (need > (int64_t)na->staging.bytes) {
            char msg[160];
            snprintf(msg, sizeof(msg),
                     "no-alloc transpose needs %lld bytes of staging, the fixed buffer holds %zu",
                     (long long)need, na->staging.bytes);
            set_message(msg);
            return IPT_ERR_TOO_LARGE;
        }
    }
    t.sync = &na->sync;
    t.scratch = &na->staging;

    if (!na->zeroed) {
        // Clear the barrier state before its first use.  Under stream capture
        // the clearing kernel becomes part of the graph, and the next eager
        // call on the device clears it again.
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(hs, &cs) != hipSuccess) {
            (void)hipGetLastError();
            cs = hipStreamCaptureStatusNone;
        }
        const uint32_t words = (uint32_t)(kNoAllocSyncBytes / sizeof(uint32_t));
        hipLaunchKernelGGL(k_noalloc_zero, dim3(256), dim3(256), 0, hs, (uint32_t*)na->ptr, words);
        IPT_LAUNCH_CHECK("k_noalloc_zero");
        if (cs == hipStreamCaptureStatusNone) {
            std::lock_guard<std::mutex> g(g_na_mu);
            na->zeroed = true;
        }
    }
    return run_variant(ctx, data, s, v, t, hs);
}

}  // extern "C"
