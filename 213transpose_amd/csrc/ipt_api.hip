// ipt_api.hip — C ABI, per-device context, dispatcher and launchers.
//
// The dispatcher mirrors inplace::transpose<T> (src/transpose.cu:99-124):
// pick a decomposition for the slice shape, then run its passes as
// stream-ordered kernels.  Like the reference it prefetches managed memory
// to the current device first; unlike it, it never prints, never allocates
// per call (scratch is cached per device), refuses memory the device cannot
// reach instead of faulting, and reports every failure as an ipt_status.
#include "../../include/inplace_transpose.h"
#include "ipt_internal.h"
#include "ipt_select.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace ipt {

// ---------------------------------------------------------------- errors

static thread_local char g_last_error[256] = "";

void set_error(const char* what, hipError_t e) {
    snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
}

void set_message(const char* what) { snprintf(g_last_error, sizeof(g_last_error), "%s", what); }

// ---------------------------------------------------------------- devices

static std::mutex g_ctx_mu;
static std::map<int, std::unique_ptr<DeviceCtx>> g_ctx;

ipt_status get_ctx(DeviceCtx** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("hipGetDevice", e);
        return IPT_ERR_NO_DEVICE;
    }
    std::lock_guard<std::mutex> g(g_ctx_mu);
    auto it = g_ctx.find(dev);
    if (it != g_ctx.end()) {
        *out = it->second.get();
        return IPT_OK;
    }
    std::unique_ptr<DeviceCtx> c(new DeviceCtx());
    c->dev = dev;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) {
        set_error("hipGetDeviceProperties", e);
        return IPT_ERR_NO_DEVICE;
    }
    c->ncu = prop.multiProcessorCount;
    int nxcd = 1;
    if (hipDeviceGetAttribute(&nxcd, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || nxcd < 1) {
        (void)hipGetLastError();
        nxcd = 1;
    }
    c->nxcd = nxcd;
    int pageable = 0;
    if (hipDeviceGetAttribute(&pageable, hipDeviceAttributePageableMemoryAccess, dev) != hipSuccess) {
        (void)hipGetLastError();
        pageable = 0;
    }
    c->pageable = pageable != 0;
    c->sync.zero_at_rest = true;
    c->sync.wd_host = &c->wd_host;  // created with the arena's first buffer
    c->lds_max = (int64_t)prop.sharedMemPerBlock;
    snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
    *out = c.get();
    g_ctx[dev] = std::move(c);
    return IPT_OK;
}

bool watchdog_word(uint32_t** host, uint64_t* dev_addr) {
    *host = nullptr;
    *dev_addr = 0;
    void* h = nullptr;
    void* d = nullptr;
    if (hipHostMalloc(&h, 64, hipHostMallocMapped) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(h);
        return false;
    }
    __atomic_store_n((uint32_t*)h, 0u, __ATOMIC_RELEASE);
    *host = (uint32_t*)h;
    *dev_addr = (uint64_t)(uintptr_t)d;
    return true;
}

hipError_t ctl_init(uint32_t* ctl, uint64_t wd_addr) {
    const uint32_t w[2] = {(uint32_t)wd_addr, (uint32_t)(wd_addr >> 32)};
    return hipMemcpy(ctl + 6, w, sizeof(w), hipMemcpyHostToDevice);
}

static bool stream_capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return cs != hipStreamCaptureStatusNone;
}

// Grow-only arena.  Growth synchronises the device once (the old buffer may
// still be in use by queued kernels) and is refused during stream capture
// (IPT_ERR_UNSUPPORTED: run the shape once outside the capture first).  A
// buffer a capture has used is retired rather than freed (its graph may be
// replayed later); growth at least doubles, so retired buffers stay below
// the live one in total.
ipt_status arena_acquire(Arena& a, size_t bytes, hipStream_t s, void** ptr) {
    if (a.fixed) {
        if (bytes > a.bytes) {
            snprintf(g_last_error, sizeof(g_last_error),
                     "needs %zu bytes of library scratch, the fixed buffer holds %zu", bytes, a.bytes);
            return IPT_ERR_TOO_LARGE;
        }
        *ptr = a.ptr;
        return IPT_OK;
    }
    const bool capturing = stream_capturing(s);
    if (a.bytes < bytes) {
        if (capturing) {
            snprintf(g_last_error, sizeof(g_last_error),
                     "library arena must grow to %zu bytes, not allowed during stream capture: "
                     "run this shape once before capturing it", bytes);
            return IPT_ERR_UNSUPPORTED;
        }
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) { set_error("hipDeviceSynchronize", e); return IPT_ERR_HIP; }
        if (a.ptr) {
            if (a.captured) a.retired.push_back(a.ptr);
            else (void)hipFree(a.ptr);
        }
        const size_t old = a.bytes;
        a.ptr = nullptr;
        a.bytes = 0;
        a.captured = false;
        a.ev_valid = false;
        size_t want = std::max(bytes, 2 * old);
        want = a.zero_at_rest ? std::max(want, (size_t)1 << 20) : std::max(want, (size_t)64 << 20);
        e = hipMalloc(&a.ptr, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            e = hipMalloc(&a.ptr, bytes);
            want = bytes;
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            a.ptr = nullptr;
            set_error("hipMalloc(arena)", e);
            return IPT_ERR_OUT_OF_MEMORY;
        }
        if (a.zero_at_rest) {
            if (a.wd_host && !*a.wd_host) (void)watchdog_word(a.wd_host, &a.wd_addr);
            e = hipMemset(a.ptr, 0, want);
            if (e == hipSuccess && a.wd_addr) e = ctl_init((uint32_t*)a.ptr, a.wd_addr);
            if (e == hipSuccess) e = hipDeviceSynchronize();
            if (e != hipSuccess) { set_error("hipMemset(arena)", e); return IPT_ERR_HIP; }
        }
        a.bytes = want;
        a.last = s;
    }
    if (capturing) a.captured = true;
    if (a.ev_valid && a.last != s) {
        hipError_t e = hipStreamWaitEvent(s, a.ev, 0);
        if (e != hipSuccess) { set_error("hipStreamWaitEvent", e); return IPT_ERR_HIP; }
    }
    *ptr = a.ptr;
    return IPT_OK;
}

// Zero the barrier arena (after the device went idle) and rewrite its fixed
// control words; clears the host watchdog word.  Called with ctx->mu held.
static hipError_t reset_sync_arena(DeviceCtx* ctx) {
    hipError_t e = hipSuccess;
    if (ctx->sync.ptr) {
        e = hipMemset(ctx->sync.ptr, 0, ctx->sync.bytes);
        if (e == hipSuccess && ctx->sync.wd_addr) e = ctl_init((uint32_t*)ctx->sync.ptr, ctx->sync.wd_addr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    if (ctx->wd_host) __atomic_store_n(ctx->wd_host, 0u, __ATOMIC_RELEASE);
    return e;
}

ipt_status watchdog_check(DeviceCtx* ctx) {
    if (!ctx->wd_host || __atomic_load_n(ctx->wd_host, __ATOMIC_ACQUIRE) == 0) return IPT_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (__atomic_load_n(ctx->wd_host, __ATOMIC_ACQUIRE) == 0) return IPT_OK;  // handled meanwhile
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = reset_sync_arena(ctx);
    if (e != hipSuccess) {
        set_error("watchdog reset", e);
        return IPT_ERR_HIP;
    }
    snprintf(g_last_error, sizeof(g_last_error),
             "a cooperative transpose enqueued earlier on device %d gave up waiting at a group "
             "barrier (watchdog): its output is not valid; this call did nothing", ctx->dev);
    return IPT_ERR_HIP;
}

// Event to attach to the next launch that uses the arena on stream s (the
// launch completes it, as arena_release's record would); null while s is
// being captured, where the caller records with arena_release instead.
ipt_status arena_event(Arena& a, hipStream_t s, hipEvent_t* ev) {
    *ev = nullptr;
    if (a.fixed || stream_capturing(s)) return IPT_OK;
    if (!a.ev) {
        hipError_t e = hipEventCreateWithFlags(&a.ev, hipEventDisableTiming);
        if (e != hipSuccess) { set_error("hipEventCreate", e); return IPT_ERR_HIP; }
    }
    a.ev_valid = true;
    a.last = s;
    *ev = a.ev;
    return IPT_OK;
}

ipt_status arena_release(Arena& a, hipStream_t s) {
    if (a.fixed) return IPT_OK;
    if (!a.ev) {
        hipError_t e = hipEventCreateWithFlags(&a.ev, hipEventDisableTiming);
        if (e != hipSuccess) { set_error("hipEventCreate", e); return IPT_ERR_HIP; }
    }
    hipError_t e = hipEventRecord(a.ev, s);
    if (e != hipSuccess) { set_error("hipEventRecord", e); return IPT_ERR_HIP; }
    a.ev_valid = true;
    a.last = s;
    return IPT_OK;
}

// ---------------------------------------------------------------- shapes

Shape make_shape(int esz, int64_t d1, int64_t d2, int64_t d3) {
    Shape s;
    s.esz = esz;
    s.d1 = d1;
    s.d2 = d2;
    s.d3 = d3;
    return s;
}

ipt_status validate(const void* data, int esz, int64_t d1, int64_t d2, int64_t d3) {
    if (!data) return IPT_ERR_INVALID_ARG;
    if (esz != 4 && esz != 8) return IPT_ERR_INVALID_ARG;
    if (d1 <= 0 || d2 <= 0 || d3 <= 0) return IPT_ERR_INVALID_ARG;
    // Total bytes must fit comfortably in 63 bits, and one slice in 2^40
    // elements (the 64-bit plan keeps every numerator below 2^63).
    unsigned __int128 slice = (unsigned __int128)d1 * (unsigned __int128)d2;
    unsigned __int128 total = slice * (unsigned __int128)d3 * (unsigned)esz;
    if (slice > ((unsigned __int128)1 << 40) || total > ((unsigned __int128)1 << 62))
        return IPT_ERR_TOO_LARGE;
    if ((uintptr_t)data % (uintptr_t)esz) return IPT_ERR_INVALID_ARG;
    return IPT_OK;
}

// Memory kind of a pointer: hipMemoryType's value (unified memory counts as
// managed); pageable host memory when HIP does not know the pointer, whether
// it says so with an error or with hipMemoryTypeUnregistered.
static int memory_kind(const void* data, int* device) {
    hipPointerAttribute_t at{};
    if (device) *device = -1;
    if (hipPointerGetAttributes(&at, data) != hipSuccess) {
        (void)hipGetLastError();
        return IPT_MEM_PAGEABLE;
    }
    if (at.isManaged || at.type == hipMemoryTypeManaged || at.type == hipMemoryTypeUnified)
        return IPT_MEM_MANAGED;
    if (at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeArray) {
        if (device) *device = at.device;
        return IPT_MEM_DEVICE;
    }
    if (at.type == hipMemoryTypeHost) return IPT_MEM_HOST;
    return IPT_MEM_PAGEABLE;
}

// The reference prefetches managed memory to the current device before its
// passes (src/transpose.cu:100-104) and exits on any other pointer.  Here
// every pointer the device reaches is accepted; one it does not reach is
// refused before anything is enqueued, since a kernel touching it would
// fault.  A prefetch that the system cannot honour leaves the data where it
// is, still reachable, and is not an error.
ipt_status prepare_memory(const DeviceCtx* ctx, const void* data, size_t bytes, hipStream_t s,
                          bool prefetch, int* kind) {
    int owner = -1;
    const int k = memory_kind(data, &owner);
    if (kind) *kind = k;
    switch (k) {
    case IPT_MEM_DEVICE:
        if (owner != ctx->dev) {
            snprintf(g_last_error, sizeof(g_last_error),
                     "data is device memory of device %d, the current device is %d", owner, ctx->dev);
            return IPT_ERR_INVALID_ARG;
        }
        return IPT_OK;
    case IPT_MEM_HOST:
        return IPT_OK;
    case IPT_MEM_MANAGED: {
        if (!prefetch || bytes == 0) return IPT_OK;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
            (void)hipGetLastError();
            return IPT_OK;
        }
        if (cs != hipStreamCaptureStatusNone) return IPT_OK;
        if (hipMemPrefetchAsync(data, bytes, ctx->dev, s) != hipSuccess) (void)hipGetLastError();
        return IPT_OK;
    }
    default:
        if (ctx->pageable) return IPT_OK;
        snprintf(g_last_error, sizeof(g_last_error),
                 "data is pageable host memory, which device %d cannot access", ctx->dev);
        return IPT_ERR_INVALID_ARG;
    }
}

// IPT_VARIANT_AUTO: the learned selector's candidate when a selector is
// loaded and no tune is forced, else the heuristic's variant.  When the
// block-cycle path is chosen but cannot run (unaligned data, more saved
// chunk heads than the scratch cap, or stream capture before the shape's
// first use) the decomposition runs instead.
ipt_status run_auto(DeviceCtx* ctx, void* data, const Shape& s, const ipt_tune* tune,
                    hipStream_t st) {
    int v;
    Tune t;
    if (!tune && selector_loaded()) {
        const Candidate& c = kCandidates[select_candidate(ctx, s)];
        v = c.variant;
        t = resolve_tune(ctx, &c.tune);
    } else {
        v = choose_variant(ctx, s);
        t = resolve_tune(ctx, tune);
    }
    if (v == IPT_VARIANT_BLOCK_CYCLE) {
        const ipt_status r = run_variant(ctx, data, s, v, t, st);
        if (r != IPT_ERR_UNSUPPORTED) return r;
        v = s.d1 >= s.d2 ? IPT_VARIANT_C2R : IPT_VARIANT_R2C;
        t = resolve_tune(ctx, tune);
    }
    return run_variant(ctx, data, s, v, t, st);
}

}  // namespace ipt

using namespace ipt;

// ---------------------------------------------------------------- C ABI

extern "C" {

const char* ipt_status_string(ipt_status s) {
    switch (s) {
    case IPT_OK: return "ok";
    case IPT_ERR_INVALID_ARG: return "invalid argument";
    case IPT_ERR_UNSUPPORTED: return "variant unsupported for this shape";
    case IPT_ERR_TOO_LARGE: return "size beyond the index plan";
    case IPT_ERR_HIP: return "HIP runtime error";
    case IPT_ERR_NO_DEVICE: return "no usable GPU";
    case IPT_ERR_OUT_OF_MEMORY: return "out of device memory";
    case IPT_ERR_IO: return "file missing or corrupt";
    case IPT_ERR_NOT_READY: return "object not ready";
    }
    return "unknown status";
}

const char* ipt_last_error(void) { return g_last_error; }

int ipt_memory_kind(const void* data, int* device) {
    if (device) *device = -1;
    if (!data) return IPT_MEM_PAGEABLE;
    return memory_kind(data, device);
}

ipt_status ipt_transpose_variant(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                                 int variant, const ipt_tune* tune, void* stream) {
    ipt_status st = validate(data, elem_bytes, d1, d2, d3);
    if (st != IPT_OK) return st;
    if (variant < 0 || variant >= IPT_VARIANT_COUNT) return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    if ((st = get_ctx(&ctx)) != IPT_OK) return st;
    if ((st = watchdog_check(ctx)) != IPT_OK) return st;
    const size_t bytes = (size_t)d1 * (size_t)d2 * (size_t)d3 * (size_t)elem_bytes;
    if ((st = prepare_memory(ctx, data, bytes, (hipStream_t)stream, true)) != IPT_OK) return st;
    Shape s = make_shape(elem_bytes, d1, d2, d3);
    if (variant == IPT_VARIANT_AUTO) return run_auto(ctx, data, s, tune, (hipStream_t)stream);
    return run_variant(ctx, data, s, variant, resolve_tune(ctx, tune), (hipStream_t)stream);
}

ipt_status ipt_transpose_213(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                             void* stream) {
    return ipt_transpose_variant(data, elem_bytes, d1, d2, d3, IPT_VARIANT_AUTO, nullptr, stream);
}

ipt_status ipt_c2r_transpose(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                             void* stream) {
    return ipt_transpose_variant(data, elem_bytes, d1, d2, d3, IPT_VARIANT_C2R, nullptr, stream);
}

ipt_status ipt_r2c_transpose(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                             void* stream) {
    // Reference argument order: slices of d1 rows x d2 columns, i.e. the 213
    // transpose of (d1' = d2, d2' = d1) done with the R2C view R = d2, C = d1.
    return ipt_transpose_variant(data, elem_bytes, d2, d1, d3, IPT_VARIANT_R2C, nullptr, stream);
}

ipt_status ipt_transpose(int row_major, void* data, int elem_bytes, int64_t m, int64_t n,
                         void* stream) {
    if (row_major) return ipt_transpose_213(data, elem_bytes, n, m, 1, stream);
    return ipt_transpose_213(data, elem_bytes, m, n, 1, stream);
}

// Permutations of rank N <= kMaxRank in place.  A move exchanges two
// adjacent groups of positions B = [i, j) and C = [j, k) (fastest first),
// the positions before (A) kept innermost and those after (D) outermost:
// one in-place pass,
//   A = 1            the 213 transpose of (|B|, |C|, |D|);
//   |A|*esz = 4, 8   the same on super-elements of |A| elements;
//   otherwise        a transpose of |D| matrices of |C| x |B| chunks of |A|
//                    elements (the block-cycle chunk step, chunks of any
//                    size moved in parts; IPT_ERR_UNSUPPORTED past its
//                    chunk-count and head-scratch limits).
// A move with |B| = 1 or |C| = 1 moves no data.  The layout state (cur[p] =
// the source dim at position p, the convention of `perm`) is searched
// breadth-first, moving no data counting zero, for the fewest passes: rank
// 3 needs one for every permutation but the reversal, which takes two.
static constexpr int kMaxRank = 6;

static ipt_status permute_move(DeviceCtx* ctx, void* data, int esz, const int64_t* sizes,
                               int rank, const int* cur, int i, int j, int k, void* stream) {
    int64_t g[4] = {1, 1, 1, 1};  // |A|, |B|, |C|, |D|
    for (int p = 0; p < rank; p++) g[p < i ? 0 : p < j ? 1 : p < k ? 2 : 3] *= sizes[cur[p]];
    if (g[1] == 1 || g[2] == 1) return IPT_OK;
    if (g[0] == 1) return ipt_transpose_213(data, esz, g[1], g[2], g[3], stream);
    const int64_t e = g[0] * esz;
    if (e == 4 || (e == 8 && (uintptr_t)data % 8 == 0))
        return ipt_transpose_213(data, (int)e, g[1], g[2], g[3], stream);
    return run_chunk_transpose(ctx, data, esz, g[0], g[2], g[1], g[3], (hipStream_t)stream);
}

ipt_status ipt_permute(void* data, int elem_bytes, int rank, const int64_t* dims, const int* perm,
                       void* stream) {
    if (!data || !dims || !perm || rank < 1 || rank > kMaxRank) return IPT_ERR_INVALID_ARG;
    int64_t total = 1;
    int seen[kMaxRank] = {0};
    for (int p = 0; p < rank; p++) {
        if (dims[p] <= 0 || perm[p] < 0 || perm[p] >= rank || seen[perm[p]]) return IPT_ERR_INVALID_ARG;
        seen[perm[p]] = 1;
        if (total > ((int64_t)1 << 62) / dims[p]) return IPT_ERR_TOO_LARGE;
        total *= dims[p];
    }
    ipt_status st = validate(data, elem_bytes, total, 1, 1);
    if (st != IPT_OK) return st;
    // States: permutations of rank positions, coded in base `rank`.
    auto encode = [rank](const int* c) {
        int v = 0;
        for (int p = 0; p < rank; p++) v = v * rank + c[p];
        return v;
    };
    int states = 1;
    for (int p = 0; p < rank; p++) states *= rank;
    std::vector<int> dist(states, -1), prev(states, -1), mi(states), mj(states), mk(states);
    std::vector<std::vector<int>> cfg(states);
    std::deque<int> dq;
    std::vector<int> start(rank);
    for (int p = 0; p < rank; p++) start[p] = p;
    const int s0 = encode(start.data());
    dist[s0] = 0;
    cfg[s0] = start;
    dq.push_back(s0);
    while (!dq.empty()) {  // 0-1 breadth-first search
        const int sc = dq.front();
        dq.pop_front();
        const std::vector<int> c = cfg[sc];
        for (int i = 0; i < rank; i++)
            for (int j = i + 1; j < rank; j++)
                for (int k = j + 1; k <= rank; k++) {
                    std::vector<int> n;
                    n.insert(n.end(), c.begin(), c.begin() + i);
                    n.insert(n.end(), c.begin() + j, c.begin() + k);
                    n.insert(n.end(), c.begin() + i, c.begin() + j);
                    n.insert(n.end(), c.begin() + k, c.end());
                    int64_t gb = 1, gc = 1;
                    for (int p = i; p < j; p++) gb *= dims[c[p]];
                    for (int p = j; p < k; p++) gc *= dims[c[p]];
                    const int w = (gb == 1 || gc == 1) ? 0 : 1;
                    const int sn = encode(n.data());
                    if (dist[sn] >= 0 && dist[sn] <= dist[sc] + w) continue;
                    dist[sn] = dist[sc] + w;
                    prev[sn] = sc;
                    mi[sn] = i;
                    mj[sn] = j;
                    mk[sn] = k;
                    cfg[sn] = n;
                    if (w == 0) dq.push_front(sn);
                    else dq.push_back(sn);
                }
    }
    const int target = encode(perm);
    if (dist[target] < 0) return IPT_ERR_INVALID_ARG;
    std::vector<int> path;
    for (int sc = target; sc != s0; sc = prev[sc]) path.push_back(sc);
    DeviceCtx* ctx;
    st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    if ((st = watchdog_check(ctx)) != IPT_OK) return st;
    st = prepare_memory(ctx, data, (size_t)total * (size_t)elem_bytes, (hipStream_t)stream, true);
    if (st != IPT_OK) return st;
    std::vector<int> cur = start;
    for (auto it = path.rbegin(); it != path.rend(); ++it) {
        st = permute_move(ctx, data, elem_bytes, dims, rank, cur.data(), mi[*it], mj[*it], mk[*it],
                          stream);
        if (st != IPT_OK) return st;
        cur = cfg[*it];
    }
    return IPT_OK;
}

ipt_status ipt_permute3(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                        const int* perm, void* stream) {
    const int64_t dims[3] = {d1, d2, d3};
    return ipt_permute(data, elem_bytes, 3, dims, perm, stream);
}

int ipt_variant_feasible(int variant, int elem_bytes, int64_t d1, int64_t d2, int64_t d3) {
    if (elem_bytes != 4 && elem_bytes != 8) return 0;
    if (d1 <= 0 || d2 <= 0 || d3 <= 0) return 0;
    DeviceCtx* ctx;
    if (get_ctx(&ctx) != IPT_OK) return 0;
    return variant_feasible(ctx, make_shape(elem_bytes, d1, d2, d3), variant) ? 1 : 0;
}

ipt_status ipt_plan_query(int elem_bytes, int64_t d1, int64_t d2, int64_t d3, ipt_plan_info* out) {
    if (!out) return IPT_ERR_INVALID_ARG;
    if (elem_bytes != 4 && elem_bytes != 8) return IPT_ERR_INVALID_ARG;
    if (d1 <= 0 || d2 <= 0 || d3 <= 0) return IPT_ERR_INVALID_ARG;
    Shape s = make_shape(elem_bytes, d1, d2, d3);
    // View of the heuristic's decomposition: C2R keeps (R, C) = (d2, d1),
    // R2C uses (d1, d2) (src/transpose.cu:116-123).
    const bool c2r = d1 >= d2;
    HostPlan h = HostPlan::make(c2r ? d2 : d1, c2r ? d1 : d2);
    out->R = h.R; out->C = h.C; out->c = h.c; out->a = h.a; out->b = h.b; out->k = h.k; out->q = h.q;
    out->index_bits = h.fits_index32() ? 32 : h.fits_index64() ? 64 : 0;
    DeviceCtx* ctx = nullptr;
    out->heuristic_variant = (get_ctx(&ctx) == IPT_OK) ? choose_variant(ctx, s)
                                                       : choose_variant(nullptr, s);
    return IPT_OK;
}

ipt_status ipt_debug_eval_map(int map_id, int64_t R, int64_t C, int index_bits, const int64_t* i,
                              const int64_t* j, int64_t n, int64_t* out) {
    if (R <= 0 || C <= 0 || n < 0 || (n > 0 && (!i || !j || !out))) return IPT_ERR_INVALID_ARG;
    if (map_id < 0 || map_id > MAP_R2C_SHUFFLE) return IPT_ERR_INVALID_ARG;
    HostPlan h = HostPlan::make(R, C);
    if (index_bits == 32) {
        if (!h.fits_index32()) return IPT_ERR_UNSUPPORTED;
        return eval_map_host<uint32_t>(map_id, h, i, j, n, out);
    }
    if (index_bits != 64) return IPT_ERR_INVALID_ARG;
    if (!h.fits_index64()) return IPT_ERR_TOO_LARGE;
    return eval_map_host<uint64_t>(map_id, h, i, j, n, out);
}

ipt_status ipt_debug_run_pass(void* data, int elem_bytes, int64_t R, int64_t C, int64_t d3,
                              int map_id, const ipt_tune* tune, void* stream) {
    ipt_status st = validate(data, elem_bytes, C, R, d3);
    if (st != IPT_OK) return st;
    if (map_id < 0 || map_id > MAP_R2C_SHUFFLE) return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    if ((st = get_ctx(&ctx)) != IPT_OK) return st;
    if ((st = watchdog_check(ctx)) != IPT_OK) return st;
    st = prepare_memory(ctx, data, (size_t)R * (size_t)C * (size_t)d3 * (size_t)elem_bytes,
                        (hipStream_t)stream, true);
    if (st != IPT_OK) return st;
    return run_pass(ctx, data, elem_bytes, R, C, d3, map_id, resolve_tune(ctx, tune),
                    (hipStream_t)stream);
}

ipt_status ipt_debug_copy(void* dst, const void* src, size_t bytes, void* stream) {
    if (!dst || !src) return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    if ((st = prepare_memory(ctx, dst, bytes, (hipStream_t)stream, true)) != IPT_OK) return st;
    if ((st = prepare_memory(ctx, src, bytes, (hipStream_t)stream, true)) != IPT_OK) return st;
    return run_copy(ctx, dst, src, bytes, (hipStream_t)stream);
}

ipt_status ipt_transpose_oop(void* dst, const void* src, int elem_bytes, int64_t d1, int64_t d2,
                             int64_t d3, void* stream) {
    ipt_status st = validate(src, elem_bytes, d1, d2, d3);
    if (st != IPT_OK) return st;
    if ((st = validate(dst, elem_bytes, d1, d2, d3)) != IPT_OK) return st;
    const size_t bytes = (size_t)d1 * (size_t)d2 * (size_t)d3 * (size_t)elem_bytes;
    const uintptr_t a = (uintptr_t)dst, b = (uintptr_t)src;
    if (a < b + bytes && b < a + bytes) return IPT_ERR_INVALID_ARG;  // must not overlap
    DeviceCtx* ctx;
    if ((st = get_ctx(&ctx)) != IPT_OK) return st;
    if ((st = prepare_memory(ctx, dst, bytes, (hipStream_t)stream, true)) != IPT_OK) return st;
    if ((st = prepare_memory(ctx, src, bytes, (hipStream_t)stream, true)) != IPT_OK) return st;
    return run_oop(ctx, dst, src, make_shape(elem_bytes, d1, d2, d3), (hipStream_t)stream);
}

ipt_status ipt_device_check(void) {
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) { set_error("hipDeviceSynchronize", e); return IPT_ERR_HIP; }
    bool na_tripped = false;
    if ((st = noalloc_check(ctx, &na_tripped)) != IPT_OK) return st;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (!ctx->sync.ptr) {
        if (na_tripped) set_message("group barrier watchdog tripped (no-alloc path)");
        return na_tripped ? IPT_ERR_HIP : IPT_OK;
    }
    uint32_t flag = 0;
    e = hipMemcpy(&flag, (uint32_t*)ctx->sync.ptr + 2, sizeof(flag), hipMemcpyDeviceToHost);
    if (e != hipSuccess) { set_error("hipMemcpy(sync flag)", e); return IPT_ERR_HIP; }
    const bool host_flag = ctx->wd_host && __atomic_load_n(ctx->wd_host, __ATOMIC_ACQUIRE) != 0;
    if (flag || host_flag || na_tripped) {
        // A group barrier gave up waiting: restore the zero-at-rest state.
        (void)reset_sync_arena(ctx);
        snprintf(g_last_error, sizeof(g_last_error), "group barrier watchdog tripped");
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

ipt_status ipt_debug_set_watchdog(uint32_t polls, uint32_t extra) {
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    std::lock_guard<std::mutex> g(ctx->mu);
    void* sync;
    st = arena_acquire(ctx->sync, 32 * sizeof(uint32_t), nullptr, &sync);
    if (st != IPT_OK) return st;
    const uint32_t w[2] = {polls, extra};
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy((uint32_t*)sync + 3, w, sizeof(w), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error("ipt_debug_set_watchdog", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

ipt_status ipt_debug_chunk_moves(int64_t rc, int64_t cc, int64_t ell, int builder,
                                 int64_t counts[3], uint32_t* items, uint32_t* msrc,
                                 uint32_t* mdst, uint32_t* hpos, int64_t cap) {
    if (rc < 1 || cc < 1 || rc * cc >= kBcMaxChunks || ell < 1 || ell > kBcMaxSeg ||
        (builder != 0 && builder != 1) || !counts || cap < 0)
        return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    ChunkMoves m;
    if ((st = chunk_moves_debug(rc, cc, ell, builder == 1, &m)) != IPT_OK) return st;
    counts[0] = m.nitems;
    counts[1] = m.nmoves;
    counts[2] = m.nheads;
    if ((int64_t)m.nitems > cap || (int64_t)m.nmoves > cap || (int64_t)m.nheads > cap) {
        chunk_moves_free(&m);
        return IPT_ERR_TOO_LARGE;
    }
    hipError_t e = hipSuccess;
    if (m.nitems) e = hipMemcpy(items, m.items, (size_t)m.nitems * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && m.nmoves) e = hipMemcpy(msrc, m.msrc, (size_t)m.nmoves * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && m.nmoves) e = hipMemcpy(mdst, m.mdst, (size_t)m.nmoves * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && m.nheads) e = hipMemcpy(hpos, m.hpos, (size_t)m.nheads * 4, hipMemcpyDeviceToHost);
    chunk_moves_free(&m);
    if (e != hipSuccess) {
        set_error("hipMemcpy", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

ipt_status ipt_debug_arena_info(int which, void** ptr, size_t* bytes, int64_t* retired) {
    if ((which != 0 && which != 1) || !ptr || !bytes || !retired) return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    std::lock_guard<std::mutex> g(ctx->mu);
    const Arena& a = which == 0 ? ctx->sync : ctx->scratch;
    *ptr = a.ptr;
    *bytes = a.bytes;
    *retired = (int64_t)a.retired.size();
    return IPT_OK;
}

ipt_status ipt_debug_arena_grow(int which, size_t bytes) {
    if (which != 0 && which != 1) return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    std::lock_guard<std::mutex> g(ctx->mu);
    Arena& a = which == 0 ? ctx->sync : ctx->scratch;
    void* p;
    st = arena_acquire(a, bytes, nullptr, &p);
    if (st != IPT_OK) return st;
    return arena_release(a, nullptr);
}

const char* ipt_version(void) {
    static char buf[128];
    DeviceCtx* ctx;
    if (get_ctx(&ctx) != IPT_OK) {
        snprintf(buf, sizeof(buf), "inplace_hip (no device)");
    } else {
        snprintf(buf, sizeof(buf), "inplace_hip %s lds=%lld cus=%d", ctx->arch,
                 (long long)ctx->lds_max, ctx->ncu);
    }
    return buf;
}

}  // extern "C"
