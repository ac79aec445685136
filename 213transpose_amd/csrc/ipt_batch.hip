// ipt_batch.hip — batched uneven transposes: one persistent launch over a
// device work-item table, captured once into a hipGraph.
//
// No reference counterpart: the reference transposes one tensor per call
// (src/transpose.cu:99-124), launching several kernels each time.  Here a
// batch of instances of any shapes is planned once (items, barrier state,
// launch geometry), and every run is one graph launch.
#include "ipt_internal.h"
#include <hip/hip_ext.h>
#include "ipt_kernels.h"
#include "ipt_select.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <new>
#include <vector>

namespace ipt {

// Work queues.  Items sit in nq + 2 segments of one table, seg[q] .. seg[q+1]:
//   queue 0       groups of the cooperative kernel too wide for one XCD's
//                 share of the grid (taken first, by every workgroup);
//   queue 1 + x   groups placed on XCD x (workgroup i runs on XCD i % nq):
//                 the members of a group then share an L2, which halves the
//                 cost of their stores (each writes lines another just read);
//   queue nq + 1  items with no traffic between workgroups (whole slices,
//                 tile pairs), shared by all.
// A workgroup drains queue 0, its own XCD's queue, the shared queue, then
// the other XCDs' queues.  Taking from a queue is one atomic on its counter
// ctl[kQueueBase + q], so items of uneven cost balance across the grid.
// Deadlock freedom: within a queue items are taken in order, so only its
// head group can be partly taken; once any workgroup has left queue 0 every
// group there is fully taken (and completes).  A stuck state needs every
// workgroup waiting in a partly taken group, at most one per XCD queue, each
// with at most g - 1 members: impossible while nq * (g - 1) < grid, which the
// host guarantees for every group it places in an XCD queue (queue 0 needs
// only g <= grid).  ctl[kQueueDone] counts finished workgroups; the last one
// zeroes the counters (zero at rest, so graph replays start empty).
constexpr int kQueueDone = 9, kQueueBase = 16;
constexpr uint32_t kMaxXcdQueues = 8;

__global__ void __launch_bounds__(kBatchBlock)
k_batch(const BatchInst* __restrict__ insts, const BatchItem* __restrict__ items,
        const uint32_t* __restrict__ seg, uint32_t nq, uint32_t* __restrict__ ctl,
        uint32_t* __restrict__ bars) {
    IPT_DYN_LDS(unsigned char, lds);
    __shared__ uint32_t next;
    const uint32_t own = blockIdx.x % nq;
    for (uint32_t step = 0; step < nq + 2; step++) {
        const uint32_t q = step == 0 ? 0u
                         : step == 1 ? 1 + own
                         : step == 2 ? nq + 1
                                     : 1 + (own + step - 2) % nq;
        const uint32_t lo = seg[q], hi = seg[q + 1];
        if (lo == hi) continue;
        // ahead: a take made (thread 0) as soon as the current cooperative
        // item's barrier released the member, so the atomic's round trip
        // overlaps the item's stores.  The deadlock argument is unchanged: a
        // member is released once the members it waits for have arrived, and
        // every member arrives before it waits; by the time a workgroup takes
        // any later item, every slot of its current group has been taken
        // (earlier in the same queue, or the queue is exhausted), so the
        // members still waiting in that group wait only for taken members.
        uint32_t ahead = 0;
        bool have_ahead = false;
        auto take_ahead = [&]() {
            if (threadIdx.x == 0) {
                ahead = lo + atomic_add_agent(&ctl[kQueueBase + q], 1u);
                have_ahead = true;
            }
        };
        for (;;) {
            if (threadIdx.x == 0) {
                next = have_ahead ? ahead : lo + atomic_add_agent(&ctl[kQueueBase + q], 1u);
                have_ahead = false;
            }
            __syncthreads();
            const uint32_t s = next;
            __syncthreads();  // `next` is rewritten by the next take
            if (s >= hi) break;
            const BatchItem it = items[s];
            const BatchInst I = insts[it.inst];
            if (I.esz == 4)
                batch_item<uint32_t>(I, it, ctl, bars, reinterpret_cast<uint32_t*>(lds), take_ahead);
            else
                batch_item<uint64_t>(I, it, ctl, bars, reinterpret_cast<uint64_t*>(lds), take_ahead);
            __syncthreads();  // LDS is reused by the next item
        }
    }
    if (threadIdx.x == 0 && atomic_add_agent(&ctl[kQueueDone], 1u) == gridDim.x - 1) {
        for (uint32_t q = 0; q < nq + 2; q++) atomic_store_agent(&ctl[kQueueBase + q], 0u);
        atomic_store_agent(&ctl[kQueueDone], 0u);
    }
}

}  // namespace ipt

using namespace ipt;

struct ipt_batch_s {
    DeviceCtx* ctx = nullptr;
    BatchInst* d_insts = nullptr;
    BatchItem* d_items = nullptr;
    uint32_t* d_seg = nullptr;  // queue segments of d_items (nq + 3 words)
    uint32_t nq = 1;            // XCD queues
    uint32_t* d_sync = nullptr;  // 32 control words + barrier state (zero at rest)
    uint64_t sync_words = 0;
    uint32_t* wd_host = nullptr;  // host-mapped watchdog word of this batch's barriers
    uint64_t wd_addr = 0;         // its device address (control words 6, 7)
    uint64_t nitems = 0;
    size_t lds = 0;
    int grid = 0;
    struct Big {
        void* data;
        Shape s;
        int variant;
        Tune tune;
    };
    std::vector<Big> big;  // instances run by the per-call launchers
    // Barrier state of the separate cooperative launches among `big`: fixed
    // (allocated and zeroed when the batch is planned, never grown, so a
    // captured run needs no allocation), wired to this batch's watchdog word.
    Arena big_sync;
    // Chunk steps of block-cycle instances (their blocks are k_batch items).
    struct Chunk {
        void* data;
        uint64_t nslices, slice_vecs, group;
        BlockCyclePlan plan;
        size_t moves;  // index into `moves`
    };
    std::vector<Chunk> pre, post;  // before / after k_batch
    std::vector<ChunkMoves> moves;  // owned by the batch
    void* d_head = nullptr;         // saved heads, shared by the chunk steps in turn
    // Managed instances (data, bytes), prefetched to the device before each run.
    std::vector<std::pair<void*, size_t>> managed;
    Tune tune{};
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ipt_batch_info info{};
    std::vector<int32_t> routes;  // ipt_batch_route per descriptor
    // Timing of the persistent kernel in eager runs (ipt_batch_set_timing):
    // events the launch itself records around k_batch on the run's stream.
    bool timing = false;
    hipEvent_t k0 = nullptr, k1 = nullptr;
    bool timed = false;  // the last eager run recorded them
};

namespace {

enum class Route { Skip, Slice, Gcs, Swap, Big };

// Batch heuristic for one instance.  `forced` is the caller's variant
// (IPT_VARIANT_AUTO = choose).
Route plan_instance(const DeviceCtx* ctx, const Shape& s, int forced, int64_t lds_cap, int grid,
                   const Tune& t, int* big_variant) {
    if (s.d1 == 1 || s.d2 == 1) return Route::Skip;
    const int64_t slice_bytes = s.d2 * slice_pitch(s.d1) * s.esz;
    const GcsGeom gg = gcs_geom(ctx, s, lds_cap, false, kBatchJrClass, false, true);
    const bool gcs_ok = gg.ok && (int64_t)gg.g <= grid && gg.lds <= (size_t)lds_cap;
    switch (forced) {
    case IPT_VARIANT_SLICE_LDS: return slice_bytes <= lds_cap ? Route::Slice : Route::Big;
    case IPT_VARIANT_GCS:
        if (gcs_ok) return Route::Gcs;
        *big_variant = IPT_VARIANT_GCS;  // a separate launch, if one runs it
        return Route::Big;
    case IPT_VARIANT_SQUARE_SWAP: return s.d1 == s.d2 ? Route::Swap : Route::Big;
    case IPT_VARIANT_C2R:
    case IPT_VARIANT_R2C:
    case IPT_VARIANT_BLOCK_CYCLE: *big_variant = forced; return Route::Big;
    default: break;
    }
    if (slice_bytes <= lds_cap) return Route::Slice;
    if (s.d1 == s.d2) return Route::Swap;  // faster than the cooperative route on squares
    if (gcs_ok) return Route::Gcs;
    // Too large for one group of the batch kernel: one separate cooperative
    // launch (one sweep, the per-call kernel's larger register tiles) where
    // that fits; else the block-cycle path with its blocks folded into the
    // batch (two sweeps) when a plan fits, which the caller tries next; else
    // C2R / R2C as separate launches.
    // (uneven 4096-matrix batch: three f64 instances of 88-96 MB as separate
    // launches instead of folded block-cycle, graph 4.23 -> 4.08 ms)
    *big_variant = gcs_geom(ctx, s, lds_max_of(ctx), true, 3, true).ok ? IPT_VARIANT_GCS
                                                                       : IPT_VARIANT_BLOCK_CYCLE;
    (void)t;
    return Route::Big;
}

// Candidate for an AUTO instance of a batch: a measured choice if the
// autotuner recorded one; else, with predictions y, the fastest predicted
// route of the persistent kernel (whole slice, cooperative group, square
// swap) when any fits, and only otherwise the fastest predicted candidate
// with separate launches (one cooperative launch of the per-call kernel,
// block cycle, C2R, R2C: such launches cost more inside a batch than their
// standalone timings show).  -1 = let plan_instance's heuristic decide.
int batch_choice(const DeviceCtx* ctx, const Shape& s, const float* y, int64_t lds_cap, int grid) {
    const int tuned = tuned_candidate(ctx, s);
    if (tuned >= 0) return tuned;
    if (!y || !selector_domain(s.d1, s.d2, s.d3, s.esz) || s.d1 == s.d2) return -1;
    const int64_t slice_bytes = s.d2 * slice_pitch(s.d1) * s.esz;
    const GcsGeom gg = gcs_geom(ctx, s, lds_cap, false, kBatchJrClass, false, true);
    const bool route_ok[4] = {slice_bytes <= lds_cap, s.d1 == s.d2,
                              gg.ok && (int64_t)gg.g <= grid, gg.ok && (int64_t)gg.g <= grid};
    int best = -1;
    for (int k = 0; k < 4; k++)
        if (route_ok[k] && (best < 0 || y[k] < y[best])) best = k;
    if (best >= 0) return best;
    for (int k = 2; k < kNumCandidates; k++)  // 2, 3: one separate cooperative launch
        if (candidate_ok(ctx, s, k) && (best < 0 || y[k] < y[best])) best = k;
    return best;
}

ipt_status run_chunks(ipt_batch* b, const std::vector<ipt_batch::Chunk>& cs, hipStream_t st) {
    for (const auto& c : cs) {
        const ipt_status r = chunk_step_launch(b->ctx, c.data, c.nslices, c.slice_vecs, c.plan,
                                               b->moves[c.moves], b->d_head, c.group, st);
        if (r != IPT_OK) return r;
    }
    return IPT_OK;
}

ipt_status enqueue(ipt_batch* b, hipStream_t st, bool timed = false) {
    ipt_status r = run_chunks(b, b->pre, st);
    if (r != IPT_OK) return r;
    if (b->nitems) {
        if (timed)
            hipExtLaunchKernelGGL(k_batch, dim3(b->grid), dim3(kBatchBlock), b->lds, st, b->k0, b->k1,
                                  0, b->d_insts, b->d_items, (const uint32_t*)b->d_seg, b->nq,
                                  b->d_sync, b->d_sync + 32);
        else
            hipLaunchKernelGGL(k_batch, dim3(b->grid), dim3(kBatchBlock), b->lds, st, b->d_insts,
                               b->d_items, (const uint32_t*)b->d_seg, b->nq, b->d_sync,
                               b->d_sync + 32);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            set_error("k_batch", e);
            return IPT_ERR_HIP;
        }
    }
    r = run_chunks(b, b->post, st);
    if (r != IPT_OK) return r;
    for (const auto& x : b->big) {
        r = run_variant(b->ctx, x.data, x.s, x.variant, x.tune, st);
        if (r != IPT_OK) return r;
    }
    return IPT_OK;
}

void release(ipt_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->d_insts) (void)hipFree(b->d_insts);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_seg) (void)hipFree(b->d_seg);
    if (b->d_sync) (void)hipFree(b->d_sync);
    if (b->big_sync.ptr) (void)hipFree(b->big_sync.ptr);
    for (auto& m : b->moves) chunk_moves_free(&m);
    if (b->d_head) (void)hipFree(b->d_head);
    if (b->wd_host) (void)hipHostFree(b->wd_host);
    if (b->k0) (void)hipEventDestroy(b->k0);
    if (b->k1) (void)hipEventDestroy(b->k1);
    delete b;
}

// The batch's barrier state is dirty when one of its group barriers gave up
// (the host-mapped word says so without synchronising): wait for the
// device, zero the state, rewrite the control block's fixed words.
ipt_status reset_batch_sync(ipt_batch* b) {
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemset(b->d_sync, 0, (size_t)b->sync_words * sizeof(uint32_t));
    if (e == hipSuccess && b->wd_addr) e = ctl_init(b->d_sync, b->wd_addr);
    if (e == hipSuccess && b->big_sync.ptr) e = hipMemset(b->big_sync.ptr, 0, b->big_sync.bytes);
    if (e == hipSuccess && b->big_sync.ptr && b->wd_addr)
        e = ctl_init((uint32_t*)b->big_sync.ptr, b->wd_addr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (b->wd_host) __atomic_store_n(b->wd_host, 0u, __ATOMIC_RELEASE);
    if (e != hipSuccess) {
        set_error("batch barrier reset", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

// IPT_ERR_HIP when an earlier run of this batch tripped its watchdog (its
// output is not valid); the state is restored, so the next run is normal.
ipt_status batch_watchdog_check(ipt_batch* b) {
    if (!b->wd_host || __atomic_load_n(b->wd_host, __ATOMIC_ACQUIRE) == 0) return IPT_OK;
    const ipt_status r = reset_batch_sync(b);
    if (r != IPT_OK) return r;
    set_message("an earlier run of this batch gave up waiting at a group barrier (watchdog): its "
                "output is not valid; this call did nothing");
    return IPT_ERR_HIP;
}

}  // namespace

extern "C" {

ipt_status ipt_batch_create(const ipt_matrix_desc* descs, int64_t count, const ipt_tune* tune,
                            ipt_batch** out) {
    if (!out || count < 0 || (count > 0 && !descs)) return IPT_ERR_INVALID_ARG;
    *out = nullptr;
    for (int64_t i = 0; i < count; i++) {
        ipt_status st = validate(descs[i].data, descs[i].elem_bytes, descs[i].d1, descs[i].d2,
                                 descs[i].d3);
        if (st != IPT_OK) return st;
        if (descs[i].variant < 0 || descs[i].variant >= IPT_VARIANT_COUNT)
            return IPT_ERR_INVALID_ARG;
    }
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    std::vector<std::pair<void*, size_t>> managed;
    for (int64_t i = 0; i < count; i++) {
        const size_t bytes = (size_t)descs[i].d1 * (size_t)descs[i].d2 * (size_t)descs[i].d3 *
                             (size_t)descs[i].elem_bytes;
        int kind = 0;
        if ((st = prepare_memory(ctx, descs[i].data, bytes, nullptr, false, &kind)) != IPT_OK)
            return st;
        if (kind == IPT_MEM_MANAGED) managed.emplace_back(descs[i].data, bytes);
    }
    ipt_batch* b = new (std::nothrow) ipt_batch;
    if (!b) return IPT_ERR_OUT_OF_MEMORY;
    b->ctx = ctx;
    b->managed = std::move(managed);
    b->tune = resolve_tune(ctx, tune);
    // One workgroup of kBatchBlock threads per CU at this LDS size (1 WG/CU
    // is the residency the group barriers rely on).
    const int64_t lds_cap = std::min<int64_t>(b->tune.lds_cap, lds_max_of(ctx)) - 64;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch, kBatchBlock, (size_t)lds_cap) !=
            hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    const int grid = per_cu * ncu_of(ctx);

    // Predictions for every AUTO instance, scored on the GPU in one call.
    std::vector<Shape> auto_shapes;
    for (int64_t i = 0; i < count; i++)
        if (descs[i].variant == IPT_VARIANT_AUTO)
            auto_shapes.push_back(make_shape(descs[i].elem_bytes, descs[i].d1, descs[i].d2, descs[i].d3));
    std::vector<float> pred;
    const bool have_pred = predict_gpu(auto_shapes, pred);
    size_t next_auto = 0;

    std::vector<BatchInst> insts;
    // Cooperative groups (one slice of an instance: g consecutive items).
    struct Group {
        uint32_t inst, g;
        uint64_t sl, bytes;
    };
    std::vector<Group> groups;
    std::vector<BatchItem> swap_items;
    std::vector<std::pair<uint64_t, BatchItem>> slice_sorted;
    uint64_t bar_words_total = 0;
    size_t lds_need = 0;
    // Items of one instance on a persistent-kernel route.
    auto add_items = [&](void* data, const Shape& s, Route p) -> ipt_status {
        BatchInst I{};
        I.data = (uint64_t)(uintptr_t)data;
        I.R = (uint32_t)s.d2;
        I.C = (uint32_t)s.d1;
        I.esz = (uint32_t)s.esz;
        const uint32_t idx = (uint32_t)insts.size();
        if (p == Route::Slice) {
            I.kind = BK_SLICE;
            I.P = (uint32_t)slice_pitch(s.d1);
            I.dS = Div32::make(I.R * I.C);
            I.dC = Div32::make(I.C);
            I.dR = Div32::make(I.R);
            const uint64_t slice_bytes = (uint64_t)I.R * I.P * s.esz;
            const uint64_t per = std::max<uint64_t>(1, std::min<uint64_t>(
                (uint64_t)b->tune.lds_bytes / slice_bytes, (uint64_t)lds_cap / slice_bytes));
            lds_need = std::max<size_t>(lds_need, std::min<uint64_t>(per, (uint64_t)s.d3) * slice_bytes);
            for (uint64_t sl = 0; sl < (uint64_t)s.d3; sl += per) {
                BatchItem it{idx, (uint32_t)std::min<uint64_t>(per, (uint64_t)s.d3 - sl), sl};
                slice_sorted.push_back({slice_bytes * it.sub, it});
            }
            b->info.n_slice_items += ((uint64_t)s.d3 + per - 1) / per;
        } else if (p == Route::Gcs) {
            const GcsGeom gg = gcs_geom(ctx, s, lds_cap, false, kBatchJrClass, false, true);
            I.kind = BK_GCS;
            I.jr = (uint32_t)gcs_jr_class(s.esz, gg.K, gg.KL, kBatchBlock / 64, rows_off_lines(s));
            I.g = gg.g;
            I.K = gg.K;
            I.KL = gg.KL;
            I.ntiles = gg.ntiles;
            I.dntj = Div32::make(gg.ntj);
            // members own blocks of adjacent tiles: uneven 4096-matrix batch
            // 4.48 -> 4.29 ms (write traffic 1.13 -> 1.07x, most rows not
            // 128-byte aligned)
            I.ord = banded_order(kBatchBlock / 64, gg.ntiles / gg.ntj, gg.ntj);
            I.bar = bar_words_total;
            bar_words_total += bar_words(gg.g) * (uint64_t)s.d3;
            lds_need = std::max(lds_need, gg.lds);
            for (uint64_t sl = 0; sl < (uint64_t)s.d3; sl++)
                groups.push_back({idx, gg.g, sl, (uint64_t)s.d1 * s.d2 * s.esz});
            b->info.n_gcs_items += (uint64_t)gg.g * s.d3;
        } else {
            const uint64_t TS = s.esz == 4 ? 64 : 32;
            const bool vec = ((uint64_t)s.d1 * s.esz) % 16 == 0 && (uintptr_t)data % 16 == 0;
            // the buffer-resource form takes any side for 4-byte elements (as
            // the per-call square swap)
            const bool rs = (vec || s.esz == 4) && (uint64_t)s.d1 * s.d1 * s.esz < ((uint64_t)1 << 31);
            I.kind = rs ? BK_SWAP_RS : vec ? BK_SWAP_VEC : BK_SWAP;
            const uint64_t nt = ((uint64_t)s.d1 + TS - 1) / TS;
            const uint64_t pairs = nt * (nt + 1) / 2;
            if (pairs >= ((uint64_t)1 << 32)) return IPT_ERR_TOO_LARGE;
            lds_need = std::max<size_t>(lds_need, 2 * TS * (TS + 1) * s.esz);
            for (uint64_t sl = 0; sl < (uint64_t)s.d3; sl++)
                for (uint64_t t = 0; t < pairs; t++) swap_items.push_back({idx, (uint32_t)t, sl});
            b->info.n_swap_items += pairs * s.d3;
        }
        insts.push_back(I);
        return IPT_OK;
    };
    // A block-cycle instance: its blocks on a persistent-kernel route, its
    // chunk step recorded with move lists of its own (shared by instances of
    // one chunk-matrix shape).  false = the instance cannot go this way.
    std::map<std::tuple<int64_t, int64_t, int64_t>, size_t> moves_of;
    size_t head_bytes = 0;
    size_t big_sync_bytes = 0;  // barrier state of the separate cooperative launches
    // Plans are made for the batch kernel's register-tile class; a chunk-size
    // range (a tuned or learned chunk class) that leaves no block this kernel
    // runs falls back to the whole range before the instance leaves the batch.
    auto add_block_cycle = [&](void* data, const Shape& s, const Tune& t, ipt_status* err) -> bool {
        *err = IPT_OK;
        BlockCyclePlan bp;
        Route rb = Route::Big;
        for (int pass = 0; pass < 2 && rb == Route::Big; pass++) {
            if (pass == 1 && t.chunk_min <= 0 && t.chunk_max <= 0) break;
            bp = block_cycle_for(block_cycle_plan(ctx, s, pass ? 0 : t.chunk_min,
                                                  pass ? 0 : t.chunk_max, kBatchJrClass),
                                 data);
            if (!bp.ok) continue;
            int bv = IPT_VARIANT_AUTO;
            rb = plan_instance(ctx, bp.blk, bp.block_variant, lds_cap, grid, b->tune, &bv);
            if (rb == Route::Skip) rb = Route::Big;
        }
        if (rb == Route::Big) return false;
        const auto key = std::make_tuple(bp.rc, bp.cc, bp.ell);
        auto it = moves_of.find(key);
        if (it == moves_of.end()) {
            ChunkMoves m;
            *err = chunk_moves_create(bp, &m);
            if (*err != IPT_OK) return false;
            b->moves.push_back(m);
            it = moves_of.emplace(key, b->moves.size() - 1).first;
        }
        const ChunkMoves& m = b->moves[it->second];
        const uint64_t per_slice = chunk_head_bytes(bp, m);
        if (per_slice > kScratchCap) return false;
        const uint64_t group =
            per_slice ? std::max<uint64_t>(1, kScratchCap / per_slice) : (uint64_t)s.d3;
        head_bytes = std::max<size_t>(head_bytes, std::min<uint64_t>(group, s.d3) * per_slice);
        *err = add_items(data, bp.blk, rb);
        if (*err != IPT_OK) return false;
        const ipt_batch::Chunk c{data, (uint64_t)s.d3, (uint64_t)s.d1 * s.d2 * s.esz / (uint64_t)bp.vb, group,
                                 bp, it->second};
        (bp.chunk_first ? b->pre : b->post).push_back(c);
        b->info.n_block_cycle++;
        return true;
    };
    for (int64_t i = 0; i < count; i++) {
        const ipt_matrix_desc& d = descs[i];
        const Shape s = make_shape(d.elem_bytes, d.d1, d.d2, d.d3);
        int big_variant = IPT_VARIANT_AUTO;
        int forced = d.variant;
        Tune big_tune = b->tune;
        if (forced == IPT_VARIANT_AUTO) {
            const float* y = have_pred ? &pred[(next_auto++) * kNumCandidates] : nullptr;
            const int k = batch_choice(ctx, s, y, lds_cap, grid);
            if (k >= 0) {
                forced = kCandidates[k].variant;
                if (forced == IPT_VARIANT_C2R || forced == IPT_VARIANT_R2C ||
                    forced == IPT_VARIANT_BLOCK_CYCLE || forced == IPT_VARIANT_GCS)
                    big_tune = resolve_tune(ctx, &kCandidates[k].tune);
            }
        }
        big_tune.opt_scratch = false;  // a captured batch cannot grow the scratch arena
        const Route p = plan_instance(ctx, s, forced, lds_cap, grid, b->tune, &big_variant);
        if (p == Route::Skip) {
            b->info.skipped++;
            b->routes.push_back(IPT_ROUTE_SKIP);
            continue;
        }
        b->info.bytes += (uint64_t)s.d1 * s.d2 * s.d3 * s.esz;
        if (p != Route::Big) {
            st = add_items(d.data, s, p);
            if (st != IPT_OK) {
                release(b);
                return st;
            }
            b->routes.push_back(p == Route::Slice ? IPT_ROUTE_SLICE
                                : p == Route::Gcs ? IPT_ROUTE_GCS : IPT_ROUTE_SWAP);
            continue;
        }
        if (big_variant == IPT_VARIANT_AUTO) big_variant = choose_variant(ctx, s);
        if (big_variant == IPT_VARIANT_BLOCK_CYCLE) {
            if (add_block_cycle(d.data, s, big_tune, &st)) {
                b->routes.push_back(IPT_ROUTE_BLOCK_CYCLE);
                continue;
            }
            if (st != IPT_OK) {
                release(b);
                return st;
            }
            big_variant = s.d1 >= s.d2 ? IPT_VARIANT_C2R : IPT_VARIANT_R2C;
        }
        const bool c2r = big_variant == IPT_VARIANT_C2R;
        if ((big_variant == IPT_VARIANT_C2R || big_variant == IPT_VARIANT_R2C) &&
            decomp_needs_scratch(s, c2r, b->tune)) {
            release(b);
            return IPT_ERR_UNSUPPORTED;  // scratch tiers are not used inside a batch
        }
        if (!variant_feasible(ctx, s, big_variant)) {
            release(b);
            return IPT_ERR_UNSUPPORTED;
        }
        if (big_variant == IPT_VARIANT_GCS) {
            const GcsGeom gg = gcs_geom(ctx, s, big_tune.lds_cap, true, 3, true);
            if (!gg.ok) {
                release(b);
                return IPT_ERR_UNSUPPORTED;
            }
            big_sync_bytes = std::max<size_t>(
                big_sync_bytes, (size_t)(32 + bar_words(gg.g) * (uint64_t)s.d3) * sizeof(uint32_t));
        }
        b->big.push_back({d.data, s, big_variant, big_tune});
        b->info.n_big++;
        b->routes.push_back(big_variant == IPT_VARIANT_GCS ? IPT_ROUTE_GCS_LAUNCH : IPT_ROUTE_SEPARATE);
    }
    // Queues (see k_batch).  Heavy work first everywhere; the members of a
    // group stay consecutive.  Groups go to the least loaded XCD queue, in
    // decreasing size, when nq * (g - 1) < grid; the rest to queue 0.
    uint64_t n_gcs_items = 0;
    for (const Group& x : groups) n_gcs_items += x.g;
    const uint64_t nitems_total = n_gcs_items + swap_items.size() + slice_sorted.size();
    const int grid_final = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)grid, nitems_total));
    uint32_t nq = (uint32_t)std::min<int>(std::max(ctx->nxcd, 1), (int)kMaxXcdQueues);
    if ((uint32_t)grid_final < 2 * nq) nq = 1;
    std::stable_sort(groups.begin(), groups.end(),
                     [](const Group& x, const Group& y) { return x.bytes > y.bytes; });
    std::stable_sort(slice_sorted.begin(), slice_sorted.end(),
                     [](const auto& x, const auto& y) { return x.first > y.first; });
    std::vector<std::vector<BatchItem>> queues(nq + 2);
    std::vector<uint64_t> load(nq, 0);
    for (const Group& x : groups) {
        size_t q = 0;
        if ((uint64_t)nq * (x.g - 1) < (uint64_t)grid_final) {
            const size_t k = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
            load[k] += x.bytes;
            q = 1 + k;
        }
        for (uint32_t m = 0; m < x.g; m++) queues[q].push_back({x.inst, m, x.sl});
    }
    queues[nq + 1] = swap_items;
    for (const auto& x : slice_sorted) queues[nq + 1].push_back(x.second);
    std::vector<BatchItem> items;
    std::vector<uint32_t> seg(nq + 3, 0);
    items.reserve(nitems_total);
    for (uint32_t q = 0; q < nq + 2; q++) {
        seg[q] = (uint32_t)items.size();
        items.insert(items.end(), queues[q].begin(), queues[q].end());
    }
    seg[nq + 2] = (uint32_t)items.size();

    if (!b->pre.empty() || !b->post.empty()) {
        st = chunk_kernels_setup(ctx);
        if (st != IPT_OK) {
            release(b);
            return st;
        }
    }
    // queue positions count in 32 bits, and each counter overshoots by at
    // most one take per workgroup
    if (nitems_total + (uint64_t)grid_final >= ((uint64_t)1 << 32)) {
        release(b);
        return IPT_ERR_TOO_LARGE;
    }
    b->nitems = items.size();
    b->nq = nq;
    b->lds = std::max<size_t>(lds_need, 16);
    b->grid = grid_final;
    b->info.count = count;
    b->info.n_instances = (int64_t)insts.size();
    b->info.grid = b->grid;
    b->info.lds_bytes = (int64_t)b->lds;
    hipError_t e = hipSuccess;
    if (!insts.empty()) e = hipMalloc(&b->d_insts, insts.size() * sizeof(BatchInst));
    if (e == hipSuccess && !items.empty()) e = hipMalloc(&b->d_items, items.size() * sizeof(BatchItem));
    if (e == hipSuccess) e = hipMalloc(&b->d_seg, seg.size() * sizeof(uint32_t));
    if (e == hipSuccess)
        e = hipMemcpy(b->d_seg, seg.data(), seg.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    b->sync_words = 32 + bar_words_total;
    const size_t sync_bytes = b->sync_words * sizeof(uint32_t);
    if (e == hipSuccess) e = hipMalloc(&b->d_sync, sync_bytes);
    if (e == hipSuccess) e = hipMemset(b->d_sync, 0, sync_bytes);
    if (e == hipSuccess && watchdog_word(&b->wd_host, &b->wd_addr)) e = ctl_init(b->d_sync, b->wd_addr);
    if (e == hipSuccess && big_sync_bytes) {
        e = hipMalloc(&b->big_sync.ptr, big_sync_bytes);
        if (e == hipSuccess) e = hipMemset(b->big_sync.ptr, 0, big_sync_bytes);
        if (e == hipSuccess && b->wd_addr) e = ctl_init((uint32_t*)b->big_sync.ptr, b->wd_addr);
        b->big_sync.bytes = big_sync_bytes;
        b->big_sync.fixed = true;
        for (auto& x : b->big)
            if (x.variant == IPT_VARIANT_GCS) x.tune.sync = &b->big_sync;
    }
    if (e == hipSuccess && !insts.empty())
        e = hipMemcpy(b->d_insts, insts.data(), insts.size() * sizeof(BatchInst), hipMemcpyHostToDevice);
    if (e == hipSuccess && !items.empty())
        e = hipMemcpy(b->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice);
    if (e == hipSuccess && head_bytes) e = hipMalloc(&b->d_head, head_bytes);
    if (e == hipSuccess && b->lds > 64 * 1024)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_batch),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("ipt_batch_create", e);
        release(b);
        return e == hipErrorOutOfMemory ? IPT_ERR_OUT_OF_MEMORY : IPT_ERR_HIP;
    }
    *out = b;
    return IPT_OK;
}

// Managed instances are prefetched to the device on the caller's stream
// before each run (outside the graph), as the per-call entry points do.
static ipt_status prefetch_managed(ipt_batch* b, hipStream_t st) {
    for (const auto& m : b->managed) {
        const ipt_status r = prepare_memory(b->ctx, m.first, m.second, st, true);
        if (r != IPT_OK) return r;
    }
    return IPT_OK;
}

ipt_status ipt_batch_run_eager(ipt_batch* b, void* stream) {
    if (!b) return IPT_ERR_INVALID_ARG;
    ipt_status r = batch_watchdog_check(b);
    if (r != IPT_OK) return r;
    r = prefetch_managed(b, (hipStream_t)stream);
    if (r != IPT_OK) return r;
    const bool timed = b->timing && b->nitems;
    r = enqueue(b, (hipStream_t)stream, timed);
    b->timed = timed && r == IPT_OK;
    return r;
}

ipt_status ipt_batch_set_timing(ipt_batch* b, int on) {
    if (!b) return IPT_ERR_INVALID_ARG;
    if (on && !b->k0) {
        hipError_t e = hipEventCreate(&b->k0);
        if (e == hipSuccess) e = hipEventCreate(&b->k1);
        if (e != hipSuccess) {
            set_error("ipt_batch_set_timing", e);
            return IPT_ERR_HIP;
        }
    }
    b->timing = on != 0;
    return IPT_OK;
}

ipt_status ipt_batch_kernel_ms(ipt_batch* b, float* ms) {
    if (!b || !ms) return IPT_ERR_INVALID_ARG;
    if (!b->timed) return IPT_ERR_NOT_READY;
    hipError_t e = hipEventSynchronize(b->k1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, b->k0, b->k1);
    if (e != hipSuccess) {
        set_error("ipt_batch_kernel_ms", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

ipt_status ipt_batch_run(ipt_batch* b, void* stream) {
    if (!b) return IPT_ERR_INVALID_ARG;
    const ipt_status wd = batch_watchdog_check(b);
    if (wd != IPT_OK) return wd;
    if (!b->exec) {
        hipError_t e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) {
            set_error("hipStreamBeginCapture", e);
            return IPT_ERR_HIP;
        }
        ipt_status st = enqueue(b, b->cap);
        e = hipStreamEndCapture(b->cap, &b->graph);
        if (st != IPT_OK) return st;
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            set_error("hipGraphInstantiate", e);
            return IPT_ERR_HIP;
        }
    }
    const ipt_status r = prefetch_managed(b, (hipStream_t)stream);
    if (r != IPT_OK) return r;
    hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error("hipGraphLaunch", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

ipt_status ipt_batch_get_info(const ipt_batch* b, ipt_batch_info* info) {
    if (!b || !info) return IPT_ERR_INVALID_ARG;
    *info = b->info;
    return IPT_OK;
}

ipt_status ipt_batch_get_routes(const ipt_batch* b, int32_t* routes) {
    if (!b || (!routes && !b->routes.empty())) return IPT_ERR_INVALID_ARG;
    std::copy(b->routes.begin(), b->routes.end(), routes);
    return IPT_OK;
}

ipt_status ipt_batch_check(ipt_batch* b) {
    if (!b) return IPT_ERR_INVALID_ARG;
    hipError_t e = hipDeviceSynchronize();
    uint32_t flag = 0;
    if (e == hipSuccess) e = hipMemcpy(&flag, b->d_sync + 2, sizeof(flag), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error("ipt_batch_check", e);
        return IPT_ERR_HIP;
    }
    if (flag || (b->wd_host && __atomic_load_n(b->wd_host, __ATOMIC_ACQUIRE))) {
        // A group barrier gave up waiting: restore the zero-at-rest state.
        (void)reset_batch_sync(b);
        set_message("group barrier watchdog tripped in a batch");
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

ipt_status ipt_batch_debug_set_watchdog(ipt_batch* b, uint32_t polls, uint32_t extra) {
    if (!b) return IPT_ERR_INVALID_ARG;
    const uint32_t w[2] = {polls, extra};
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(b->d_sync + 3, w, sizeof(w), hipMemcpyHostToDevice);
    if (e == hipSuccess && b->big_sync.ptr)
        e = hipMemcpy((uint32_t*)b->big_sync.ptr + 3, w, sizeof(w), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error("ipt_batch_debug_set_watchdog", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

ipt_status ipt_batch_destroy(ipt_batch* b) {
    if (!b) return IPT_ERR_INVALID_ARG;
    (void)hipDeviceSynchronize();
    release(b);
    return IPT_OK;
}

}  // extern "C"
