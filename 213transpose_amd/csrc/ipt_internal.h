// ipt_internal.h — host-side declarations shared by the library's sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "../../include/inplace_transpose.h"
#include "ipt_plan.h"

namespace ipt {

// A device buffer owned by the library and shared by every call on one
// device.  Users on different streams are ordered through `ev` (the next user
// waits for the previous user's last kernel).  `zero_at_rest` buffers are
// cleared when allocated and every kernel that uses them restores zeros; for
// the barrier arena, control words 6 and 7 then hold `wd_addr` (see
// report_watchdog).  A buffer that a stream capture has used is never freed:
// the captured graph's nodes keep its address, so growth retires it to
// `retired` (kept until the process ends) instead; growth itself (an
// allocation and a device synchronise) is refused during capture.
struct Arena {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    bool ev_valid = false;
    hipStream_t last = nullptr;
    bool zero_at_rest = false;
    bool captured = false;        // a stream capture has recorded `ptr`
    uint64_t wd_addr = 0;         // device address of the owner's watchdog word
    uint32_t** wd_host = nullptr; // where the owner keeps that word's host view
    std::vector<void*> retired;   // buffers captures may still reference
    // Fixed: `ptr` / `bytes` are set once and never change (memory the
    // library does not allocate); acquiring more than `bytes` fails with
    // IPT_ERR_TOO_LARGE, and no ordering events are kept (users that
    // overlap on one device race; the owner documents it).
    bool fixed = false;
};

// Per-device state: properties read on first use of the device (not at
// static init, so a later hipSetDevice is honoured) and the cached arenas.
struct DeviceCtx {
    int dev = 0;
    int ncu = 0;
    int nxcd = 1;  // accelerator complex dies (XCDs); workgroup i runs on XCD i % nxcd
    int64_t lds_max = 0;
    char arch[64] = {0};
    std::mutex mu;  // guards the arenas and serialises enqueues that use them
    Arena scratch;  // fallback-tier staging
    Arena sync;     // group-barrier counters (zero at rest)
    bool pageable = false;  // kernels may touch pageable host memory (HMM)
    // Host-mapped pinned word that a group barrier of the per-call kernels
    // sets when it gives up waiting (report_watchdog); the next call on the
    // device reads it with a plain host load (watchdog_check).
    uint32_t* wd_host = nullptr;
};

// Pinned host word mapped for the device (hipHostMalloc, mapped): *host is
// the host view, *dev_addr the address kernels store to.  False when the
// system offers no such memory (the sticky device flag still reports).
bool watchdog_word(uint32_t** host, uint64_t* dev_addr);
// Writes the barrier control block's fixed words (the watchdog address) after
// the block was zeroed.  Synchronous.
hipError_t ctl_init(uint32_t* ctl, uint64_t wd_addr);

// One 213 problem: d3 slices of d2 rows x d1 columns, esz-byte elements.
struct Shape {
    int esz = 4;
    int64_t d1 = 0, d2 = 0, d3 = 0;
};

// Resolved tuning knobs (see ipt_tune).
struct Tune {
    int lds_cap;    // hard LDS cap per workgroup: decides LDS tier vs fallback
    int lds_bytes;  // LDS packing target per workgroup (<= lds_cap)
    int block;      // threads per workgroup
    int strip;      // max column-strip width (elements)
    int grid_mult;  // workgroups per CU (x occupancy) for grid-stride kernels
    int64_t chunk_min = 0, chunk_max = 0;  // block-cycle chunk bytes (0: the defaults)
    int tile = 0;   // square-swap tile side (0: default)
    bool opt_scratch = true;  // column passes may take the scratch tier by choice
    // Arenas of the run (null: the device's own): the barrier state of the
    // cooperative kernel and the staging of the decomposition's scratch tiers.
    Arena* sync = nullptr;
    Arena* scratch = nullptr;
};

constexpr int64_t kDefaultLds = 160 * 1024;       // gfx950 LDS per workgroup
constexpr size_t kScratchCap = (size_t)256 << 20;  // fallback-tier scratch bound

void set_error(const char* what, hipError_t e);
void set_message(const char* what);  // the detail of a non-HIP failure

// Return IPT_ERR_HIP (recording the call) when a launch or HIP call failed.
#define IPT_LAUNCH_CHECK(what)                          \
    do {                                                \
        hipError_t e_ = hipGetLastError();              \
        if (e_ != hipSuccess) {                         \
            set_error(what, e_);                        \
            return IPT_ERR_HIP;                         \
        }                                               \
    } while (0)

#define IPT_HIP_CHECK(call)                             \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) {                         \
            set_error(#call, e_);                       \
            return IPT_ERR_HIP;                         \
        }                                               \
    } while (0)

ipt_status get_ctx(DeviceCtx** out);
// IPT_ERR_HIP (with the detail in ipt_last_error) when a cooperative kernel
// enqueued earlier on this device gave up at a group barrier: its output is
// not valid.  Waits for the device, restores the barrier arena and clears
// the flag, so the call after it runs normally.  A plain host load when
// nothing tripped.
ipt_status watchdog_check(DeviceCtx* ctx);
// ipt_device_check's part for ipt_transpose_213_noalloc's barrier state on
// this device: *tripped when one of its barriers gave up (then cleared).
ipt_status noalloc_check(const DeviceCtx* ctx, bool* tripped);
ipt_status arena_acquire(Arena& a, size_t bytes, hipStream_t s, void** ptr);
ipt_status arena_event(Arena& a, hipStream_t s, hipEvent_t* ev);
ipt_status arena_release(Arena& a, hipStream_t s);
ipt_status validate(const void* data, int esz, int64_t d1, int64_t d2, int64_t d3);
// Refuse memory the current device cannot reach and prefetch managed memory
// (when `prefetch`) to the device on stream s; *kind (when not null)
// receives the ipt_memory_kind_t of `data`.
ipt_status prepare_memory(const DeviceCtx* ctx, const void* data, size_t bytes, hipStream_t s,
                          bool prefetch, int* kind = nullptr);
Shape make_shape(int esz, int64_t d1, int64_t d2, int64_t d3);

Tune resolve_tune(const DeviceCtx* ctx, const ipt_tune* t);
bool variant_feasible(const DeviceCtx* ctx, const Shape& s, int variant);
int choose_variant(const DeviceCtx* ctx, const Shape& s);
// Candidate (index into the selector's candidate list) the dispatcher runs for
// IPT_VARIANT_AUTO: the learned selector's choice when one is loaded, else the
// heuristic's.
int select_candidate(const DeviceCtx* ctx, const Shape& s);
bool selector_loaded();
bool predict_gpu(const std::vector<Shape>& shapes, std::vector<float>& y);
int tuned_candidate(const DeviceCtx* ctx, const Shape& s);  // -1: no measured choice
bool candidate_ok(const DeviceCtx* ctx, const Shape& s, int k);
ipt_status run_variant(DeviceCtx* ctx, void* data, const Shape& s, int variant, const Tune& t,
                       hipStream_t stream);
ipt_status run_auto(DeviceCtx* ctx, void* data, const Shape& s, const ipt_tune* tune,
                    hipStream_t stream);
ipt_status run_copy(DeviceCtx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st);
ipt_status run_oop(DeviceCtx* ctx, void* dst, const void* src, const Shape& s, hipStream_t st);
ipt_status run_pass(DeviceCtx* ctx, void* data, int esz, int64_t R, int64_t C, int64_t d3,
                    int map_id, const Tune& t, hipStream_t stream);

constexpr int kGcsBlock = 512;  // 8 waves per workgroup

struct GcsGeom {
    int ts;           // tile side
    uint32_t ntj, ntiles;
    uint32_t g;       // workgroups per slice
    uint32_t K;       // tiles per member (max)
    uint32_t KL;      // of which staged in LDS; the rest stay in registers
    size_t lds;       // bytes per workgroup
    uint32_t grid;    // packed: k * g, whole groups on one workgroup per CU (0: not packed)
    bool ok;
};

// spread: when every slice fits one round, use more members per slice.
// jr_cls: the largest register-tile class the kernel that runs it has
// (gcs_jr_of; the batch kernel stops at kBatchJrClass).
// pack: when slices take several rounds, size groups so that a whole number
// of them fills the CUs, and launch exactly those workgroups.
// tail: rows off the 128-byte line grid keep one register tile per wave free
// (the line after a wave's run of tiles in the line-aligned loads of
// gcs_body); classes below the largest only.
GcsGeom gcs_geom(const DeviceCtx* ctx, const Shape& s, int64_t lds_cap, bool spread = true,
                 int jr_cls = 3, bool pack = false, bool tail = false);
// Whether rows of this shape are off the 128-byte line grid (the line-aligned
// loads then want the tail register tile).
inline bool rows_off_lines(const Shape& s) { return ((s.d1 * s.esz) & 127) != 0; }
int64_t slice_pitch(int64_t C);
int64_t lds_max_of(const DeviceCtx* ctx);
int ncu_of(const DeviceCtx* ctx);
bool decomp_needs_scratch(const Shape& s, bool c2r, const Tune& t);
int64_t decomp_scratch_need(const Shape& s, bool c2r, const Tune& t);

// Block-cycle path (ipt_block_cycle.hip): chunks of T elements along one
// side moved by cycle following, blocks of other x T transposed in one sweep.
constexpr int64_t kBcMinChunk = 512;             // bytes per moved chunk, at least
constexpr int64_t kBcMaxChunk = 32 * 1024;       // and at most (larger chunks mean skinny
                                                 // blocks, measured slower)
constexpr int64_t kBcStageBytes = 64 * 1024;     // chunk parts staged per segment
constexpr int64_t kBcPartBytes = 2048;           // chunks above twice this move in parts of >= it
constexpr int64_t kBcLdsBlockBytes = 64 * 1024;  // blocks up to this go to the whole-slice kernel
constexpr uint32_t kBcMaxSeg = 128;              // chunk moves per work item
constexpr int64_t kBcMaxChunks = (int64_t)1 << 23;  // chunks per slice (host plan size)
constexpr size_t kBcPlanCacheBytes = (size_t)256 << 20;  // move plans kept per process

struct BlockCyclePlan {
    bool ok = false;
    bool chunk_first = false;  // chunks along the columns: chunk step, then block step
    int64_t T = 0;             // elements per chunk
    int64_t nb = 0;            // blocks (chunks along the chunked side) per slice
    int64_t other = 0;         // the side that is not chunked
    int64_t rc = 0, cc = 0;    // chunk matrix transposed by the chunk step
    int64_t vb = 16;           // bytes per vector the chunk step moves (16, 8 or 4)
    int64_t cv = 0;            // vectors per chunk
    int64_t nsub = 1;          // parts a chunk is moved in (each part follows the same cycles)
    int64_t ell = 0;           // chunks per segment (at most)
    Shape blk;                 // the block step as a 213 problem
    int block_variant = 0;     // its one-sweep kernel
};
// The cheapest plan with chunks of chunk_min .. chunk_max bytes (0: the
// defaults kBcMinChunk .. kBcMaxChunk) whose blocks a one-sweep kernel runs,
// the cooperative kernel with register tiles of class jr_cls at most (the
// persistent batch kernel stops at kBatchJrClass).
BlockCyclePlan block_cycle_plan(const DeviceCtx* ctx, const Shape& s, int64_t chunk_min = 0,
                                int64_t chunk_max = 0, int jr_cls = 2);
// The plan for data at `data`: vectors narrowed to the data's alignment.
BlockCyclePlan block_cycle_for(const BlockCyclePlan& p, const void* data);
// Parts and segment length of a plan whose T, blk.esz and vb are set.
void chunk_parts(BlockCyclePlan* p);
// Vectors per part (the last part may be shorter) and the number of parts.
inline int64_t chunk_part_vecs(const BlockCyclePlan& p) { return (p.cv + p.nsub - 1) / p.nsub; }
inline int64_t chunk_nparts(const BlockCyclePlan& p) {
    const int64_t pv = chunk_part_vecs(p);
    return (p.cv + pv - 1) / pv;
}
// [slices][rows][cols][inner] -> [slices][cols][rows][inner] by the chunk
// step alone.
ipt_status run_chunk_transpose(DeviceCtx* ctx, void* data, int esz, int64_t inner, int64_t rows,
                               int64_t cols, int64_t slices, hipStream_t st);
ipt_status run_block_cycle(DeviceCtx* ctx, void* data, const Shape& s, const Tune& t,
                           hipStream_t st);

// The chunk step's move lists in device memory (see ipt_block_cycle.hip).
struct ChunkMoves {
    uint2* items = nullptr;    // (first move, moves)
    uint32_t* msrc = nullptr;  // source position, or a flagged head slot
    uint32_t* mdst = nullptr;  // destination position
    uint32_t* hpos = nullptr;  // position of each saved head
    uint32_t nitems = 0, nheads = 0, nmoves = 0;
    size_t bytes = 0;  // device bytes held
};
// Builds the move lists of plan p on the device (returns when they are
// complete); the caller owns them and frees them with chunk_moves_free.
// Allocates: never under capture.
ipt_status chunk_moves_create(const BlockCyclePlan& p, ChunkMoves* out);
// The lists of the rc x cc chunk transpose with segments of ell, built on
// the device (as chunk_moves_create) or by the host walk.
ipt_status chunk_moves_debug(int64_t rc, int64_t cc, int64_t ell, bool device, ChunkMoves* out);
void chunk_moves_free(ChunkMoves* m);
// Raises the chunk kernel's dynamic-LDS limit on this device (once).
ipt_status chunk_kernels_setup(const DeviceCtx* ctx);
// Saved-head bytes per slice.
inline uint64_t chunk_head_bytes(const BlockCyclePlan& p, const ChunkMoves& m) {
    return (uint64_t)m.nheads * (uint64_t)p.cv * (uint64_t)p.vb;
}
// Chunk step over `nslices` slices of `slice_vecs` vectors (p.vb bytes) at data,
// `group` slices at a time through `head` (group * chunk_head_bytes bytes).
// Launches only: nothing is allocated, so it may be captured.
ipt_status chunk_step_launch(const DeviceCtx* ctx, void* data, uint64_t nslices,
                             uint64_t slice_vecs, const BlockCyclePlan& p, const ChunkMoves& m,
                             void* head, uint64_t group, hipStream_t st);

template <typename Ix>
ipt_status eval_map_host(int map_id, const HostPlan& h, const int64_t* i, const int64_t* j,
                         int64_t n, int64_t* out) {
    Plan<Ix> p = Plan<Ix>::make(h);
    for (int64_t t = 0; t < n; t++) {
        if (i[t] < 0 || i[t] >= h.R || j[t] < 0 || j[t] >= h.C) return IPT_ERR_INVALID_ARG;
        Ix ii = (Ix)i[t], jj = (Ix)j[t];
        Ix r;
        switch (map_id) {
        case MAP_C2R_PRE: r = MapC2RPre::src(p, ii, jj); break;
        case MAP_C2R_POST: r = MapC2RPost::src(p, ii, jj); break;
        case MAP_R2C_PRE: r = MapR2CPre::src(p, ii, jj); break;
        case MAP_R2C_POST: r = MapR2CPost::src(p, ii, jj); break;
        case MAP_C2R_SHUFFLE: r = MapC2RShuffle::src(p, ii, jj); break;
        default: r = MapR2CShuffle::src(p, ii, jj); break;
        }
        out[t] = (int64_t)r;
    }
    return IPT_OK;
}

}  // namespace ipt
