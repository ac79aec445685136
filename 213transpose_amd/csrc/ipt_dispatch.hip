// ipt_dispatch.hip — variant choice and the launchers of every pass.
//
// Variants (ipt_variant):
//   SLICE_LDS    whole slice staged in one workgroup's LDS, written back
//                transposed: one read and one write of every element.
//   SQUARE_SWAP  d1 == d2: tile pairs swapped through LDS, one sweep.
//   C2R / R2C    the three-pass decomposition (src/transpose.cu:30-89), each
//                pass a column or row gather; tiers per pass: LDS when the
//                column strip / row fits, else scratch gather + copy back.
//   GCS          a group of workgroups stages a slice in LDS, meets at a
//                per-slice barrier, writes it transposed: one sweep.
#include "ipt_internal.h"
#include "ipt_kernels.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <tuple>

namespace ipt {

// ---------------------------------------------------------------------------
// Plain copy (the bandwidth ceiling reported beside the transposes, and the
// calibration of the traffic counters): each workgroup streams whole blocks
// of bv 16-byte words, 8 per thread in flight, with non-temporal loads and
// stores.  n16 = number of 16-byte words.  Measured on MI355X, 7.5 GB:
// 5.38 TB/s with 256 KB blocks, against 4.54-4.79 TB/s for a grid-stride
// copy of one word per thread (4.26 for hipMemcpy).
__global__ void __launch_bounds__(256)
k_copy16(u32x4* __restrict__ dst, const u32x4* __restrict__ src, uint64_t n16, uint64_t bv) {
    constexpr uint32_t U = 8;
    const uint64_t nb = (n16 + bv - 1) / bv;
    for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint64_t e = umin64(n16, (b + 1) * bv);
        for (uint64_t i = b * bv + threadIdx.x; i < e; i += U * 256) {
            u32x4 r[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                if (i + u * 256 < e) r[u] = __builtin_nontemporal_load(src + i + u * 256);
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                if (i + u * 256 < e) __builtin_nontemporal_store(r[u], dst + i + u * 256);
        }
    }
}

// Out-of-place transpose (the comparator; the reference times cuTT's
// out-of-place plan for it, src/test_cutt.cu:18-41): dst slice s (C rows x R
// columns) = transpose of src slice s (R rows x C columns).  Tiles of TR x TS
// (TS * esz = 256 bytes) staged in LDS with an odd row pitch; VEC moves
// 16-byte vectors (R and C multiples of 16 / esz, 16-byte aligned pointers),
// else one element per lane.  Both sides are coalesced.  Measured on MI355X
// against the earlier form whose loads each sat behind a bounds branch (and
// so waited for the one before): f32 7 x 4194304 977 -> 1974 GB/s, f32 4099
// x 4093 3937 -> 4535 GB/s, the rest unchanged.
template <typename T, bool VEC, uint32_t TR>
__global__ void __launch_bounds__(256)
k_oop_transpose(T* __restrict__ dst, const T* __restrict__ src, uint32_t R, uint32_t C,
                uint64_t nslices, uint32_t nti, uint32_t ntj) {
    constexpr uint32_t TS = 256 / sizeof(T), TP = TS + 1;  // tile: TR rows x TS columns
    constexpr uint32_t PER = VEC ? 16 / sizeof(T) : 1;
    constexpr uint32_t VPR = TS / PER, RPP = 256 / VPR;   // loads: vectors per row, rows per pass
    constexpr uint32_t NP = TR / RPP;
    constexpr uint32_t VPO = TR / PER, RPO = 256 / VPO;   // stores: per output row, rows per pass
    constexpr uint32_t NO = TS / RPO;
    static_assert(NP >= 1 && NO >= 1 && TR % RPP == 0 && TS % RPO == 0, "tile shape");
    using V = union { uint4 v; T e[16 / sizeof(T)]; };
    __shared__ T tile[TR * TP];
    const uint64_t per_slice = (uint64_t)nti * ntj, total = per_slice * nslices;
    const uint32_t x = (threadIdx.x % VPR) * PER, y0 = threadIdx.x / VPR;
    const uint32_t xo = (threadIdx.x % VPO) * PER, yo0 = threadIdx.x / VPO;
    for (uint64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const uint64_t sl = w / per_slice, t = w - sl * per_slice;
        const uint32_t r0 = (uint32_t)(t / ntj) * TR, c0 = (uint32_t)(t % ntj) * TS;
        const T* sp = src + sl * (uint64_t)R * C;
        T* dp = dst + sl * (uint64_t)R * C;
        // Every pass's load first, at addresses clamped into the slice (a
        // branch around each load would wait for the previous one): the
        // tile's loads are all in flight at once; lanes outside the matrix
        // then drop what they loaded.
        const uint32_t cc = min(c0 + x, C - PER);
        V u[NP];
#pragma unroll
        for (uint32_t p = 0; p < NP; p++) {
            const uint32_t r = min(r0 + y0 + p * RPP, R - 1);
            if constexpr (VEC) u[p].v = *reinterpret_cast<const uint4*>(sp + (uint64_t)r * C + cc);
            else u[p].e[0] = sp[(uint64_t)r * C + cc];
        }
        const bool col_in = c0 + x < C;
#pragma unroll
        for (uint32_t p = 0; p < NP; p++) {
            const uint32_t y = y0 + p * RPP;
            if (col_in && r0 + y < R) {
#pragma unroll
                for (uint32_t q = 0; q < PER; q++) tile[y * TP + x + q] = u[p].e[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t p = 0; p < NO; p++) {  // output row c0 + y, columns r0 + xo ..
            const uint32_t y = yo0 + p * RPO;
            const uint32_t orow = c0 + y, ocol = r0 + xo;
            if (orow < C && ocol < R) {
                if constexpr (VEC) {
                    V o;
#pragma unroll
                    for (uint32_t q = 0; q < PER; q++) o.e[q] = tile[(xo + q) * TP + y];
                    *reinterpret_cast<uint4*>(dp + (uint64_t)orow * R + ocol) = o.v;
                } else {
                    dp[(uint64_t)orow * R + ocol] = tile[xo * TP + y];
                }
            }
        }
        __syncthreads();
    }
}

int64_t lds_max_of(const DeviceCtx* ctx) { return ctx ? ctx->lds_max : kDefaultLds; }
int ncu_of(const DeviceCtx* ctx) { return ctx ? ctx->ncu : 256; }

Tune resolve_tune(const DeviceCtx* ctx, const ipt_tune* t) {
    Tune r;
    r.lds_cap = (int)lds_max_of(ctx);
    r.lds_bytes = (int)std::min<int64_t>(64 * 1024, lds_max_of(ctx));
    r.block = 256;
    r.strip = 64;
    r.grid_mult = 2;
    if (t) {
        if (t->lds_bytes > 0) {
            r.lds_cap = (int)std::min<int64_t>(t->lds_bytes, lds_max_of(ctx));
            r.lds_bytes = std::min(r.lds_bytes, r.lds_cap);
        }
        if (t->block >= 64 && t->block <= 1024 && (t->block & (t->block - 1)) == 0) r.block = t->block;
        if (t->strip > 0) r.strip = t->strip;
        if (t->grid_mult > 0) r.grid_mult = t->grid_mult;
        if (t->chunk_min > 0) r.chunk_min = t->chunk_min;
        if (t->chunk_max > 0) r.chunk_max = t->chunk_max;
        if (t->tile > 0) r.tile = t->tile;
    }
    return r;
}

// Odd LDS row pitch for the whole-slice kernel.
int64_t slice_pitch(int64_t C) { return C | 1; }

static bool slice_lds_fits(const DeviceCtx* ctx, const Shape& s) {
    return s.d2 * slice_pitch(s.d1) * s.esz <= lds_max_of(ctx);
}

// ---------------------------------------------------------------- GCS geometry

// Geometry of the cooperative kernel.  Each member holds at most KLmax wave
// tiles in LDS (JL per wave) plus KRmax in registers (JR per wave).  When
// every slice fits in one round of one workgroup per CU, slices are spread
// over all CUs (larger g, less data per member); otherwise members are
// filled to capacity.  Slices of 2 GiB or more never fit (buffer offsets are
// 32-bit, and such a slice is far beyond what 256 members hold).  Register
// tiles of class 2 are used only when a slice needs them to fit, or when
// packed groups of them share a round two or more to one of class 1: members
// that large were slower wherever a smaller class also fits (f32 3000 x 3000
// x 8: 3446 vs 4578 GB/s); class 3 only when class 2 cannot hold the slice
// in one group.
// Packing (several rounds of groups): g members per slice with gmin <= g
// leave ncu mod g workgroups that start the next round's first group early,
// and that group's other members then wait at its barrier for them to finish
// a whole slice first.  Groups of ncu / k members (k = ncu / gmin groups per
// round) launched on exactly k * g workgroups avoid it.  Measured on MI355X
// (GB/s, default -> packed): f32 1669 x 5801 x 25 (9646 tiles) 2525 -> 2873,
// and 3294 with the larger register tiles that fit two groups per round;
// f32 2000 x 2048 x 60 4677 -> 4720; f64 1200 x 1000 x 80 4448 -> 4609.  The
// larger register tiles pay only there: elsewhere they lost (f32 1500 x 1300
// x 150: 13 groups per round 3553, 9 groups with class 1 3794).
//
// gcs_geom_cls: the geometry with register tiles of class jr_cls.
static GcsGeom gcs_geom_cls(const DeviceCtx* ctx, const Shape& s, int64_t lds_cap, bool spread,
                            int jr_cls, bool pack, bool tail) {
    GcsGeom gg{};
    const int64_t R = s.d2, C = s.d1, esz = s.esz;
    gg.ts = esz == 4 ? 32 : 16;
    const int64_t tile_bytes = (int64_t)gg.ts * gg.ts * esz;
    const int64_t cap = std::min<int64_t>(lds_cap, lds_max_of(ctx)) - 64;  // static LDS headroom
    const int64_t nti = (R + gg.ts - 1) / gg.ts, ntj = (C + gg.ts - 1) / gg.ts;
    const int64_t ntiles = nti * ntj;
    constexpr int64_t nw = kGcsBlock / 64;
    const int64_t klmax = std::min<int64_t>(cap / tile_bytes, gcs_wave_lds_tiles(esz) * nw);
    const int64_t krmax = (int64_t)(gcs_jr_of(esz, jr_cls) - (tail ? 1 : 0)) * nw;
    gg.ok = klmax >= 1 && R * C * esz < ((int64_t)1 << 31);
    if (!gg.ok) return gg;
    const int64_t ncu = ncu_of(ctx);
    const int64_t gmin = (ntiles + klmax + krmax - 1) / (klmax + krmax);
    int64_t g = gmin;
    gg.grid = 0;
    if (spread && s.d3 * gmin <= ncu) {
        g = std::max<int64_t>(gmin, std::min<int64_t>(ncu / s.d3, ntiles));
    } else if (pack && gmin > 1 && gmin <= ncu) {
        const int64_t k = ncu / gmin;  // groups per round
        g = std::min<int64_t>(ncu / k, ntiles);
        gg.grid = (uint32_t)(k * g);
    }
    const int64_t K = (ntiles + g - 1) / g;
    gg.ntj = (uint32_t)ntj;
    gg.ntiles = (uint32_t)ntiles;
    gg.g = (uint32_t)g;
    gg.K = (uint32_t)K;
    gg.KL = (uint32_t)std::min(K, klmax);
    gg.lds = (size_t)(gg.KL * tile_bytes);
    // A group must never need more workgroups than one per CU (the residency
    // every launch of this kernel is guaranteed).
    gg.ok = g <= ncu && (uint64_t)s.d3 * (uint64_t)g < ((uint64_t)1 << 31);
    return gg;
}

GcsGeom gcs_geom(const DeviceCtx* ctx, const Shape& s, int64_t lds_cap, bool spread, int jr_cls,
                 bool pack, bool tail) {
    tail = tail && rows_off_lines(s);
    if (jr_cls <= 1) return gcs_geom_cls(ctx, s, lds_cap, spread, jr_cls, pack, tail);
    const GcsGeom small = gcs_geom_cls(ctx, s, lds_cap, spread, 1, pack, tail);
    if (!small.ok) {
        const GcsGeom c2 = gcs_geom_cls(ctx, s, lds_cap, spread, 2, pack, false);
        return c2.ok || jr_cls < 3 ? c2 : gcs_geom_cls(ctx, s, lds_cap, spread, 3, pack, false);
    }
    if (!pack || small.grid == 0 || small.grid / small.g != 1) return small;
    const GcsGeom big = gcs_geom_cls(ctx, s, lds_cap, spread, 2, pack, false);
    return big.ok && big.grid && big.grid / big.g >= 2 ? big : small;
}

bool variant_feasible(const DeviceCtx* ctx, const Shape& s, int variant) {
    switch (variant) {
    case IPT_VARIANT_AUTO: return true;
    case IPT_VARIANT_SLICE_LDS: return slice_lds_fits(ctx, s);
    case IPT_VARIANT_SQUARE_SWAP: return s.d1 == s.d2;
    case IPT_VARIANT_C2R:
    case IPT_VARIANT_R2C: return true;
    case IPT_VARIANT_GCS: return gcs_geom(ctx, s, lds_max_of(ctx)).ok;
    case IPT_VARIANT_BLOCK_CYCLE: return block_cycle_plan(ctx, s).ok;
    default: return false;
    }
}

// Heuristic dispatch (the fallback when no learned selector is loaded).
// Mirrors src/transpose.cu:105-123 for the decomposition choice (C2R when
// d1 >= d2, else R2C with swapped arguments) and adds the one-sweep variants
// first where they apply.  Square slices beyond the small whole-slice case
// take the square swap: it beat the cooperative kernel on every square
// measured (1.1x with many slices, up to 3.2x for one slice: a tile pair is
// read and written by one workgroup, so its stores never cross an XCD, and
// workgroups overlap their loads and stores), e.g. f32 4096^2 7402 vs 4189
// GB/s, f64 2048^2 6665 vs 2094, f32 512^2 x 64 7559 vs 6701.
int choose_variant(const DeviceCtx* ctx, const Shape& s) {
    const int64_t slice_bytes = s.d1 * s.d2 * s.esz;
    if (slice_bytes <= 16 * 1024 && slice_lds_fits(ctx, s)) return IPT_VARIANT_SLICE_LDS;
    if (s.d1 == s.d2) return IPT_VARIANT_SQUARE_SWAP;
    if (gcs_geom(ctx, s, lds_max_of(ctx)).ok) return IPT_VARIANT_GCS;
    if (slice_lds_fits(ctx, s)) return IPT_VARIANT_SLICE_LDS;
    if (block_cycle_plan(ctx, s).ok) return IPT_VARIANT_BLOCK_CYCLE;
    return s.d1 >= s.d2 ? IPT_VARIANT_C2R : IPT_VARIANT_R2C;
}

// Occupancy and the dynamic-LDS attribute belong to a kernel on a device and
// do not change; both are cached, since querying or setting them costs host
// time on every call.
static std::mutex g_kern_mu;
static std::map<std::tuple<int, const void*, int, size_t>, int> g_occupancy;
static std::set<std::pair<int, const void*>> g_lds_raised;

static int occupancy(const DeviceCtx* ctx, const void* fn, int block, size_t lds) {
    const auto key = std::make_tuple(ctx->dev, fn, block, lds);
    {
        std::lock_guard<std::mutex> g(g_kern_mu);
        auto it = g_occupancy.find(key);
        if (it != g_occupancy.end()) return it->second;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds) != hipSuccess ||
        per_cu <= 0) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    std::lock_guard<std::mutex> g(g_kern_mu);
    g_occupancy[key] = per_cu;
    return per_cu;
}

// Workgroups for a grid-stride kernel: occupancy x CUs x grid_mult, capped by
// the number of work items.
template <typename K>
static ipt_status grid_for(const DeviceCtx* ctx, K kernel, int block, size_t lds, uint64_t items,
                           int grid_mult, int* grid) {
    const int per_cu = occupancy(ctx, reinterpret_cast<const void*>(kernel), block, lds);
    uint64_t g = (uint64_t)per_cu * (uint64_t)ncu_of(ctx) * (uint64_t)std::max(grid_mult, 1);
    g = std::min<uint64_t>(g, items);
    *grid = (int)std::max<uint64_t>(g, 1);
    return IPT_OK;
}

// Dynamic LDS above 64 KB must be allowed per kernel; the first use on a
// device raises the limit to the device's maximum once.
template <typename K>
static ipt_status allow_lds(const DeviceCtx* ctx, K kernel, size_t lds) {
    if (lds <= 64 * 1024) return IPT_OK;
    const void* fn = reinterpret_cast<const void*>(kernel);
    std::lock_guard<std::mutex> g(g_kern_mu);
    if (g_lds_raised.count({ctx->dev, fn})) return IPT_OK;
    IPT_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)std::max<int64_t>((int64_t)lds, lds_max_of(ctx))));
    g_lds_raised.insert({ctx->dev, fn});
    return IPT_OK;
}

// Whether the C2R (c2r) or R2C decomposition of s falls to a scratch tier
// (a row longer, or a column taller, than the LDS cap).
bool decomp_needs_scratch(const Shape& s, bool c2r, const Tune& t) {
    const int64_t R = c2r ? s.d2 : s.d1, C = c2r ? s.d1 : s.d2;
    return C * s.esz + 16 > t.lds_cap || R * s.esz > t.lds_cap;
}

// Column strip of the LDS tier of a column pass (R rows x C columns): the
// widest power of two within the LDS budget; wider than the budget only when
// even a 1-column strip does not fit it.
static int64_t col_strip_width(int64_t R, int64_t C, int64_t esz, const Tune& t) {
    const int64_t wmax = std::min<int64_t>(std::max(t.strip, 1), t.block);
    int64_t W = 1;
    while (W * 2 <= C && W * 2 <= wmax) W *= 2;
    if (W < C && W * 2 <= wmax) W *= 2;
    while (W > 1 && R * W * esz > t.lds_bytes) W /= 2;
    while (W > 1 && R * W * esz > t.lds_cap) W /= 2;
    return W;
}

// Bytes of scratch the C2R (c2r) or R2C decomposition of s cannot do with
// less (0: LDS tiers only): a 256-byte column strip of the view's R rows when
// a column is taller than the LDS cap or the pass takes the scratch tier by
// choice (every column map here moves 256-byte strips there), a whole row when
// a row is longer than the LDS cap.
int64_t decomp_scratch_need(const Shape& s, bool c2r, const Tune& t) {
    const int64_t R = c2r ? s.d2 : s.d1, C = c2r ? s.d1 : s.d2;
    int64_t need = 0;
    const int64_t W = col_strip_width(R, C, s.esz, t);
    const bool by_choice = t.opt_scratch && W < C && W * s.esz < 64;
    if (R * s.esz > t.lds_cap || by_choice) need = std::max<int64_t>(need, R * 256);
    if (C * s.esz + 16 > t.lds_cap) need = std::max<int64_t>(need, C * s.esz);
    return need;
}

// ---------------------------------------------------------------- SLICE_LDS

template <typename T>
static ipt_status run_slice_lds(DeviceCtx* ctx, T* data, const Shape& s, const Tune& t,
                                hipStream_t st) {
    const uint32_t R = (uint32_t)s.d2, C = (uint32_t)s.d1;
    const uint32_t P = (uint32_t)slice_pitch(C);
    const int64_t slice_bytes = (int64_t)R * P * sizeof(T);
    int64_t budget = std::max<int64_t>(t.lds_bytes, slice_bytes);
    uint32_t per = (uint32_t)std::max<int64_t>(1, budget / slice_bytes);
    per = (uint32_t)std::min<uint64_t>(per, (uint64_t)s.d3);
    const size_t lds = (size_t)per * slice_bytes;
    // Vector form: every slice a multiple of 16 bytes at a 16-byte aligned
    // base; it runs 256-thread workgroups (its compile-time block).
    const bool vec = ((uint64_t)R * C * sizeof(T)) % 16 == 0 && (uintptr_t)data % 16 == 0;
    auto k = vec ? k_slice_lds<T, true> : k_slice_lds<T, false>;
    const int block = vec ? 256 : t.block;
    ipt_status stt = allow_lds(ctx, k, lds);
    if (stt != IPT_OK) return stt;
    const uint64_t items = ((uint64_t)s.d3 + per - 1) / per;
    int grid;
    grid_for(ctx, k, block, lds, items, t.grid_mult, &grid);
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, st, data, R, C, P, (uint64_t)s.d3, per,
                       Div32::make(R * C), Div32::make(C), Div32::make(R));
    IPT_LAUNCH_CHECK("k_slice_lds");
    return IPT_OK;
}

// ---------------------------------------------------------------- SQUARE_SWAP

// Square slices with tiles of TS x TS (the default 64 / 32 for 4- / 8-byte
// elements, or half of that: the `tile` knob).
template <typename T, int TS>
static ipt_status run_square_ts(DeviceCtx* ctx, T* data, const Shape& s, const Tune& t,
                                hipStream_t st) {
    const uint32_t N = (uint32_t)s.d1;
    const uint32_t nt = (N + TS - 1) / TS;
    const uint64_t pairs = (uint64_t)nt * (nt + 1) / 2;
    const bool vec = ((uint64_t)N * sizeof(T)) % 16 == 0 && (uintptr_t)data % 16 == 0;
    // Slices below 2^31 bytes go through a buffer resource (32-bit offsets),
    // one grid round of resident workgroups (measured best on 16384^2 f32):
    // 16-byte rows, and 4-byte elements with any side (f32 4095^2 4565 ->
    // 4754 GB/s, 1001^2 x 8 3466 -> 4041 over the scalar form).  Odd f64
    // sides keep the scalar form (2049^2 4690 vs 4389 through the resource).
    const bool rs = (vec || sizeof(T) == 4) && (uint64_t)N * N * sizeof(T) < ((uint64_t)1 << 31);
    auto k = rs ? k_square_swap_rs<T, TS> : vec ? k_square_swap<T, TS, true> : k_square_swap<T, TS, false>;
    // Pairs in diagonal order where the row pitch is a power of two of 16 KB
    // or more (swap_pair).
    const uint64_t pitch = (uint64_t)N * sizeof(T);
    const bool dord = pitch >= 16384 && (pitch & (pitch - 1)) == 0;
    int grid;
    grid_for(ctx, k, kSwapBlock, 0, pairs * (uint64_t)s.d3, rs ? 1 : t.grid_mult, &grid);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kSwapBlock), 0, st, data, N, pairs, (uint64_t)s.d3, dord);
    IPT_LAUNCH_CHECK("k_square_swap");
    return IPT_OK;
}

// Squares whose row pitch is a power of two of 64 KB or more (HBM-sized:
// f32 16384^2 and up): tiles of 512-byte row segments (k_square_swap_big),
// pairs in diagonal order, non-temporal loads and stores; measured on f32
// 16384^2, GB/s: 64-wide tiles 4746 in column order, 4952 in diagonal
// order; 128-wide 5012, with non-temporal accesses 5067.  Slices below
// 2^31 bytes with 16-byte rows (or 4-byte elements).
template <typename T>
static ipt_status run_square_big(DeviceCtx* ctx, T* data, const Shape& s, hipStream_t st) {
    constexpr int TS = sizeof(T) == 4 ? 128 : 64;
    const uint32_t N = (uint32_t)s.d1;
    const uint64_t nt = (N + TS - 1) / TS;
    const uint64_t pairs = nt * (nt + 1) / 2;
    const size_t lds = (size_t)2 * TS * (TS + 1) * sizeof(T);
    auto k = k_square_swap_big<T, TS>;
    ipt_status stt = allow_lds(ctx, k, lds);
    if (stt != IPT_OK) return stt;
    int grid;
    grid_for(ctx, k, kSwapBigBlock, lds, pairs * (uint64_t)s.d3, 1, &grid);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kSwapBigBlock), lds, st, data, N, pairs, (uint64_t)s.d3);
    IPT_LAUNCH_CHECK("k_square_swap_big");
    return IPT_OK;
}

template <typename T>
static ipt_status run_square(DeviceCtx* ctx, T* data, const Shape& s, const Tune& t,
                             hipStream_t st) {
    constexpr int TS = sizeof(T) == 4 ? 64 : 32;
    if (t.tile > 0 && t.tile <= TS / 2) return run_square_ts<T, TS / 2>(ctx, data, s, t, st);
    const uint64_t pitch = (uint64_t)s.d1 * sizeof(T);
    const bool vec = pitch % 16 == 0 && (uintptr_t)data % 16 == 0;
    if (pitch >= 65536 && (pitch & (pitch - 1)) == 0 && (vec || sizeof(T) == 4) &&
        pitch * (uint64_t)s.d1 < ((uint64_t)1 << 31))
        return run_square_big<T>(ctx, data, s, st);
    return run_square_ts<T, TS>(ctx, data, s, t, st);
}

// ---------------------------------------------------------------- GCS

// Signature of the GCS kernel (both element sizes).
template <typename T>
using GcsKernel = void (*)(T*, uint32_t, uint32_t, uint64_t, uint32_t, uint32_t, uint32_t, Div32,
                           uint32_t, uint32_t*, uint32_t*, uint32_t, TileOrder);

template <typename T>
static ipt_status launch_gcs(DeviceCtx* ctx, GcsKernel<T> k, T* d, const Shape& s,
                             const GcsGeom& gg, Arena& sync_arena, const TileOrder& ord,
                             hipStream_t st) {
    ipt_status stt = allow_lds(ctx, k, gg.lds);
    if (stt != IPT_OK) return stt;
    const int per_cu = occupancy(ctx, reinterpret_cast<const void*>(k), kGcsBlock, gg.lds);
    const uint64_t slots = (uint64_t)s.d3 * gg.g;
    const uint64_t resident = (uint64_t)per_cu * (uint64_t)ncu_of(ctx);
    // Never more workgroups than can be resident at once: every member of a
    // group must be running while the others wait for it.
    uint64_t want = std::min<uint64_t>(slots, resident);
    // packed: whole groups per round over every resident workgroup (k * g
    // at one workgroup per CU; twice that where the LDS allows two)
    if (gg.grid) want = std::min<uint64_t>(want, resident / gg.g * gg.g);
    const int grid = (int)std::max<uint64_t>(1, want);
    if ((uint64_t)grid < gg.g) return IPT_ERR_UNSUPPORTED;  // a group must be co-resident
    if (gg.g == 1) {  // whole slices per workgroup: no barrier state needed
        hipLaunchKernelGGL(k, dim3(grid), dim3(kGcsBlock), gg.lds, st, d, (uint32_t)s.d2,
                           (uint32_t)s.d1, (uint64_t)s.d3, gg.g, gg.K, gg.KL, Div32::make(gg.ntj),
                           gg.ntiles, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t)ctx->nxcd, ord);
        IPT_LAUNCH_CHECK("k_gcs");
        return IPT_OK;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    void* sync;
    stt = arena_acquire(sync_arena, (size_t)(32 + bar_words(gg.g) * (uint64_t)s.d3) * sizeof(uint32_t),
                        st, &sync);
    if (stt != IPT_OK) return stt;
    uint32_t* ctl = (uint32_t*)sync;
    // The arena's ordering event rides on the launch itself (a separate
    // event record between kernels costs ~3 us of queue time per call);
    // during stream capture it is recorded as its own node.
    hipEvent_t ev = nullptr;
    stt = arena_event(sync_arena, st, &ev);
    if (stt != IPT_OK) return stt;
    if (ev) {
        hipExtLaunchKernelGGL(k, dim3(grid), dim3(kGcsBlock), gg.lds, st, nullptr, ev, 0, d,
                              (uint32_t)s.d2, (uint32_t)s.d1, (uint64_t)s.d3, gg.g, gg.K, gg.KL,
                              Div32::make(gg.ntj), gg.ntiles, ctl, ctl + 32, (uint32_t)ctx->nxcd, ord);
        IPT_LAUNCH_CHECK("k_gcs");
        return IPT_OK;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(kGcsBlock), gg.lds, st, d, (uint32_t)s.d2,
                       (uint32_t)s.d1, (uint64_t)s.d3, gg.g, gg.K, gg.KL, Div32::make(gg.ntj),
                       gg.ntiles, ctl, ctl + 32, (uint32_t)ctx->nxcd, ord);
    IPT_LAUNCH_CHECK("k_gcs");
    return arena_release(sync_arena, st);
}

template <typename T>
static ipt_status run_gcs(DeviceCtx* ctx, T* data, const Shape& s, const Tune& t, hipStream_t st) {
    const GcsGeom gg = gcs_geom(ctx, s, t.lds_cap, true, 3, true);
    if (!gg.ok) return IPT_ERR_UNSUPPORTED;
    Arena& a = t.sync ? *t.sync : ctx->sync;
    // Banded tile ownership and write-back stores (the L2 merges the partial
    // lines of neighbouring tiles before they leave it) where they pay:
    // slices of 1 MiB or more in packed launches (groups take several
    // rounds, so each round's dirty lines are written back while later
    // rounds run) or in single-round launches whose members hold 56 tiles or
    // more.  Elsewhere the strided order and device-scope stores, whose
    // launches end with no dirty lines to write back: short single-round
    // launches lose more to that end-of-kernel write-back than the merging
    // saves.  Measured on MI355X, us per call with the default dispatch,
    // device scope strided -> write-back banded: packed f32 513 x 769 x 300
    // 268 -> 224, 1000 x 3000 x 40 266 -> 254; single round, K >= 56: f64
    // 1000 x 256 x 17 16.0 -> 14.2, f32 3001 x 2999 x 2 34.2 -> 32.5; single
    // round, K <= 48: f32 3000 x 1000 (K = 12) 16.6 -> 16.7, 1500 x 6000 (35)
    // 26.5 -> 26.9, 700 x 5000 x 3 (41) 22.9 -> 24.2; 786 KB slices: f32 512
    // x 384 x 64 17.8 -> 19.8.  Banded order with device-scope stores, and
    // non-temporal stores, were slower than either.
    const bool wide = (uint64_t)s.d1 * (uint64_t)s.d2 * sizeof(T) >= ((uint64_t)1 << 20) &&
                      (gg.grid != 0 || gg.K >= 56);
    const TileOrder ord = wide ? banded_order(kGcsBlock / 64, gg.ntiles / gg.ntj, gg.ntj) : TileOrder{};
    const int cls = gcs_jr_class(sizeof(T), gg.K, gg.KL, kGcsBlock / 64);
    if (wide) {
        switch (cls) {
        case 0: return launch_gcs<T>(ctx, k_gcs<T, kGcsBlock, gcs_jr_of(sizeof(T), 0), 0>, data, s, gg, a, ord, st);
        case 1: return launch_gcs<T>(ctx, k_gcs<T, kGcsBlock, gcs_jr_of(sizeof(T), 1), 0>, data, s, gg, a, ord, st);
        case 2: return launch_gcs<T>(ctx, k_gcs<T, kGcsBlock, gcs_jr_of(sizeof(T), 2), 0>, data, s, gg, a, ord, st);
        default: return launch_gcs<T>(ctx, k_gcs<T, kGcsBlock, gcs_jr_of(sizeof(T), 3), 0>, data, s, gg, a, ord, st);
        }
    }
    switch (cls) {
    case 0: return launch_gcs<T>(ctx, k_gcs<T, kGcsBlock, gcs_jr_of(sizeof(T), 0), kStoreDevice>, data, s, gg, a, ord, st);
    case 1: return launch_gcs<T>(ctx, k_gcs<T, kGcsBlock, gcs_jr_of(sizeof(T), 1), kStoreDevice>, data, s, gg, a, ord, st);
    case 2: return launch_gcs<T>(ctx, k_gcs<T, kGcsBlock, gcs_jr_of(sizeof(T), 2), kStoreDevice>, data, s, gg, a, ord, st);
    default: return launch_gcs<T>(ctx, k_gcs<T, kGcsBlock, gcs_jr_of(sizeof(T), 3), kStoreDevice>, data, s, gg, a, ord, st);
    }
}

// ---------------------------------------------------------------- passes

// Column gather pass new[i][j] = old[g(i,j)][j] over every slice.
// Tiers: a strip of W columns x all R rows in LDS (one sweep), else the
// strip gathered into scratch and copied back (two sweeps).  The LDS strip
// narrows as R grows; below 64 bytes of row segment its reads waste most of
// every line, so the pass takes the scratch tier instead, with 256-byte
// strips: a map whose source row is constant across runs of columns (a
// rotation, Map::kRunConstant) gathers row segments directly; the two
// diagonal maps (Map::kDiag) are walked in source order so their diagonal
// side stays in L2 (k_col_diag_to_scratch).  Measured on f32 R=28876 x
// C=9286 (strips of one column in LDS): C2R 13.9 -> 8.7 ms with the first
// alone.  Not under stream capture (the arena cannot grow there) and not
// where the caller rules scratch out (batches).
constexpr int64_t kDiagBandBytes = 128 * 1024;  // LDS band of the diagonal column maps

template <typename T, typename V, typename Ix, bool C2R_POST>
static ipt_status launch_diag(DeviceCtx* ctx, T* data, T* scr, const Plan<Ix>& p, uint32_t wshift,
                              Div64 dstrips, uint64_t first, uint64_t n, uint32_t kb,
                              hipStream_t st) {
    if constexpr (sizeof(V) < sizeof(T)) {
        return IPT_ERR_INVALID_ARG;  // 8-byte elements are always 8-byte aligned
    } else {
        const uint32_t W = 1u << wshift;
        const size_t lds = (size_t)(kb + W - 1) * W * sizeof(T);
        auto k = k_col_diag_to_scratch<T, V, Ix, C2R_POST>;
        ipt_status stt = allow_lds(ctx, k, lds);
        if (stt != IPT_OK) return stt;
        const uint32_t nkb = (uint32_t)(((uint64_t)p.R + kb - 1) / kb);
        const uint64_t units = n * nkb;
        const int gd = (int)std::min<uint64_t>(units, (uint64_t)ncu_of(ctx) * 8);
        hipLaunchKernelGGL(k, dim3(gd), dim3(kDiagBlock), lds, st, (const T*)data, scr, p, wshift,
                           dstrips, first, (uint32_t)n, kb, Div32::make(nkb));
        IPT_LAUNCH_CHECK("k_col_diag_to_scratch");
        return IPT_OK;
    }
}

template <typename T, typename Ix, typename Map>
static ipt_status col_op(DeviceCtx* ctx, T* data, const Plan<Ix>& p, const HostPlan& h,
                         uint64_t d3, const Tune& t, hipStream_t st) {
    const int64_t R = h.R, C = h.C, esz = sizeof(T);
    const int64_t lds_max = t.lds_cap;
    if (R * esz <= lds_max) {
        const int64_t W = col_strip_width(R, C, esz, t);
        bool by_choice = false;
        if (t.opt_scratch && W < C && W * esz < 64) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            by_choice = hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone;
        }
        if (!by_choice) {
            const uint32_t nstrips = (uint32_t)((C + W - 1) / W);
            const size_t lds = (size_t)(R * W * esz);
            auto k = k_col_gather_lds<T, Ix, Map>;
            ipt_status stt = allow_lds(ctx, k, lds);
            if (stt != IPT_OK) return stt;
            int grid;
            grid_for(ctx, k, t.block, lds, d3 * nstrips, t.grid_mult, &grid);
            hipLaunchKernelGGL(k, dim3(grid), dim3(t.block), lds, st, data, p, d3, (uint32_t)W, nstrips);
            IPT_LAUNCH_CHECK("k_col_gather_lds");
            return IPT_OK;
        }
    }
    // Scratch tier: gather strips into scratch, copy back.  Wide strips pay
    // where a strip's row segment is contiguous on one side (every map but
    // a general one; all four column maps here are such).
    const uint32_t W = (Map::kRunConstant || Map::kDiag) ? (uint32_t)(256 / esz)
                                                          : (sizeof(T) == 4 ? 16u : 8u);
    const uint32_t wshift = (uint32_t)ceil_log2(W);
    const uint64_t nstrips = ((uint64_t)C + W - 1) / W;
    const uint64_t items = d3 * nstrips;
    const uint64_t per_item = (uint64_t)R * W;
    Arena& arena = t.scratch ? *t.scratch : ctx->scratch;
    const uint64_t cap = arena.fixed ? arena.bytes : kScratchCap;
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(cap / esz, (uint64_t)INT32_MAX) / per_item);
    if (per_item > (uint64_t)INT32_MAX) return IPT_ERR_TOO_LARGE;  // one strip beyond 32-bit element math
    const Div64 dstrips = Div64::make(nstrips);
    std::lock_guard<std::mutex> g(ctx->mu);
    void* scr;
    ipt_status stt = arena_acquire(arena, (size_t)(std::min(chunk, items) * per_item * esz), st, &scr);
    if (stt != IPT_OK) return stt;
    for (uint64_t first = 0; first < items; first += chunk) {
        const uint64_t n = std::min(chunk, items - first);
        const uint32_t total = (uint32_t)(n * per_item);
        const Div32 drow = Div32::make((uint32_t)(n * W));  // a scratch row: the chunk's n strips
        const int grid = (int)std::min<uint64_t>(((uint64_t)total + 255) / 256, (uint64_t)ncu_of(ctx) * 16);
        if constexpr (Map::kDiag != 0) {
            // k rows per unit: a band of kb + W - 1 rows of W * esz = 256 bytes
            // in at most kDiagBandBytes of LDS; row segments move as the widest
            // vectors the row pitch and the data's alignment allow
            const uint32_t kb = (uint32_t)std::min<int64_t>(R, kDiagBandBytes / (W * esz) - (W - 1));
            const uintptr_t align = (uintptr_t)data | (uintptr_t)(C * esz);
            if (align % 16 == 0)
                stt = launch_diag<T, uint4, Ix, Map::kDiag == 1>(ctx, data, (T*)scr, p, wshift, dstrips,
                                                                 first, n, kb, st);
            else if (align % 8 == 0)
                stt = launch_diag<T, uint2, Ix, Map::kDiag == 1>(ctx, data, (T*)scr, p, wshift, dstrips,
                                                                 first, n, kb, st);
            else
                stt = launch_diag<T, uint32_t, Ix, Map::kDiag == 1>(ctx, data, (T*)scr, p, wshift,
                                                                    dstrips, first, n, kb, st);
            if (stt != IPT_OK) return stt;
        } else {
            hipLaunchKernelGGL((k_col_gather_to_scratch<T, Ix, Map>), dim3(grid), dim3(256), 0, st,
                               data, (T*)scr, p, wshift, dstrips, first, total, drow);
            IPT_LAUNCH_CHECK("k_col_gather_to_scratch");
        }
        hipLaunchKernelGGL((k_col_scratch_back<T>), dim3(grid), dim3(256), 0, st, data,
                           (const T*)scr, (uint64_t)R, (uint64_t)C, wshift, dstrips, first, total, drow);
        IPT_LAUNCH_CHECK("k_col_scratch_back");
    }
    return arena_release(arena, st);
}

// Row gather pass new[i][j] = old[i][s_i(j)] over every row of every slice.
template <typename T, typename Ix, typename Map>
static ipt_status row_op(DeviceCtx* ctx, T* data, const Plan<Ix>& p, const HostPlan& h,
                         uint64_t d3, const Tune& t, hipStream_t st) {
    const int64_t R = h.R, C = h.C, esz = sizeof(T);
    const int64_t row_bytes = C * esz;
    const uint64_t nrows = d3 * (uint64_t)R;
    if (row_bytes + 16 <= t.lds_cap) {
        uint32_t per = (uint32_t)std::max<int64_t>(1, (t.lds_bytes - 16) / row_bytes);
        per = (uint32_t)std::min<uint64_t>(per, nrows);
        const size_t lds = (size_t)(per * row_bytes + 16);
        auto k = k_row_gather_lds<T, Ix, Map>;
        ipt_status stt = allow_lds(ctx, k, lds);
        if (stt != IPT_OK) return stt;
        int grid;
        grid_for(ctx, k, t.block, lds, (nrows + per - 1) / per, t.grid_mult, &grid);
        hipLaunchKernelGGL(k, dim3(grid), dim3(t.block), lds, st, data, p, nrows, per,
                           Div32::make((uint32_t)C));
        IPT_LAUNCH_CHECK("k_row_gather_lds");
        return IPT_OK;
    }
    // Rows longer than LDS: gather whole rows into scratch, copy back.
    Arena& arena = t.scratch ? *t.scratch : ctx->scratch;
    const uint64_t cap = arena.fixed ? arena.bytes : kScratchCap;
    const uint64_t chunk = std::max<uint64_t>(1, cap / (uint64_t)row_bytes);
    std::lock_guard<std::mutex> g(ctx->mu);
    void* scr;
    ipt_status stt =
        arena_acquire(arena, (size_t)(std::min(chunk, nrows) * (uint64_t)row_bytes), st, &scr);
    if (stt != IPT_OK) return stt;
    for (uint64_t g0 = 0; g0 < nrows; g0 += chunk) {
        const uint64_t n = std::min(chunk, nrows - g0);
        const uint64_t elems = n * (uint64_t)C;
        const int grid = (int)std::min<uint64_t>((elems + 255) / 256, (uint64_t)ncu_of(ctx) * 16);
        hipLaunchKernelGGL((k_row_gather_to_scratch<T, Ix, Map>), dim3(grid), dim3(256), 0, st,
                           (const T*)data, (T*)scr, p, g0, n);
        IPT_LAUNCH_CHECK("k_row_gather_to_scratch");
        IPT_HIP_CHECK(hipMemcpyAsync(data + g0 * (uint64_t)C, scr, (size_t)(elems * esz),
                                     hipMemcpyDeviceToDevice, st));
    }
    return arena_release(arena, st);
}

// ---------------------------------------------------------------- C2R / R2C

// inplace::c2r::transpose (src/transpose.cu:30-47) on the view h (R x C).
template <typename T, typename Ix>
static ipt_status run_c2r_ix(DeviceCtx* ctx, T* data, const HostPlan& h, uint64_t d3,
                             const Tune& t, hipStream_t st) {
    const Plan<Ix> p = Plan<Ix>::make(h);
    ipt_status r = IPT_OK;
    if (h.c > 1) r = col_op<T, Ix, MapC2RPre>(ctx, data, p, h, d3, t, st);
    if (r == IPT_OK) r = row_op<T, Ix, MapC2RShuffle>(ctx, data, p, h, d3, t, st);
    if (r == IPT_OK) r = col_op<T, Ix, MapC2RPost>(ctx, data, p, h, d3, t, st);
    return r;
}

// inplace::r2c::transpose (src/transpose.cu:70-89) on the view h (R x C).
template <typename T, typename Ix>
static ipt_status run_r2c_ix(DeviceCtx* ctx, T* data, const HostPlan& h, uint64_t d3,
                             const Tune& t, hipStream_t st) {
    const Plan<Ix> p = Plan<Ix>::make(h);
    ipt_status r = col_op<T, Ix, MapR2CPre>(ctx, data, p, h, d3, t, st);
    if (r == IPT_OK) r = row_op<T, Ix, MapR2CShuffle>(ctx, data, p, h, d3, t, st);
    if (r == IPT_OK && h.c > 1) r = col_op<T, Ix, MapR2CPost>(ctx, data, p, h, d3, t, st);
    return r;
}

template <typename T>
static ipt_status run_decomp(DeviceCtx* ctx, T* data, const Shape& s, bool c2r, const Tune& t,
                             hipStream_t st) {
    // C2R transposes slices of R = d2 rows x C = d1 columns; R2C produces the
    // same result with the view R = d1, C = d2 (the data is its C x R form).
    const HostPlan h = c2r ? HostPlan::make(s.d2, s.d1) : HostPlan::make(s.d1, s.d2);
    const uint64_t d3 = (uint64_t)s.d3;
    if (!h.fits_index64()) return IPT_ERR_TOO_LARGE;  // no exact plan for this view
    if (h.fits_index32())
        return c2r ? run_c2r_ix<T, uint32_t>(ctx, data, h, d3, t, st)
                   : run_r2c_ix<T, uint32_t>(ctx, data, h, d3, t, st);
    return c2r ? run_c2r_ix<T, uint64_t>(ctx, data, h, d3, t, st)
               : run_r2c_ix<T, uint64_t>(ctx, data, h, d3, t, st);
}

template <typename T>
static ipt_status run_typed(DeviceCtx* ctx, T* data, const Shape& s, int variant, const Tune& t,
                            hipStream_t st) {
    switch (variant) {
    case IPT_VARIANT_SLICE_LDS: return run_slice_lds<T>(ctx, data, s, t, st);
    case IPT_VARIANT_SQUARE_SWAP: return run_square<T>(ctx, data, s, t, st);
    case IPT_VARIANT_C2R: return run_decomp<T>(ctx, data, s, true, t, st);
    case IPT_VARIANT_R2C: return run_decomp<T>(ctx, data, s, false, t, st);
    case IPT_VARIANT_GCS: return run_gcs<T>(ctx, data, s, t, st);
    case IPT_VARIANT_BLOCK_CYCLE: return run_block_cycle(ctx, data, s, t, st);
    default: return IPT_ERR_INVALID_ARG;
    }
}

ipt_status run_variant(DeviceCtx* ctx, void* data, const Shape& s, int variant, const Tune& t,
                       hipStream_t stream) {
    if (!variant_feasible(ctx, s, variant)) return IPT_ERR_UNSUPPORTED;
    // A slice with one row or one column is its own transpose in memory.
    if (s.d1 == 1 || s.d2 == 1) return IPT_OK;
    if (s.esz == 4) return run_typed<uint32_t>(ctx, (uint32_t*)data, s, variant, t, stream);
    return run_typed<uint64_t>(ctx, (uint64_t*)data, s, variant, t, stream);
}

template <typename T, typename Ix>
static ipt_status run_pass_ix(DeviceCtx* ctx, T* data, const HostPlan& h, uint64_t d3, int map_id,
                              const Tune& t, hipStream_t st) {
    const Plan<Ix> p = Plan<Ix>::make(h);
    switch (map_id) {
    case MAP_C2R_PRE: return col_op<T, Ix, MapC2RPre>(ctx, data, p, h, d3, t, st);
    case MAP_C2R_POST: return col_op<T, Ix, MapC2RPost>(ctx, data, p, h, d3, t, st);
    case MAP_R2C_PRE: return col_op<T, Ix, MapR2CPre>(ctx, data, p, h, d3, t, st);
    case MAP_R2C_POST: return col_op<T, Ix, MapR2CPost>(ctx, data, p, h, d3, t, st);
    case MAP_C2R_SHUFFLE: return row_op<T, Ix, MapC2RShuffle>(ctx, data, p, h, d3, t, st);
    case MAP_R2C_SHUFFLE: return row_op<T, Ix, MapR2CShuffle>(ctx, data, p, h, d3, t, st);
    default: return IPT_ERR_INVALID_ARG;
    }
}

ipt_status run_pass(DeviceCtx* ctx, void* data, int esz, int64_t R, int64_t C, int64_t d3,
                    int map_id, const Tune& t, hipStream_t stream) {
    const HostPlan h = HostPlan::make(R, C);
    if (!h.fits_index64()) return IPT_ERR_TOO_LARGE;
    if (h.fits_index32()) {
        if (esz == 4) return run_pass_ix<uint32_t, uint32_t>(ctx, (uint32_t*)data, h, d3, map_id, t, stream);
        return run_pass_ix<uint64_t, uint32_t>(ctx, (uint64_t*)data, h, d3, map_id, t, stream);
    }
    if (esz == 4) return run_pass_ix<uint32_t, uint64_t>(ctx, (uint32_t*)data, h, d3, map_id, t, stream);
    return run_pass_ix<uint64_t, uint64_t>(ctx, (uint64_t*)data, h, d3, map_id, t, stream);
}

template <typename T, uint32_t TR>
static ipt_status launch_oop_tr(DeviceCtx* ctx, T* dst, const T* src, const Shape& s, hipStream_t st) {
    constexpr uint32_t TS = 256 / sizeof(T), PER = 16 / sizeof(T);
    const uint32_t R = (uint32_t)s.d2, C = (uint32_t)s.d1;
    const uint32_t nti = (R + TR - 1) / TR, ntj = (C + TS - 1) / TS;
    const uint64_t tiles = (uint64_t)nti * ntj * (uint64_t)s.d3;
    const int grid = (int)std::min<uint64_t>(tiles, (uint64_t)ncu_of(ctx) * 8);
    const bool vec = R % PER == 0 && C % PER == 0 && ((uintptr_t)dst | (uintptr_t)src) % 16 == 0;
    if (vec)
        hipLaunchKernelGGL((k_oop_transpose<T, true, TR>), dim3(grid), dim3(256), 0, st, dst, src, R,
                           C, (uint64_t)s.d3, nti, ntj);
    else
        hipLaunchKernelGGL((k_oop_transpose<T, false, TR>), dim3(grid), dim3(256), 0, st, dst, src,
                           R, C, (uint64_t)s.d3, nti, ntj);
    IPT_LAUNCH_CHECK("k_oop_transpose");
    return IPT_OK;
}

template <typename T>
static ipt_status launch_oop(DeviceCtx* ctx, T* dst, const T* src, const Shape& s, hipStream_t st) {
    // square tiles: half-height ones gained under 10 % on 12 MB slices (whose
    // calls are launch-bound) and lost up to 25 % on skinny and large ones
    return launch_oop_tr<T, 256 / sizeof(T)>(ctx, dst, src, s, st);
}

ipt_status run_oop(DeviceCtx* ctx, void* dst, const void* src, const Shape& s, hipStream_t st) {
    if (s.d1 >= ((int64_t)1 << 32) || s.d2 >= ((int64_t)1 << 32)) return IPT_ERR_TOO_LARGE;
    if (s.esz == 4) return launch_oop<uint32_t>(ctx, (uint32_t*)dst, (const uint32_t*)src, s, st);
    return launch_oop<uint64_t>(ctx, (uint64_t*)dst, (const uint64_t*)src, s, st);
}

ipt_status run_copy(DeviceCtx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (((uintptr_t)dst | (uintptr_t)src | bytes) % 16) return IPT_ERR_INVALID_ARG;
    const uint64_t n16 = bytes / 16;
    if (n16 == 0) return IPT_OK;
    // blocks of 256 KB where the copy is large enough to give every one of
    // 8 workgroups per CU several; smaller blocks (down to 16 KB) below that
    const uint64_t slots = (uint64_t)ncu_of(ctx) * 8;
    const uint64_t bv = std::max<uint64_t>(1024, std::min<uint64_t>(16384, (n16 / slots) / 2048 * 2048));
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((n16 + bv - 1) / bv, slots));
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, st, (u32x4*)dst, (const u32x4*)src, n16, bv);
    IPT_LAUNCH_CHECK("k_copy16");
    return IPT_OK;
}

}  // namespace ipt

using namespace ipt;

ipt_status ipt_debug_gcs_deps(int elem_bytes, int64_t d1, int64_t d2, uint32_t g, uint32_t K,
                              int banded, uint32_t* lo, uint32_t* hi) {
    if ((elem_bytes != 4 && elem_bytes != 8) || d1 <= 0 || d2 <= 0 || g == 0 || K == 0 || !lo || !hi)
        return IPT_ERR_INVALID_ARG;
    const uint64_t TS = elem_bytes == 4 ? 32 : 16;
    if ((uint64_t)d1 * (uint64_t)d2 >= ((uint64_t)1 << 31)) return IPT_ERR_TOO_LARGE;
    const uint32_t R = (uint32_t)d2, C = (uint32_t)d1;
    const uint32_t nti = (uint32_t)((R + TS - 1) / TS), ntj = (uint32_t)((C + TS - 1) / TS);
    const uint32_t ntiles = nti * ntj;
    if ((uint64_t)g * K < ntiles || (uint64_t)g * K >= ((uint64_t)1 << 31)) return IPT_ERR_INVALID_ARG;
    const TileOrder ord = banded ? banded_order(kGcsBlock / 64, nti, ntj) : TileOrder{};
    for (uint32_t m = 0; m < g; m++) {
        uint32_t nk;
        if (banded) nk = m * K < ntiles ? std::min(K, ntiles - m * K) : 0u;
        else nk = m < ntiles ? std::min(K, (ntiles - m + g - 1) / g) : 0u;
        if (elem_bytes == 4) gcs_deps<32>(m, nk, K, ntiles, g, R, C, ntj, ord, &lo[m], &hi[m]);
        else gcs_deps<16>(m, nk, K, ntiles, g, R, C, ntj, ord, &lo[m], &hi[m]);
    }
    return IPT_OK;
}
