// ipt_block_cycle.hip — the block-cycle path (IPT_VARIANT_BLOCK_CYCLE).
//
// One side of an R x C slice (the chunked side, length L = nb * T) is cut
// into chunks of T elements; the slice is then transposed in two steps:
//   chunk step  the chunks move as a matrix of chunks is transposed, by
//               segmented cycle following (the reference's cycle-following
//               permute, src/permute.cu:18-137, rebuilt on whole chunks);
//   block step  nb blocks of other x T (or T x other) elements, each a
//               contiguous slice, are transposed in place by a one-sweep
//               kernel: the square swap, the cooperative kernel (GCS) or the
//               whole-slice kernel.
// Chunking the columns (C = nb * T) runs the chunk step first, chunking the
// rows (R = nb * T) runs the block step first.  With a short side s <= 128
// and blocks that fit LDS this is the skinny path (src/skinny.cu:313-363,
// rebuilt); with larger blocks it covers big rectangular slices that fit
// neither the square swap nor one cooperative group.  Cycles are found and
// cut into segments on the device once per chunk-matrix shape (cached per
// device);
// every segment's first chunk is saved to scratch, and then each segment is
// moved by one workgroup with no synchronisation between segments.  Both
// steps stream contiguous vectors (16 bytes where the chunk size and the
// data's alignment allow, else 8 or 4): about two sweeps.
#include "ipt_internal.h"
#include "ipt_kernels.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <set>
#include <tuple>
#include <vector>

namespace ipt {

// ---------------------------------------------------------------------------
// Chunk permutation.  A slice is a sequence of N = rc * cc chunks of cv
// vectors (V: 16, 8 or 4 bytes), and the content of chunk q moves to
//     P(q) = (q mod cc) * rc + q / cc
// (the transpose of an rc x cc matrix of chunks).  The cycles of P are
// found once per shape (on the device, below) and written as a move plan:
// work item i is moves [m0, m0 + nm) of (source, destination) chunk
// positions, nm <= ell.
//   - A cycle of at most ell positions is moved whole by one item (several
//     short cycles share an item); all its chunks are staged before any is
//     written, so it needs nothing else.
//   - A longer cycle is cut into segments of ell positions, one item each:
//     the segment's first chunk was saved beforehand by k_chunk_save (its
//     source is head slot `s`, flagged kHeadFlag; it is always move 0), the
//     item stages the rest, and writes every chunk to the next position of
//     the cycle, the last one to the first position of the next segment,
//     whose content was saved.
// Fixed points have no moves.  An item reads only positions it also writes,
// and no position is written by two items, so once the heads are saved (a
// kernel boundary) items need no synchronisation at all.
constexpr int kChunkBlock = 256;
constexpr uint32_t kHeadFlag = 0x80000000u;

// head[(slice * nheads + h) * cv + v] = vector v of chunk hpos[h] of the
// slice.  Lanes past the end repeat the last vector (same address, value).
template <typename V>
__global__ void __launch_bounds__(kChunkBlock)
k_chunk_save(const V* __restrict__ data, V* __restrict__ head, const uint32_t* __restrict__ hpos,
             uint32_t nheads, uint64_t nslices, uint64_t slice_vecs, uint32_t cv, Div32 dcv) {
    const uint64_t per_slice = (uint64_t)nheads * cv;
    const uint64_t total = nslices * per_slice;
    constexpr uint32_t U = 4 * 16 / sizeof(V);  // 16 KB in flight per workgroup
    for (uint64_t e0 = (uint64_t)blockIdx.x * kChunkBlock * U; e0 < total;
         e0 += (uint64_t)gridDim.x * kChunkBlock * U) {
        V r[U];
        uint64_t dst[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint64_t e = umin64(e0 + u * kChunkBlock + threadIdx.x, total - 1);
            const uint64_t sl = e / per_slice;
            const uint32_t x = (uint32_t)(e - sl * per_slice);
            const uint32_t h = dcv.div(x), v = x - h * cv;
            dst[u] = e;
            r[u] = data[sl * slice_vecs + (uint64_t)hpos[h] * cv + v];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) head[dst[u]] = r[u];
    }
}

// One workgroup per (slice, item, part): the item's moves are read into LDS,
// part `part` (vectors [v0, v0 + pv)) of its source chunks staged in LDS (the
// saved head's from `head`, the rest from the slice), then each staged part
// written to its destination.  Every part of a chunk follows the same cycle,
// so the parts of one item are independent work; splitting large chunks
// lets a segment hold up to kBcStageBytes / part chunks instead of
// kBcStageBytes / chunk, and fewer segments means fewer saved heads.
// Staging lanes past the end repeat the last vector, so the loads carry no
// branches and a batch of U of them is in flight at once.  dpv divides by
// the part length, dpv_last by the last part's (cv - (nparts - 1) * pv).
template <typename V>
__global__ void __launch_bounds__(kChunkBlock)
k_chunk_cycle(V* __restrict__ data, const V* __restrict__ head, const uint2* __restrict__ items,
              const uint32_t* __restrict__ msrc, const uint32_t* __restrict__ mdst, uint32_t nitems,
              uint32_t nheads, uint64_t nslices, uint64_t slice_vecs, uint32_t cv, uint32_t nparts,
              Div32 dpv, Div32 dpv_last) {
    IPT_DYN_LDS(V, stage);
    __shared__ uint32_t src[kBcMaxSeg], dst[kBcMaxSeg];
    // 32 KB of loads in flight per workgroup whatever the vector width (with
    // 8 loads of 4 bytes per thread the chunk step of an f32 1 GB slice ran
    // at 3.1 TB/s)
    constexpr uint32_t U = 8 * 16 / sizeof(V);
    const uint64_t total_items = nslices * nitems * nparts;
    for (uint64_t it = blockIdx.x; it < total_items; it += gridDim.x) {
        const uint64_t seg = it / nparts;
        const uint32_t part = (uint32_t)(it - seg * nparts);
        const uint64_t sl = seg / nitems;
        const uint2 item = items[seg - sl * nitems];
        const uint32_t nm = item.y;
        const Div32 dp = part + 1 == nparts ? dpv_last : dpv;
        const uint32_t pv = dp.d;
        const uint32_t v0 = part * dpv.d;
        V* base = data + sl * slice_vecs + v0;
        for (uint32_t j = threadIdx.x; j < nm; j += kChunkBlock) {
            src[j] = msrc[item.x + j];
            dst[j] = mdst[item.x + j];
        }
        __syncthreads();
        uint32_t first = 0;
        if (src[0] & kHeadFlag) {  // a cut cycle's segment: its head was saved
            const V* h = head + (sl * nheads + (src[0] & ~kHeadFlag)) * cv + v0;
            for (uint32_t v = threadIdx.x; v < pv; v += kChunkBlock) stage[v] = h[v];
            first = 1;
        }
        const uint32_t total = (nm - first) * pv;
        for (uint32_t e0 = 0; e0 < total; e0 += U * kChunkBlock) {
            V r[U];
            uint32_t at[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t e = min(e0 + u * kChunkBlock + threadIdx.x, total - 1);
                const uint32_t j = dp.div(e), v = e - j * pv;
                at[u] = first * pv + e;
                r[u] = base[(uint64_t)src[first + j] * cv + v];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) stage[at[u]] = r[u];
        }
        __syncthreads();
        const uint32_t wtotal = nm * pv;
        for (uint32_t e = threadIdx.x; e < wtotal; e += kChunkBlock) {
            const uint32_t j = dp.div(e), v = e - j * pv;
            base[(uint64_t)dst[j] * cv + v] = stage[e];
        }
        __syncthreads();  // stage, src and dst are reused by the next item
    }
}

// ---------------------------------------------------------------------------
// Plan.  Candidates: both orientations, every divisor T of the chunked side
// with T <= L / 2 (at least two blocks), a chunk of kBcMinChunk ..
// kBcMaxChunk bytes, fewer than kBcMaxChunks per slice and a block some
// one-sweep kernel runs.  Chunks move as 16-byte vectors when their size is
// a multiple of 16, else as 8- or 4-byte ones.  Each is priced in sweeps:
//   chunk step  (1 + 1.5 / ell): the saved heads are read and written once
//               and read again, 1/ell of the data; chunks under 2 KB cost
//               extra (short contiguous runs per position), and so do
//               narrower vectors (x1.15 for 8 bytes, x1.4 for 4);
//   block step  1 / efficiency of its kernel relative to a copy (measured:
//               square swap ~1, whole-slice 16-byte form ~0.9, cooperative
//               kernel ~0.95 with groups on one XCD, ~0.8 across XCDs, ~0.6
//               with the largest register-tile class, ~0.4 when a block
//               side is narrower than the wave tile, whole-slice scalar
//               form ~0.5).
// The cheapest wins; ties go to the smaller chunk (smaller blocks).
static int block_kernel(const DeviceCtx* ctx, const Shape& blk, int jr_cls, double* eff) {
    const int64_t bytes = blk.d1 * blk.d2 * blk.esz;
    if (blk.d1 == blk.d2) { *eff = 1.0; return IPT_VARIANT_SQUARE_SWAP; }
    const GcsGeom gg = gcs_geom(ctx, blk, lds_max_of(ctx), true, jr_cls, true);
    const bool slice = bytes <= kBcLdsBlockBytes && variant_feasible(ctx, blk, IPT_VARIANT_SLICE_LDS);
    if (slice && (bytes % 16) == 0) { *eff = 0.9; return IPT_VARIANT_SLICE_LDS; }
    if (gg.ok) {
        const int nx = std::max(ctx ? ctx->nxcd : 1, 1);
        const int64_t tile = blk.esz == 4 ? 32 : 16;  // wave tile side
        const bool packed2 = gg.grid && gg.grid / gg.g >= 2;  // largest class by choice
        if (std::min(blk.d1, blk.d2) < tile) *eff = 0.4;  // mostly idle lanes
        else if (gcs_jr_class(blk.esz, gg.K, gg.KL, kGcsBlock / 64) == 2 && !packed2) *eff = 0.6;
        else *eff = (int64_t)gg.g <= ncu_of(ctx) / nx ? 0.95 : 0.8;
        return IPT_VARIANT_GCS;
    }
    if (slice) { *eff = 0.5; return IPT_VARIANT_SLICE_LDS; }
    return IPT_VARIANT_AUTO;
}

static void plan_orientation(const DeviceCtx* ctx, const Shape& s, bool cols, int64_t cmin,
                             int64_t cmax, int jr_cls, BlockCyclePlan* best, double* best_cost) {
    const int64_t esz = s.esz;
    const int64_t L = cols ? s.d1 : s.d2, other = cols ? s.d2 : s.d1;
    const int64_t tmax = std::min<int64_t>(L / 2, cmax / esz);
    const int64_t tmin = (cmin + esz - 1) / esz;
    for (int64_t T = tmax; T >= tmin; T--) {
        if (L % T) continue;
        const int64_t nb = L / T;
        if (nb * other >= kBcMaxChunks) break;  // smaller T only adds chunks
        const Shape blk = cols ? make_shape(s.esz, T, other, s.d3 * nb)
                               : make_shape(s.esz, other, T, s.d3 * nb);
        double eff;
        const int v = block_kernel(ctx, blk, jr_cls, &eff);
        if (v == IPT_VARIANT_AUTO) continue;
        const int64_t chunk = T * esz;
        BlockCyclePlan c{};
        c.ok = true;
        c.chunk_first = cols;
        c.T = T;
        c.nb = nb;
        c.other = other;
        c.rc = cols ? other : nb;
        c.cc = cols ? nb : other;
        c.vb = chunk % 16 == 0 ? 16 : chunk % 8 == 0 ? 8 : 4;
        c.blk = blk;
        c.block_variant = v;
        chunk_parts(&c);
        const double short_run = chunk < 2048 ? 1.0 + (2048.0 - chunk) / 4096.0 : 1.0;
        const double narrow = c.vb == 16 ? 1.0 : c.vb == 8 ? 1.15 : 1.4;
        const double cost = (1.0 + 1.5 / c.ell) * short_run * narrow + 1.0 / eff;
        // ties go to the smaller chunk (T falls): its blocks are smaller
        if (best->ok && cost > *best_cost + 1e-9) continue;
        *best = c;
        *best_cost = cost;
    }
}

// Chunks of more than 2 * kBcPartBytes move in nsub parts of at least
// kBcPartBytes; a segment stages up to kBcStageBytes of parts.  Narrowing the
// vectors later (block_cycle_for) keeps nsub and ell: a part then holds at
// most as many bytes as before, so the staging still fits.
void chunk_parts(BlockCyclePlan* p) {
    const int64_t chunk = p->T * p->blk.esz;
    p->cv = chunk / p->vb;
    p->nsub = std::max<int64_t>(1, chunk / kBcPartBytes);
    if (p->nsub > p->cv) p->nsub = p->cv;
    const int64_t part = chunk_part_vecs(*p) * p->vb;
    p->ell = std::max<int64_t>(1, std::min<int64_t>(kBcMaxSeg, kBcStageBytes / part));
}

BlockCyclePlan block_cycle_for(const BlockCyclePlan& p, const void* data) {
    BlockCyclePlan q = p;
    if (!q.ok) return q;
    while (q.vb > 4 && (uintptr_t)data % (uintptr_t)q.vb != 0) q.vb /= 2;
    if ((uintptr_t)data % 4 != 0) {
        q.ok = false;
        return q;
    }
    q.cv = q.T * q.blk.esz / q.vb;
    return q;
}

BlockCyclePlan block_cycle_plan(const DeviceCtx* ctx, const Shape& s, int64_t chunk_min,
                                int64_t chunk_max, int jr_cls) {
    BlockCyclePlan p{};
    if (s.d1 == s.d2 || s.d1 < 2 || s.d2 < 2) return p;
    const int64_t cmin = std::max<int64_t>(chunk_min > 0 ? chunk_min : kBcMinChunk, kBcMinChunk);
    const int64_t cmax = std::min<int64_t>(chunk_max > 0 ? chunk_max : kBcMaxChunk, kBcMaxChunk);
    if (cmin > cmax) return p;
    double cost = 0;
    plan_orientation(ctx, s, true, cmin, cmax, jr_cls, &p, &cost);
    plan_orientation(ctx, s, false, cmin, cmax, jr_cls, &p, &cost);
    return p;
}

// Move lists of the chunk permutation walked on the host, cycle by cycle
// from each cycle's least position (whole short cycles bundled up to ell
// moves).  The library builds its lists on the device (below); this walk is
// the cross-check behind ipt_debug_chunk_moves.
static void build_moves(const BlockCyclePlan& p, std::vector<uint2>* items,
                        std::vector<uint32_t>* msrc, std::vector<uint32_t>* mdst,
                        std::vector<uint32_t>* hpos) {
    const uint32_t N = (uint32_t)(p.rc * p.cc), rc = (uint32_t)p.rc, cc = (uint32_t)p.cc;
    const uint32_t ell = (uint32_t)p.ell;
    std::vector<uint8_t> seen(N, 0);
    std::vector<uint32_t> cyc;
    uint32_t bundle = 0;  // moves in the open item of whole short cycles
    auto close_bundle = [&] {
        if (bundle) items->push_back(make_uint2((uint32_t)msrc->size() - bundle, bundle));
        bundle = 0;
    };
    for (uint32_t q0 = 0; q0 < N; q0++) {
        if (seen[q0]) continue;
        cyc.clear();
        uint32_t q = q0;
        do {
            seen[q] = 1;
            cyc.push_back(q);
            q = (q % cc) * rc + q / cc;
        } while (q != q0);
        const uint32_t n = (uint32_t)cyc.size();
        if (n == 1) continue;  // fixed point
        if (n <= ell) {
            if (bundle + n > ell) close_bundle();
            for (uint32_t j = 0; j < n; j++) {
                msrc->push_back(cyc[j]);
                mdst->push_back(cyc[(j + 1) % n]);
            }
            bundle += n;
            continue;
        }
        close_bundle();
        for (uint32_t a = 0; a < n; a += ell) {
            const uint32_t len = std::min(ell, n - a);
            items->push_back(make_uint2((uint32_t)msrc->size(), len));
            for (uint32_t j = 0; j < len; j++) {
                msrc->push_back(j == 0 ? (kHeadFlag | (uint32_t)hpos->size()) : cyc[a + j]);
                mdst->push_back(cyc[(a + j + 1) % n]);
            }
            hpos->push_back(cyc[a]);
        }
    }
    close_bundle();
}

// ---------------------------------------------------------------------------
// Move lists built on the device.  P(q) = (q mod cc) * rc + q / cc fixes 0
// and N - 1 and is multiplication by rc modulo M = N - 1 on the rest (q = a
// cc + b: q rc = a N + b rc = a + b rc mod M), so no cycle is longer than
// ord_M(rc), the length of the cycle of 1.  Every position learns its
// cycle's least position (its leader) and its distance to it by pointer
// doubling over records (next, least, distance to the least's first
// occurrence): after round k a record covers q, P(q), .., P^(2^k - 1)(q),
// and ceil(log2 ord_M(rc)) rounds cover every cycle.  A leader l's cycle has
// L = dist(P(l)) + 1 positions, and q is at rank (L - dist(q)) mod L of its
// cycle counted from l.  One exclusive scan over per-leader counts lays the
// cycles out, and then every position writes its own move:
//   a cycle of more than B = ell / 2 positions gets items of its own,
//     segments of up to ell positions, each with a saved head when the
//     cycle is longer than ell (the host walk's form);
//   shorter cycles are packed by the offset e of their last move among
//     the short cycles' moves: those with e in [b B, (b + 1) B) form item b,
//     which starts at most B - 1 moves before b B (at most (B - 1) + B < ell
//     moves) and is never empty: the cycle holding move min((b + 1) B,
//     total) - 1 either ends in bin b or starts after b B, and then its
//     predecessor ends in bin b.
// Work: ceil(log2 ord) gathers of 16-byte records per position plus a few
// streaming passes, against the host walk's serial pass over all N
// positions (measured on MI355X and its host, lists copied back included:
// 2^23 positions 18.8 vs 74.1 ms, 2^22 9.6 vs 37.5, 2^18 0.54 vs 1.52,
// 2^11 0.20 vs 0.12).
constexpr int kBuildBlock = 256;
constexpr int kScanPer = 4;  // elements per thread of the scan

__device__ __forceinline__ uint32_t chunk_perm(uint32_t q, uint32_t rc, Div32 dcc) {
    const uint32_t a = dcc.div(q);
    return (q - a * dcc.d) * rc + a;
}
__device__ __forceinline__ uint4 add4(uint4 a, uint4 b) {
    return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

__global__ void __launch_bounds__(kBuildBlock)
k_bc_init(uint4* __restrict__ rec, uint32_t N, uint32_t rc, Div32 dcc) {
    const uint32_t q = blockIdx.x * kBuildBlock + threadIdx.x;
    if (q < N) rec[q] = make_uint4(chunk_perm(q, rc, dcc), q, 0, 0);
}

// One doubling round: the window of q (length step) joined with the window
// of the position step places ahead; ties keep the nearer occurrence.
__global__ void __launch_bounds__(kBuildBlock)
k_bc_double(const uint4* __restrict__ in, uint4* __restrict__ out, uint32_t N, uint32_t step) {
    const uint32_t q = blockIdx.x * kBuildBlock + threadIdx.x;
    if (q >= N) return;
    const uint4 r = in[q];
    const uint4 a = in[r.x];
    out[q] = a.y < r.y ? make_uint4(a.x, a.y, step + a.z, 0) : make_uint4(a.x, r.y, r.z, 0);
}

// Leaders record their cycle length (rec[l].w) and their counts: (moves of
// own items, own items, saved heads, short-cycle moves).  Only the .z words
// are read across positions here and only .w words written.
__global__ void __launch_bounds__(kBuildBlock)
k_bc_leaders(uint4* __restrict__ rec, uint4* __restrict__ cnt, uint32_t N, uint32_t rc, Div32 dcc,
             uint32_t ell, uint32_t B) {
    const uint32_t q = blockIdx.x * kBuildBlock + threadIdx.x;
    if (q >= N) return;
    uint32_t* w = reinterpret_cast<uint32_t*>(rec);
    uint4 c = make_uint4(0, 0, 0, 0);
    if (w[4 * (size_t)q + 1] == q) {
        const uint32_t L = w[4 * (size_t)chunk_perm(q, rc, dcc) + 2] + 1;
        w[4 * (size_t)q + 3] = L;
        if (L >= 2) {
            if (L > B) {
                const uint32_t segs = (L + ell - 1) / ell;
                c = make_uint4(L, segs, L > ell ? segs : 0, 0);
            } else {
                c.w = L;
            }
        }
    }
    cnt[q] = c;
}

// Exclusive scan of uint4 counts in place: per-block sums, a single-block
// scan of those (tot = the grand total), then each block rescanned with its
// offset.  kScanPer consecutive elements per thread.
__device__ __forceinline__ uint4 block_exclusive(uint4 v, uint4* sh, uint4* total) {
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kBuildBlock; off <<= 1) {
        const uint4 o = t >= off ? sh[t - off] : make_uint4(0, 0, 0, 0);
        __syncthreads();
        sh[t] = add4(sh[t], o);
        __syncthreads();
    }
    const uint4 incl = sh[t];
    *total = sh[kBuildBlock - 1];
    __syncthreads();
    return make_uint4(incl.x - v.x, incl.y - v.y, incl.z - v.z, incl.w - v.w);
}

__global__ void __launch_bounds__(kBuildBlock)
k_bc_scan_up(const uint4* __restrict__ v, uint4* __restrict__ bsum, uint32_t N) {
    __shared__ uint4 sh[kBuildBlock];
    const uint64_t e0 = ((uint64_t)blockIdx.x * kBuildBlock + threadIdx.x) * kScanPer;
    uint4 s = make_uint4(0, 0, 0, 0);
    for (uint32_t k = 0; k < kScanPer; k++)
        if (e0 + k < N) s = add4(s, v[e0 + k]);
    uint4 total;
    (void)block_exclusive(s, sh, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBuildBlock)
k_bc_scan_top(uint4* __restrict__ bsum, uint32_t nb, uint4* __restrict__ tot) {
    __shared__ uint4 sh[kBuildBlock];
    uint4 carry = make_uint4(0, 0, 0, 0);
    for (uint32_t b0 = 0; b0 < nb; b0 += kBuildBlock) {
        const uint32_t b = b0 + threadIdx.x;
        const uint4 v = b < nb ? bsum[b] : make_uint4(0, 0, 0, 0);
        uint4 total;
        const uint4 ex = block_exclusive(v, sh, &total);
        if (b < nb) bsum[b] = add4(carry, ex);
        carry = add4(carry, total);
    }
    if (threadIdx.x == 0) *tot = carry;
}

__global__ void __launch_bounds__(kBuildBlock)
k_bc_scan_down(uint4* __restrict__ v, const uint4* __restrict__ bsum, uint32_t N) {
    __shared__ uint4 sh[kBuildBlock];
    const uint64_t e0 = ((uint64_t)blockIdx.x * kBuildBlock + threadIdx.x) * kScanPer;
    uint4 x[kScanPer];
    uint4 s = make_uint4(0, 0, 0, 0);
    for (uint32_t k = 0; k < kScanPer; k++) {
        x[k] = e0 + k < N ? v[e0 + k] : make_uint4(0, 0, 0, 0);
        s = add4(s, x[k]);
    }
    uint4 total;
    uint4 run = add4(bsum[blockIdx.x], block_exclusive(s, sh, &total));
    for (uint32_t k = 0; k < kScanPer; k++) {
        if (e0 + k < N) v[e0 + k] = run;
        run = add4(run, x[k]);
    }
}

__global__ void __launch_bounds__(kBuildBlock)
k_bc_fill(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

// Every position of a cycle of two or more writes its move; rank-0
// positions of short cycles post their offset to the bin of their cycle's
// last move, segment starts write their item (and head).  own = (moves, items) of all own-item cycles.
__global__ void __launch_bounds__(kBuildBlock)
k_bc_emit(const uint4* __restrict__ rec, const uint4* __restrict__ base, uint32_t N, uint32_t rc,
          Div32 dcc, uint32_t ell, uint32_t B, uint32_t own_moves, uint2* __restrict__ items,
          uint32_t* __restrict__ msrc, uint32_t* __restrict__ mdst, uint32_t* __restrict__ hpos,
          uint32_t* __restrict__ first) {
    const uint32_t q = blockIdx.x * kBuildBlock + threadIdx.x;
    if (q >= N) return;
    const uint4 r = rec[q];
    const uint32_t L = rec[r.y].w;
    if (L < 2) return;
    const uint32_t rank = (L - r.z) % L;
    const uint4 b = base[r.y];
    const uint32_t dst = chunk_perm(q, rc, dcc);
    if (L > B) {
        const uint32_t idx = b.x + rank, a = rank / ell, j = rank - a * ell;
        uint32_t src = q;
        if (j == 0) {
            items[b.y + a] = make_uint2(idx, min(ell, L - a * ell));
            if (L > ell) {
                hpos[b.z + a] = q;
                src = kHeadFlag | (b.z + a);
            }
        }
        msrc[idx] = src;
        mdst[idx] = dst;
    } else {
        const uint32_t idx = own_moves + b.w + rank;
        msrc[idx] = q;
        mdst[idx] = dst;
        if (rank == 0) atomicMin(&first[(b.w + L - 1) / B], b.w);
    }
}

__global__ void __launch_bounds__(kBuildBlock)
k_bc_bins(uint2* __restrict__ items, const uint32_t* __restrict__ first, uint32_t nbins,
          uint32_t short_moves, uint32_t own_moves, uint32_t own_items) {
    const uint32_t b = blockIdx.x * kBuildBlock + threadIdx.x;
    if (b >= nbins) return;
    const uint32_t s0 = first[b], s1 = b + 1 < nbins ? first[b + 1] : short_moves;
    items[own_items + b] = make_uint2(own_moves + s0, s1 - s0);
}

// Multiplicative order of a modulo m (gcd(a, m) = 1, 2 <= m < 2^32): the
// Carmichael function of m reduced by its prime factors while a still
// reaches 1.
static uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    a %= m;
    for (; e; e >>= 1, a = a * a % m)
        if (e & 1) r = r * a % m;
    return r;
}
static std::vector<uint64_t> prime_factors(uint64_t n) {
    std::vector<uint64_t> f;
    for (uint64_t p = 2; p * p <= n; p++)
        if (n % p == 0) {
            f.push_back(p);
            while (n % p == 0) n /= p;
        }
    if (n > 1) f.push_back(n);
    return f;
}
static uint64_t mul_order(uint64_t a, uint64_t m) {
    uint64_t lam = 1, n = m;
    for (uint64_t p : prime_factors(m)) {
        uint64_t pe = 1;
        int e = 0;
        while (n % p == 0) {
            n /= p;
            pe *= p;
            e++;
        }
        const uint64_t l = p == 2 ? (e <= 2 ? pe / 2 : pe / 4) : pe / p * (p - 1);
        lam = lam / std::gcd(lam, l) * l;
    }
    uint64_t ord = lam;
    for (uint64_t f : prime_factors(lam))
        while (ord % f == 0 && powmod(a, ord / f, m) == 1) ord /= f;
    return ord;
}

template <typename V>
static hipError_t upload(const std::vector<V>& h, V** d) {
    *d = nullptr;
    if (h.empty()) return hipSuccess;
    hipError_t e = hipMalloc(d, h.size() * sizeof(V));
    if (e == hipSuccess) e = hipMemcpy(*d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice);
    return e;
}

void chunk_moves_free(ChunkMoves* m) {
    (void)hipFree(m->items);
    (void)hipFree(m->msrc);
    (void)hipFree(m->mdst);
    (void)hipFree(m->hpos);
    *m = ChunkMoves{};
}

static ipt_status chunk_moves_host(const BlockCyclePlan& p, ChunkMoves* out) {
    std::vector<uint2> items;
    std::vector<uint32_t> msrc, mdst, hpos;
    build_moves(p, &items, &msrc, &mdst, &hpos);
    ChunkMoves m;
    m.nitems = (uint32_t)items.size();
    m.nheads = (uint32_t)hpos.size();
    m.nmoves = (uint32_t)msrc.size();
    m.bytes = items.size() * sizeof(uint2) + (msrc.size() + mdst.size() + hpos.size()) * 4;
    hipError_t e = upload(items, &m.items);
    if (e == hipSuccess) e = upload(msrc, &m.msrc);
    if (e == hipSuccess) e = upload(mdst, &m.mdst);
    if (e == hipSuccess) e = upload(hpos, &m.hpos);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        chunk_moves_free(&m);
        set_error("move plan upload", e);
        return IPT_ERR_OUT_OF_MEMORY;
    }
    *out = m;
    return IPT_OK;
}

// The device build, on the null stream (ordered like the synchronous copies
// it replaces); returns once the lists are complete.
static ipt_status chunk_moves_device(const BlockCyclePlan& p, ChunkMoves* out) {
    *out = ChunkMoves{};
    const uint64_t N64 = (uint64_t)p.rc * (uint64_t)p.cc;
    if (p.rc <= 1 || p.cc <= 1) return IPT_OK;  // the identity
    if (N64 >= ((uint64_t)1 << 31) || p.ell < 1) return IPT_ERR_UNSUPPORTED;
    const uint32_t N = (uint32_t)N64, rc = (uint32_t)p.rc, ell = (uint32_t)p.ell, B = ell / 2;
    const Div32 dcc = Div32::make((uint32_t)p.cc);
    const uint64_t ord = mul_order(rc, N - 1);
    uint32_t rounds = 0;
    while (((uint64_t)1 << rounds) < ord) rounds++;
    const uint32_t grid = (N + kBuildBlock - 1) / kBuildBlock;
    const uint32_t nsb = (uint32_t)((N + (uint64_t)kBuildBlock * kScanPer - 1) /
                                    ((uint64_t)kBuildBlock * kScanPer));
    hipStream_t st = nullptr;
    uint4 *ra = nullptr, *rb = nullptr, *bsum = nullptr;
    uint32_t* first = nullptr;
    ChunkMoves m;
    ipt_status result = IPT_OK;
    uint4 tot{};
    auto fail = [&](const char* what, hipError_t e, ipt_status s) {
        (void)hipGetLastError();
        set_error(what, e);
        result = s;
    };
    hipError_t e = hipMalloc(&ra, (size_t)N * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(&rb, (size_t)N * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(&bsum, ((size_t)nsb + 1) * sizeof(uint4));
    if (e != hipSuccess) fail("move plan scratch", e, IPT_ERR_OUT_OF_MEMORY);
    if (result == IPT_OK) {
        hipLaunchKernelGGL(k_bc_init, dim3(grid), dim3(kBuildBlock), 0, st, ra, N, rc, dcc);
        for (uint32_t k = 0; k < rounds; k++) {
            hipLaunchKernelGGL(k_bc_double, dim3(grid), dim3(kBuildBlock), 0, st, (const uint4*)ra,
                               rb, N, (uint32_t)1 << k);
            std::swap(ra, rb);
        }
        hipLaunchKernelGGL(k_bc_leaders, dim3(grid), dim3(kBuildBlock), 0, st, ra, rb, N, rc, dcc,
                           ell, B);
        hipLaunchKernelGGL(k_bc_scan_up, dim3(nsb), dim3(kBuildBlock), 0, st, (const uint4*)rb,
                           bsum, N);
        hipLaunchKernelGGL(k_bc_scan_top, dim3(1), dim3(kBuildBlock), 0, st, bsum, nsb, bsum + nsb);
        hipLaunchKernelGGL(k_bc_scan_down, dim3(nsb), dim3(kBuildBlock), 0, st, rb,
                           (const uint4*)bsum, N);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(&tot, bsum + nsb, sizeof(uint4), hipMemcpyDeviceToHost);
        if (e != hipSuccess) fail("move plan build", e, IPT_ERR_HIP);
    }
    // tot = (moves of own-item cycles, their items, saved heads, short-cycle moves)
    const uint32_t nbins = B ? (tot.w + B - 1) / B : 0;
    if (result == IPT_OK) {
        m.nitems = tot.y + nbins;
        m.nheads = tot.z;
        m.nmoves = tot.x + tot.w;
        m.bytes = (size_t)m.nitems * sizeof(uint2) + ((size_t)2 * m.nmoves + m.nheads) * 4;
        if (m.nitems) e = hipMalloc(&m.items, (size_t)m.nitems * sizeof(uint2));
        if (e == hipSuccess && m.nmoves) e = hipMalloc(&m.msrc, (size_t)m.nmoves * 4);
        if (e == hipSuccess && m.nmoves) e = hipMalloc(&m.mdst, (size_t)m.nmoves * 4);
        if (e == hipSuccess && m.nheads) e = hipMalloc(&m.hpos, (size_t)m.nheads * 4);
        if (e == hipSuccess && nbins) e = hipMalloc(&first, (size_t)nbins * 4);
        if (e != hipSuccess) fail("move plan lists", e, IPT_ERR_OUT_OF_MEMORY);
    }
    if (result == IPT_OK && m.nmoves) {
        if (nbins)
            hipLaunchKernelGGL(k_bc_fill, dim3((nbins + kBuildBlock - 1) / kBuildBlock),
                               dim3(kBuildBlock), 0, st, first, nbins, UINT32_MAX);
        hipLaunchKernelGGL(k_bc_emit, dim3(grid), dim3(kBuildBlock), 0, st, (const uint4*)ra,
                           (const uint4*)rb, N, rc, dcc, ell, B, tot.x, m.items, m.msrc, m.mdst,
                           m.hpos, first);
        if (nbins)
            hipLaunchKernelGGL(k_bc_bins, dim3((nbins + kBuildBlock - 1) / kBuildBlock),
                               dim3(kBuildBlock), 0, st, m.items, (const uint32_t*)first, nbins,
                               tot.w, tot.x, tot.y);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) fail("move plan build", e, IPT_ERR_HIP);
    }
    (void)hipFree(ra);
    (void)hipFree(rb);
    (void)hipFree(bsum);
    (void)hipFree(first);
    if (result != IPT_OK) {
        chunk_moves_free(&m);
        return result;
    }
    *out = m;
    return IPT_OK;
}

ipt_status chunk_moves_create(const BlockCyclePlan& p, ChunkMoves* out) {
    return chunk_moves_device(p, out);
}

ipt_status chunk_moves_debug(int64_t rc, int64_t cc, int64_t ell, bool device, ChunkMoves* out) {
    BlockCyclePlan p{};
    p.rc = rc;
    p.cc = cc;
    p.ell = ell;
    return device ? chunk_moves_device(p, out) : chunk_moves_host(p, out);
}

// Move lists used by the per-call path, built once per (device, rc, cc, ell)
// and kept in device memory; past kBcPlanCacheBytes the cache is emptied
// (after a device synchronise: queued kernels may still read the lists)
// before the next one is added.  Lists that a stream capture has used are
// not freed then but retired for the rest of the process: the captured
// graph keeps their addresses.
struct CachedMoves {
    ChunkMoves m;
    bool captured = false;
};
static std::mutex g_plan_mu;
static std::map<std::tuple<int, int64_t, int64_t, int64_t>, CachedMoves> g_plans;
static std::vector<ChunkMoves> g_retired_plans;
static size_t g_plan_bytes = 0;  // device bytes held by g_plans

static ipt_status cached_moves(const DeviceCtx* ctx, const BlockCyclePlan& p, hipStream_t st,
                               ChunkMoves* out) {
    std::lock_guard<std::mutex> g(g_plan_mu);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
        (void)hipGetLastError();
        cs = hipStreamCaptureStatusNone;
    }
    const bool capturing = cs != hipStreamCaptureStatusNone;
    const auto key = std::make_tuple(ctx->dev, p.rc, p.cc, p.ell);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
        if (capturing) it->second.captured = true;
        *out = it->second.m;
        return IPT_OK;
    }
    if (capturing) return IPT_ERR_UNSUPPORTED;  // building the lists allocates
    ChunkMoves m;
    const ipt_status stt = chunk_moves_create(p, &m);
    if (stt != IPT_OK) return stt;
    const size_t bytes = m.bytes;
    if (g_plan_bytes + bytes > kBcPlanCacheBytes && !g_plans.empty()) {
        const hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            chunk_moves_free(&m);
            set_error("hipDeviceSynchronize", e);
            return IPT_ERR_HIP;
        }
        for (auto& kv : g_plans) {
            if (kv.second.captured) g_retired_plans.push_back(kv.second.m);
            else chunk_moves_free(&kv.second.m);
        }
        g_plans.clear();
        g_plan_bytes = 0;
    }
    g_plans[key] = CachedMoves{m, false};
    g_plan_bytes += bytes;
    *out = m;
    return IPT_OK;
}

ipt_status chunk_kernels_setup(const DeviceCtx* ctx) {
    // staging (up to kBcStageBytes) plus the static move lists exceed the
    // default dynamic-LDS limit
    static std::set<int> raised;
    std::lock_guard<std::mutex> g(g_plan_mu);
    if (!raised.count(ctx->dev)) {
        for (const void* k : {reinterpret_cast<const void*>(k_chunk_cycle<uint4>),
                              reinterpret_cast<const void*>(k_chunk_cycle<uint2>),
                              reinterpret_cast<const void*>(k_chunk_cycle<uint32_t>)})
            IPT_HIP_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)kBcStageBytes));
        raised.insert(ctx->dev);
    }
    return IPT_OK;
}

template <typename V>
static ipt_status chunk_step_launch_v(const DeviceCtx* ctx, void* data, uint64_t nslices,
                                     uint64_t slice_vecs, const BlockCyclePlan& p,
                                     const ChunkMoves& m, void* head, uint64_t group,
                                     hipStream_t st) {
    const uint32_t pv = (uint32_t)chunk_part_vecs(p), nparts = (uint32_t)chunk_nparts(p);
    const size_t lds = (size_t)p.ell * pv * sizeof(V);
    const Div32 dcv = Div32::make((uint32_t)p.cv);
    const Div32 dpv = Div32::make(pv), dpv_last = Div32::make((uint32_t)p.cv - (nparts - 1) * pv);
    for (uint64_t s0 = 0; s0 < nslices; s0 += group) {
        const uint64_t ns = std::min<uint64_t>(group, nslices - s0);
        V* d = static_cast<V*>(data) + s0 * slice_vecs;
        if (m.nheads) {
            const uint64_t save_vecs = ns * m.nheads * (uint64_t)p.cv;
            const uint64_t per_wg = (uint64_t)kChunkBlock * 4 * 16 / sizeof(V);
            const int gs = (int)std::min<uint64_t>((save_vecs + per_wg - 1) / per_wg,
                                                   (uint64_t)ncu_of(ctx) * 8);
            hipLaunchKernelGGL(k_chunk_save<V>, dim3(gs), dim3(kChunkBlock), 0, st, (const V*)d,
                               (V*)head, (const uint32_t*)m.hpos, m.nheads, ns, slice_vecs,
                               (uint32_t)p.cv, dcv);
            IPT_LAUNCH_CHECK("k_chunk_save");
        }
        const int gc = (int)std::min<uint64_t>(ns * m.nitems * nparts, (uint64_t)ncu_of(ctx) * 8);
        hipLaunchKernelGGL(k_chunk_cycle<V>, dim3(gc), dim3(kChunkBlock), lds, st, d, (const V*)head,
                           (const uint2*)m.items, (const uint32_t*)m.msrc, (const uint32_t*)m.mdst,
                           m.nitems, m.nheads, ns, slice_vecs, (uint32_t)p.cv, nparts, dpv, dpv_last);
        IPT_LAUNCH_CHECK("k_chunk_cycle");
    }
    return IPT_OK;
}

ipt_status chunk_step_launch(const DeviceCtx* ctx, void* data, uint64_t nslices,
                             uint64_t slice_vecs, const BlockCyclePlan& p, const ChunkMoves& m,
                             void* head, uint64_t group, hipStream_t st) {
    if (m.nitems == 0) return IPT_OK;
    switch (p.vb) {
    case 16: return chunk_step_launch_v<uint4>(ctx, data, nslices, slice_vecs, p, m, head, group, st);
    case 8: return chunk_step_launch_v<uint2>(ctx, data, nslices, slice_vecs, p, m, head, group, st);
    case 4: return chunk_step_launch_v<uint32_t>(ctx, data, nslices, slice_vecs, p, m, head, group, st);
    default: return IPT_ERR_INVALID_ARG;
    }
}

// The chunk step of the per-call path: cached move lists, heads saved in the
// scratch arena, in groups of slices whose heads fit the scratch cap.
static ipt_status chunk_step(DeviceCtx* ctx, void* data, const Shape& s, const BlockCyclePlan& p,
                             hipStream_t st) {
    ChunkMoves m;
    ipt_status stt = cached_moves(ctx, p, st, &m);
    if (stt != IPT_OK) return stt;
    if (m.nitems == 0) return IPT_OK;
    const uint64_t heads_per_slice = chunk_head_bytes(p, m);
    if (heads_per_slice > kScratchCap) return IPT_ERR_UNSUPPORTED;
    const uint64_t group =
        heads_per_slice ? std::max<uint64_t>(1, kScratchCap / heads_per_slice) : (uint64_t)s.d3;
    stt = chunk_kernels_setup(ctx);
    if (stt != IPT_OK) return stt;
    std::lock_guard<std::mutex> g(ctx->mu);
    void* head = nullptr;
    if (heads_per_slice) {
        stt = arena_acquire(ctx->scratch,
                            (size_t)(std::min<uint64_t>(group, (uint64_t)s.d3) * heads_per_slice), st,
                            &head);
        if (stt != IPT_OK) return stt;
    }
    const uint64_t slice_vecs = (uint64_t)s.d1 * s.d2 * s.esz / (uint64_t)p.vb;
    stt = chunk_step_launch(ctx, data, (uint64_t)s.d3, slice_vecs, p, m, head, group, st);
    if (!heads_per_slice) return stt;
    const ipt_status r = arena_release(ctx->scratch, st);
    return stt != IPT_OK ? stt : r;
}

// Transpose of `slices` rows x cols matrices of `inner`-element chunks (the
// data is [slices][rows][cols][inner], slowest first; the result
// [slices][cols][rows][inner]): the chunk step alone.  Chunks of any size
// (moved in parts); UNSUPPORTED past kBcMaxChunks chunks per slice, or when
// the saved heads of one slice exceed the scratch cap.
ipt_status run_chunk_transpose(DeviceCtx* ctx, void* data, int esz, int64_t inner, int64_t rows,
                               int64_t cols, int64_t slices, hipStream_t st) {
    if (rows == 1 || cols == 1) return IPT_OK;
    const int64_t chunk = inner * esz;
    if (rows * cols >= kBcMaxChunks || chunk / 4 > (int64_t)UINT32_MAX / kBcMaxSeg)
        return IPT_ERR_UNSUPPORTED;
    BlockCyclePlan p{};
    p.ok = true;
    p.T = inner;
    p.blk = make_shape(esz, inner, 1, 1);
    p.rc = rows;
    p.cc = cols;
    p.vb = chunk % 16 == 0 ? 16 : chunk % 8 == 0 ? 8 : 4;
    chunk_parts(&p);
    p = block_cycle_for(p, data);
    if (!p.ok) return IPT_ERR_UNSUPPORTED;
    return chunk_step(ctx, data, make_shape(esz, inner, rows * cols, slices), p, st);
}

ipt_status run_block_cycle(DeviceCtx* ctx, void* data, const Shape& s, const Tune& t,
                           hipStream_t st) {
    const BlockCyclePlan p = block_cycle_for(block_cycle_plan(ctx, s, t.chunk_min, t.chunk_max), data);
    if (!p.ok) return IPT_ERR_UNSUPPORTED;
    if (p.chunk_first) {
        const ipt_status stt = chunk_step(ctx, data, s, p, st);
        return stt == IPT_OK ? run_variant(ctx, data, p.blk, p.block_variant, t, st) : stt;
    }
    const ipt_status stt = run_variant(ctx, data, p.blk, p.block_variant, t, st);
    return stt == IPT_OK ? chunk_step(ctx, data, s, p, st) : stt;
}

}  // namespace ipt
