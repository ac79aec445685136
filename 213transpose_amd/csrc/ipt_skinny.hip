// ipt_skinny.hip — the skinny path (IPT_VARIANT_SKINNY).
//
// A slice of R x C with one short side s (s <= kSkinnyMaxShort) and long
// side L = M * B is transposed in two steps:
//   block step  the slice's M blocks of B x s (tall: B rows of s) or s x B
//               (wide) elements are transposed in place, one block per LDS
//               tile (the library's whole-slice kernels on d3 * M slices);
//   chunk step  the B-element chunks are moved as an rc x cc matrix of
//               chunks is transposed: (rc, cc) = (M, s) after the block step
//               (tall), (s, M) before it (wide).
// The chunk step is cycle following (the reference's cycle-following permute,
// src/permute.cu:18-137, rebuilt): cycles are cut into segments on the host
// once per shape (cached per device), every segment's first chunk is saved
// to scratch, and then each segment is moved by one workgroup with no
// synchronisation between segments.  Both steps stream contiguous 16-byte
// vectors; the path costs two sweeps plus the saved chunks (1/ell of one).
#include "ipt_internal.h"
#include "ipt_kernels.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <vector>

namespace ipt {

// ---------------------------------------------------------------------------
// Chunk permutation.  A slice is a sequence of N = rc * cc chunks of cv
// 16-byte vectors, and the content of chunk q moves to
//     P(q) = (q mod cc) * rc + q / cc
// (the transpose of an rc x cc matrix of chunks).  Each cycle of P is cut
// into segments of at most ell consecutive positions p_0, p_1 = P(p_0), ...;
// segment w starts at segs[w].x and has segs[w].y positions.  k_chunk_save
// copies the first chunk p_0 of every segment into `bnd`; k_chunk_cycle then
// writes the content of p_j to p_{j+1} for every j < n (p_n is the first
// position of the cycle's next segment, whose content k_chunk_save saved).
// A segment's worker reads only p_1 .. p_{n-1}, which no other segment
// writes, so once the save has completed (a kernel boundary) segments need
// no synchronisation at all.
constexpr int kChunkBlock = 256;

__device__ __forceinline__ uint32_t chunk_next(uint32_t q, uint32_t rc, Div32 dcc) {
    const uint32_t a = dcc.div(q);
    return (q - a * dcc.d) * rc + a;
}

// bnd[(slice * nseg + w) * cv + v] = vector v of chunk segs[w].x of the slice.
// Lanes past the end repeat the last vector (same address, same value).
__global__ void __launch_bounds__(kChunkBlock)
k_chunk_save(const uint4* __restrict__ data, uint4* __restrict__ bnd, const uint2* __restrict__ segs,
             uint32_t nseg, uint64_t nslices, uint64_t slice_vecs, uint32_t cv, Div32 dcv) {
    const uint64_t per_slice = (uint64_t)nseg * cv;
    const uint64_t total = nslices * per_slice;
    constexpr uint32_t U = 4;
    for (uint64_t e0 = (uint64_t)blockIdx.x * kChunkBlock * U; e0 < total;
         e0 += (uint64_t)gridDim.x * kChunkBlock * U) {
        uint4 r[U];
        uint64_t dst[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint64_t e = umin64(e0 + u * kChunkBlock + threadIdx.x, total - 1);
            const uint64_t sl = e / per_slice;
            const uint32_t x = (uint32_t)(e - sl * per_slice);
            const uint32_t w = dcv.div(x), v = x - w * cv;
            dst[u] = e;
            r[u] = data[sl * slice_vecs + (uint64_t)segs[w].x * cv + v];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) bnd[dst[u]] = r[u];
    }
}

// One workgroup per (slice, segment): stage the saved first chunk and chunks
// p_1 .. p_{n-1} in LDS, then write staged chunk j to p_{j+1}.  Lanes past
// the end repeat the last vector, so the loads carry no branches and a batch
// of U of them is in flight at once.
__global__ void __launch_bounds__(kChunkBlock)
k_chunk_cycle(uint4* __restrict__ data, const uint4* __restrict__ bnd, const uint2* __restrict__ segs,
              uint32_t nseg, uint64_t nslices, uint64_t slice_vecs, uint32_t cv, Div32 dcv,
              uint32_t rc, Div32 dcc) {
    extern __shared__ __align__(16) uint4 stage[];
    __shared__ uint32_t pos[kSkinnyMaxSeg + 1];
    constexpr uint32_t U = 8;
    const uint64_t items = nslices * nseg;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint64_t sl = it / nseg;
        const uint32_t w = (uint32_t)(it - sl * nseg);
        uint4* base = data + sl * slice_vecs;
        const uint2 sg = segs[w];
        const uint32_t n = sg.y;
        if (threadIdx.x == 0) {
            uint32_t p = sg.x;
            pos[0] = p;
            for (uint32_t j = 1; j <= n; j++) {
                p = chunk_next(p, rc, dcc);
                pos[j] = p;
            }
        }
        const uint4* b = bnd + it * cv;  // slot 0: the saved first chunk
        for (uint32_t v = threadIdx.x; v < cv; v += kChunkBlock) stage[v] = b[v];
        __syncthreads();
        const uint32_t total = (n - 1) * cv;  // slots 1 .. n-1: chunks p_1 .. p_{n-1}
        for (uint32_t e0 = 0; e0 < total; e0 += U * kChunkBlock) {
            uint4 r[U];
            uint32_t at[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t e = min(e0 + u * kChunkBlock + threadIdx.x, total - 1);
                const uint32_t j = dcv.div(e), v = e - j * cv;
                at[u] = cv + e;
                r[u] = base[(uint64_t)pos[1 + j] * cv + v];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) stage[at[u]] = r[u];
        }
        __syncthreads();
        const uint32_t wtotal = n * cv;
        for (uint32_t e = threadIdx.x; e < wtotal; e += kChunkBlock) {
            const uint32_t j = dcv.div(e), v = e - j * cv;
            base[(uint64_t)pos[j + 1] * cv + v] = stage[e];
        }
        __syncthreads();  // stage and pos are reused by the next item
    }
}

// ---------------------------------------------------------------------------
// Plan: B is the largest divisor of L below L (so M >= 2) with an s x B
// block of at most kSkinnyBlockBytes and a chunk of at least
// kSkinnyMinChunk bytes (a multiple of 16); ell = segment length in chunks,
// from the staging budget.
SkinnyPlan skinny_plan(const Shape& s) {
    SkinnyPlan p{};
    const int64_t R = s.d2, C = s.d1, esz = s.esz;
    if (R == C || R < 2 || C < 2) return p;
    p.tall = R > C;
    p.s = std::min(R, C);
    p.L = std::max(R, C);
    if (p.s > kSkinnyMaxShort) return p;
    const int64_t bmax = std::min<int64_t>(p.L / 2, kSkinnyBlockBytes / (p.s * esz));  // M >= 2
    const int64_t bmin = (kSkinnyMinChunk + esz - 1) / esz;
    for (int64_t B = bmax; B >= bmin; B--)
        if (p.L % B == 0 && (B * esz) % 16 == 0) {
            p.B = B;
            break;
        }
    if (p.B == 0) return p;
    p.M = p.L / p.B;
    if (p.M < 2 || p.M * p.s >= ((int64_t)1 << 26)) return p;
    p.rc = p.tall ? p.M : p.s;
    p.cc = p.tall ? p.s : p.M;
    p.cv = p.B * esz / 16;
    p.ell = std::max<int64_t>(1, std::min<int64_t>(kSkinnyMaxSeg, kSkinnyStageBytes / (p.B * esz)));
    p.ok = true;
    return p;
}

// Segment tables, built on the host once per (device, rc, cc, ell) and kept
// in device memory until the process ends.
struct SegTable {
    uint2* dev = nullptr;
    uint32_t nseg = 0;
};
static std::mutex g_seg_mu;
static std::map<std::tuple<int, int64_t, int64_t, int64_t>, SegTable> g_segs;

static ipt_status seg_table(const DeviceCtx* ctx, const SkinnyPlan& p, hipStream_t st,
                            SegTable* out) {
    std::lock_guard<std::mutex> g(g_seg_mu);
    const auto key = std::make_tuple(ctx->dev, p.rc, p.cc, p.ell);
    auto it = g_segs.find(key);
    if (it != g_segs.end()) {
        *out = it->second;
        return IPT_OK;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return IPT_ERR_UNSUPPORTED;  // building a table allocates
    const uint32_t N = (uint32_t)(p.rc * p.cc), rc = (uint32_t)p.rc, cc = (uint32_t)p.cc;
    std::vector<uint8_t> seen(N, 0);
    std::vector<uint2> segs;
    for (uint32_t q0 = 0; q0 < N; q0++) {
        if (seen[q0]) continue;
        uint32_t q = q0, n = 0;
        do {
            if (n == 0) segs.push_back(make_uint2(q, 0));
            seen[q] = 1;
            segs.back().y++;
            n = (n + 1 == (uint32_t)p.ell) ? 0 : n + 1;
            q = (q % cc) * rc + q / cc;
        } while (q != q0);
    }
    SegTable t;
    t.nseg = (uint32_t)segs.size();
    hipError_t e = hipMalloc(&t.dev, segs.size() * sizeof(uint2));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc(segments)", e);
        return IPT_ERR_OUT_OF_MEMORY;
    }
    e = hipMemcpy(t.dev, segs.data(), segs.size() * sizeof(uint2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(t.dev);
        set_error("hipMemcpy(segments)", e);
        return IPT_ERR_HIP;
    }
    g_segs[key] = t;
    *out = t;
    return IPT_OK;
}

// The chunk step over slices [0, d3): saves then cycles, in groups of slices
// whose saved chunks fit the scratch cap.
static ipt_status chunk_step(DeviceCtx* ctx, void* data, const Shape& s, const SkinnyPlan& p,
                             hipStream_t st) {
    SegTable tab;
    ipt_status stt = seg_table(ctx, p, st, &tab);
    if (stt != IPT_OK) return stt;
    const uint64_t chunk_bytes = (uint64_t)p.cv * 16;
    const uint64_t bnd_per_slice = (uint64_t)tab.nseg * chunk_bytes;
    if (bnd_per_slice > kScratchCap) return IPT_ERR_UNSUPPORTED;
    const uint64_t group = std::max<uint64_t>(1, kScratchCap / bnd_per_slice);
    const uint64_t slice_vecs = (uint64_t)p.L * p.s * s.esz / 16;
    const size_t lds = (size_t)p.ell * chunk_bytes;
    {
        // staging (up to kSkinnyStageBytes) plus the static position table
        // exceed the default dynamic-LDS limit: raise it once per device
        static std::set<int> raised;
        std::lock_guard<std::mutex> g(g_seg_mu);
        if (!raised.count(ctx->dev)) {
            IPT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_chunk_cycle),
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)kSkinnyStageBytes));
            raised.insert(ctx->dev);
        }
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    void* bnd;
    stt = arena_acquire(ctx->scratch, (size_t)(std::min<uint64_t>(group, (uint64_t)s.d3) * bnd_per_slice),
                        st, &bnd);
    if (stt != IPT_OK) return stt;
    const Div32 dcv = Div32::make(p.cv), dcc = Div32::make((uint32_t)p.cc);
    for (uint64_t s0 = 0; s0 < (uint64_t)s.d3; s0 += group) {
        const uint64_t ns = std::min<uint64_t>(group, (uint64_t)s.d3 - s0);
        uint4* d = static_cast<uint4*>(data) + s0 * slice_vecs;
        const uint64_t save_items = ns * tab.nseg * (uint64_t)p.cv;
        const int gs = (int)std::min<uint64_t>((save_items + kChunkBlock * 4 - 1) / (kChunkBlock * 4),
                                               (uint64_t)ncu_of(ctx) * 8);
        hipLaunchKernelGGL(k_chunk_save, dim3(gs), dim3(kChunkBlock), 0, st, (const uint4*)d,
                           (uint4*)bnd, (const uint2*)tab.dev, tab.nseg, ns, slice_vecs,
                           (uint32_t)p.cv, dcv);
        IPT_LAUNCH_CHECK("k_chunk_save");
        const int gc = (int)std::min<uint64_t>(ns * tab.nseg, (uint64_t)ncu_of(ctx) * 8);
        hipLaunchKernelGGL(k_chunk_cycle, dim3(gc), dim3(kChunkBlock), lds, st, d, (const uint4*)bnd,
                           (const uint2*)tab.dev, tab.nseg, ns, slice_vecs, (uint32_t)p.cv, dcv,
                           (uint32_t)p.rc, dcc);
        IPT_LAUNCH_CHECK("k_chunk_cycle");
    }
    return arena_release(ctx->scratch, st);
}

ipt_status run_skinny(DeviceCtx* ctx, void* data, const Shape& s, const Tune& t, hipStream_t st) {
    const SkinnyPlan p = skinny_plan(s);
    if (!p.ok || (uintptr_t)data % 16 != 0) return IPT_ERR_UNSUPPORTED;
    // Block step: tall blocks are B rows of s (213 with d1 = s, d2 = B);
    // wide blocks, after the chunk step, are s rows of B (d1 = B, d2 = s).
    const Shape blk = p.tall ? make_shape(s.esz, p.s, p.B, s.d3 * p.M)
                             : make_shape(s.esz, p.B, p.s, s.d3 * p.M);
    // Blocks fit one LDS tile: the cooperative kernel's vector path (both
    // block dimensions multiples of 16 bytes) or else the whole-slice kernel.
    const GcsGeom gg = gcs_geom(ctx, blk, t.lds_cap, true, true);
    const int bv = gg.ok && gg.vec ? IPT_VARIANT_GCS : IPT_VARIANT_SLICE_LDS;
    ipt_status stt = IPT_OK;
    if (p.tall) {
        stt = run_variant(ctx, data, blk, bv, t, st);
        if (stt == IPT_OK) stt = chunk_step(ctx, data, s, p, st);
    } else {
        stt = chunk_step(ctx, data, s, p, st);
        if (stt == IPT_OK) stt = run_variant(ctx, data, blk, bv, t, st);
    }
    return stt;
}

}  // namespace ipt
