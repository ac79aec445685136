// ipt_kernels.h — gfx950 kernels of the in-place transpose.
//
// Every kernel moves payload bits only (T is an unsigned integer of the
// element width) and is a grid-stride loop over independent work items, so
// any grid size is correct.  Each workgroup stages everything it will
// overwrite in LDS (or in scratch for the fallback tiers) before the first
// store, which is what makes the passes safe in place: work items never
// share memory.
#pragma once

#include "ipt_plan.h"

#include <type_traits>

namespace ipt {

IPT_HD uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// Dynamic LDS, 16-byte aligned.
#define IPT_DYN_LDS(type, name) extern __shared__ __align__(16) unsigned char ipt_lds_raw_[]; \
    type* name = reinterpret_cast<type*>(ipt_lds_raw_)

// Copy n contiguous elements global -> LDS.  `l` must have the same address
// phase modulo 16 as `g`; the aligned body then moves 16 bytes per lane.
template <typename T>
__device__ __forceinline__ void copy_in(const T* __restrict__ g, T* l, uint32_t n) {
    constexpr uint32_t PER = 16 / sizeof(T);
    const uint32_t tid = threadIdx.x, nthr = blockDim.x;
    uint32_t head = (uint32_t)(((16u - ((uintptr_t)g & 15u)) & 15u) / sizeof(T));
    if (head > n) head = n;
    for (uint32_t e = tid; e < head; e += nthr) l[e] = g[e];
    const uint32_t nv = (n - head) / PER;
    const uint4* gv = reinterpret_cast<const uint4*>(g + head);
    uint4* lv = reinterpret_cast<uint4*>(l + head);
    for (uint32_t v = tid; v < nv; v += nthr) lv[v] = gv[v];
    for (uint32_t e = head + nv * PER + tid; e < n; e += nthr) l[e] = g[e];
}

// Store n contiguous elements g[e] = val(e); 16 bytes per lane in the
// aligned body.
template <typename T, typename F>
__device__ __forceinline__ void store_out(T* __restrict__ g, uint32_t n, F val) {
    constexpr uint32_t PER = 16 / sizeof(T);
    const uint32_t tid = threadIdx.x, nthr = blockDim.x;
    uint32_t head = (uint32_t)(((16u - ((uintptr_t)g & 15u)) & 15u) / sizeof(T));
    if (head > n) head = n;
    for (uint32_t e = tid; e < head; e += nthr) g[e] = val(e);
    const uint32_t nv = (n - head) / PER;
    uint4* gv = reinterpret_cast<uint4*>(g + head);
    for (uint32_t v = tid; v < nv; v += nthr) {
        union { uint4 v4; T e[PER]; } u;
#pragma unroll
        for (uint32_t t = 0; t < PER; t++) u.e[t] = val(head + v * PER + t);
        gv[v] = u.v4;
    }
    for (uint32_t e = head + nv * PER + tid; e < n; e += nthr) g[e] = val(e);
}

// ---------------------------------------------------------------------------
// Plain copy (calibration reference for bandwidth and traffic counters):
// 16 bytes per lane, grid-stride.  n16 = number of 16-byte words.
__global__ void __launch_bounds__(256)
k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Whole slices in LDS: one sweep.  Each work item is `per` consecutive slices
// of R x C; the slices are read once into LDS with an odd row pitch P (so a
// column walk touches distinct banks) and written back transposed.
template <typename T>
__global__ void __launch_bounds__(1024)
k_slice_lds(T* __restrict__ data, uint32_t R, uint32_t C, uint32_t P, uint64_t nslices,
            uint32_t per, Div32 dS, Div32 dC, Div32 dR) {
    IPT_DYN_LDS(T, lds);
    const uint32_t S = R * C;
    const uint32_t SP = R * P;
    for (uint64_t s0 = (uint64_t)blockIdx.x * per; s0 < nslices; s0 += (uint64_t)gridDim.x * per) {
        const uint32_t ns = (uint32_t)umin64(per, nslices - s0);
        T* base = data + s0 * S;
        const uint32_t n = ns * S;
        for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
            uint32_t sl = dS.div(e), r = e - sl * S;
            uint32_t i = dC.div(r), j = r - i * C;
            lds[sl * SP + i * P + j] = base[e];
        }
        __syncthreads();
        for (uint32_t o = threadIdx.x; o < n; o += blockDim.x) {
            uint32_t sl = dS.div(o), r = o - sl * S;
            uint32_t jj = dR.div(r), ii = r - jj * R;  // output row jj (input column), col ii
            base[o] = lds[sl * SP + ii * P + jj];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Square slices N x N: swap-transpose tile pairs (I,J) / (J,I), I <= J.
// One sweep; both tiles are staged in LDS before either is written.
template <typename T, int TS>
__global__ void __launch_bounds__(256)
k_square_swap(T* __restrict__ data, uint32_t N, uint32_t ntiles, uint64_t pairs, uint64_t d3) {
    __shared__ T A[TS][TS + 1];
    __shared__ T B[TS][TS + 1];
    const uint32_t tx = threadIdx.x % TS, ty = threadIdx.x / TS;
    constexpr uint32_t ROWS = 256 / TS;
    const uint64_t total = pairs * d3;
    for (uint64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const uint64_t sl = w / pairs;
        const uint64_t t = w - sl * pairs;
        // t = J(J+1)/2 + I with I <= J.
        uint64_t J = (uint64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while (J * (J + 1) / 2 > t) J--;
        while ((J + 1) * (J + 2) / 2 <= t) J++;
        const uint64_t I = t - J * (J + 1) / 2;
        (void)ntiles;
        T* base = data + sl * (uint64_t)N * N;
        const uint32_t r0 = (uint32_t)I * TS, c0 = (uint32_t)J * TS;
        for (uint32_t y = ty; y < TS; y += ROWS) {
            uint32_t r = r0 + y, c = c0 + tx;
            if (r < N && c < N) A[y][tx] = base[(uint64_t)r * N + c];
            uint32_t r2 = c0 + y, c2 = r0 + tx;
            if (I != J && r2 < N && c2 < N) B[y][tx] = base[(uint64_t)r2 * N + c2];
        }
        __syncthreads();
        for (uint32_t y = ty; y < TS; y += ROWS) {
            // tile (I,J) <- transpose of tile (J,I); diagonal: of itself.
            uint32_t r = r0 + y, c = c0 + tx;
            if (r < N && c < N) base[(uint64_t)r * N + c] = (I == J) ? A[tx][y] : B[tx][y];
            uint32_t r2 = c0 + y, c2 = r0 + tx;
            if (I != J && r2 < N && c2 < N) base[(uint64_t)r2 * N + c2] = A[tx][y];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Row gather with the rows in LDS: new[i][j] = old[i][s_i(j)].  Each work
// item is `per` consecutive rows (a contiguous range of per*C elements).
template <typename T, typename Ix, typename Map>
__global__ void __launch_bounds__(1024)
k_row_gather_lds(T* __restrict__ data, Plan<Ix> p, uint64_t nrows, uint32_t per, Div32 dC) {
    IPT_DYN_LDS(T, lds_base);
    const uint32_t C = (uint32_t)p.C;
    for (uint64_t g0 = (uint64_t)blockIdx.x * per; g0 < nrows; g0 += (uint64_t)gridDim.x * per) {
        const uint32_t n = (uint32_t)umin64(per, nrows - g0);
        T* row = data + g0 * C;
        T* lds = lds_base + (((uintptr_t)row & 15u) / sizeof(T));
        copy_in(row, lds, n * C);
        __syncthreads();
        const Ix i0 = (Ix)(g0 % (uint64_t)p.R);
        if (n == 1) {
            store_out(row, C, [&](uint32_t j) { return lds[(uint32_t)Map::src(p, i0, (Ix)j)]; });
        } else {
            store_out(row, n * C, [&](uint32_t e) {
                uint32_t r = dC.div(e), j = e - r * C;
                Ix i = p.dR.mod(i0 + (Ix)r);
                return lds[r * C + (uint32_t)Map::src(p, i, (Ix)j)];
            });
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Column gather with a strip of W columns x all R rows in LDS:
// new[i][j] = old[g(i,j)][j].  W is a power of two dividing blockDim.x.
template <typename T, typename Ix, typename Map>
__global__ void __launch_bounds__(1024)
k_col_gather_lds(T* __restrict__ data, Plan<Ix> p, uint64_t d3, uint32_t W, uint32_t nstrips) {
    IPT_DYN_LDS(T, lds);
    const uint32_t R = (uint32_t)p.R;
    const uint64_t C = (uint64_t)p.C;
    const uint32_t jj = threadIdx.x & (W - 1);
    const uint32_t rstep = blockDim.x / W;
    const uint64_t items = d3 * nstrips;
    for (uint64_t w = blockIdx.x; w < items; w += gridDim.x) {
        const uint64_t sl = w / nstrips;
        const uint64_t j0 = (w - sl * nstrips) * W;
        const uint32_t wn = (uint32_t)umin64(W, C - j0);
        T* base = data + sl * (uint64_t)R * C + j0;
        if (jj < wn)
            for (uint32_t i = threadIdx.x / W; i < R; i += rstep) lds[i * W + jj] = base[i * C + jj];
        __syncthreads();
        if (jj < wn) {
            const Ix j = (Ix)(j0 + jj);
            for (uint32_t i = threadIdx.x / W; i < R; i += rstep)
                base[i * C + jj] = lds[(uint32_t)Map::src(p, (Ix)i, j) * W + jj];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Fallback tiers for columns taller / rows longer than LDS: gather into
// library scratch, then copy back.  Two launches, so the kernel boundary is
// the barrier between the last read of old data and the first overwrite.

// scratch[(w-first)][i][jj] = old[g(i,j)][j] for strips w in [first, first+n).
template <typename T, typename Ix, typename Map>
__global__ void __launch_bounds__(256)
k_col_gather_to_scratch(const T* __restrict__ data, T* __restrict__ scratch, Plan<Ix> p,
                        uint32_t W, uint32_t nstrips, uint64_t first, uint64_t n) {
    const uint64_t R = p.R, C = p.C;
    const uint64_t per_item = R * W;
    const uint64_t total = n * per_item;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t wl = e / per_item, r = e - wl * per_item;
        const uint64_t i = r / W, jj = r - i * W;
        const uint64_t w = first + wl, sl = w / nstrips;
        const uint64_t j = (w - sl * nstrips) * W + jj;
        if (j >= C) continue;
        const T* base = data + sl * R * C;
        scratch[e] = base[(uint64_t)Map::src(p, (Ix)i, (Ix)j) * C + j];
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_col_scratch_back(T* __restrict__ data, const T* __restrict__ scratch, uint64_t R, uint64_t C,
                   uint32_t W, uint32_t nstrips, uint64_t first, uint64_t n) {
    const uint64_t per_item = R * W;
    const uint64_t total = n * per_item;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t wl = e / per_item, r = e - wl * per_item;
        const uint64_t i = r / W, jj = r - i * W;
        const uint64_t w = first + wl, sl = w / nstrips;
        const uint64_t j = (w - sl * nstrips) * W + jj;
        if (j >= C) continue;
        data[sl * R * C + i * C + j] = scratch[e];
    }
}

// scratch[r][j] = old[g0+r][s_i(j)], i = (g0 + r) mod R, for rows [g0, g0+n).
template <typename T, typename Ix, typename Map>
__global__ void __launch_bounds__(256)
k_row_gather_to_scratch(const T* __restrict__ data, T* __restrict__ scratch, Plan<Ix> p,
                        uint64_t g0, uint64_t n) {
    const uint64_t C = p.C;
    const uint64_t total = n * C;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = e / C, j = e - r * C;
        const uint64_t g = g0 + r;
        const Ix i = (Ix)(g % (uint64_t)p.R);
        scratch[e] = data[g * C + (uint64_t)Map::src(p, i, (Ix)j)];
    }
}

}  // namespace ipt

namespace ipt {

// ---------------------------------------------------------------------------
// Group-cooperative staged transpose (GCS): one sweep for slices larger than
// one workgroup's LDS.  Slot s is member s % g of the group that owns slice
// s / g; workgroup b takes slots b, b + grid, b + 2*grid, ... (static
// round-robin, no global counter).  Each member stages its share of the
// slice's TS x TS tiles in LDS, the group meets at a per-slice arrival
// counter, then every member writes its tiles transposed.  All reads of a
// slice complete before any write to it and slices never overlap, so no
// workgroup ever reads data another one wrote: the counters carry no data and
// relaxed agent-scope atomics suffice.
//
// Sync state (device memory, zero at rest; each launch restores it):
//   ctl[2] watchdog flag (sticky)
//   cnt[32*S] arrivals, cnt[32*S + 16] departures of slice S (one 128-byte
//   line per slice, so groups never contend on a line).  The last member to
//   depart zeroes both.
// The spin is bounded: after ~2^26 polls the member records ctl[2] and
// proceeds, so a broken invariant shows up as an error, never as a hang.
//
// Deadlock freedom: the grid never exceeds the resident workgroups and
// g <= grid.  A member of slice S waits only for the other members of S, and
// each of those reaches S after finishing slots that are smaller than S's
// first slot (s - grid < S*g because s - S*g < g <= grid); by induction on
// slot order every group completes.

IPT_HD uint32_t atomic_add_agent(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
IPT_HD uint32_t atomic_load_agent(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
IPT_HD void atomic_store_agent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr uint32_t kSpinLimit = 1u << 26;

// Wait until *ctr reaches target (thread 0 only).  Returns false on watchdog.
__device__ __forceinline__ bool spin_until(uint32_t* ctr, uint32_t target) {
    uint32_t polls = 0;
    while (atomic_load_agent(ctr) < target) {
        __builtin_amdgcn_s_sleep(1);
        if (++polls > kSpinLimit) return false;
    }
    return true;
}

template <typename T, int TS, bool VEC>
struct GcsTile {
    static constexpr uint32_t TP = TS + 1;     // padded LDS pitch: transposed reads hit distinct banks
    static constexpr uint32_t TILE = TS * TP;  // elements per staged tile
    static constexpr uint32_t PER = VEC ? 16 / sizeof(T) : 1;
    static constexpr uint32_t VPR = TS / PER;  // vectors per tile row
    static constexpr uint32_t VPT = TS * VPR;  // vectors per tile
    using V = typename std::conditional<VEC, uint4, T>::type;
};

// Global address of vector v of tile t (input layout); false when outside.
template <typename T, int TS, bool VEC>
__device__ __forceinline__ bool gcs_src(uint32_t t, uint32_t v, uint32_t R, uint32_t C, Div32 dntj,
                                        uint64_t* off) {
    using G = GcsTile<T, TS, VEC>;
    const uint32_t ti = dntj.div(t), tj = t - ti * dntj.d;
    const uint32_t r = ti * TS + v / G::VPR, c = tj * TS + (v % G::VPR) * G::PER;
    *off = (uint64_t)r * C + c;
    return r < R && c < C;
}

// Scatter one loaded vector into its staged tile (input layout, padded).
template <typename T, int TS, bool VEC>
__device__ __forceinline__ void gcs_put(T* tile, uint32_t v, const typename GcsTile<T, TS, VEC>::V& x) {
    using G = GcsTile<T, TS, VEC>;
    const T* e = reinterpret_cast<const T*>(&x);
#pragma unroll
    for (uint32_t q = 0; q < G::PER; q++) tile[(v / G::VPR) * G::TP + (v % G::VPR) * G::PER + q] = e[q];
}

// Write staged tiles [k0, k0 + n) from LDS slots [0, n) transposed: output
// row = input column.  Four vectors per thread are read from LDS before the
// first store is issued.
template <typename T, int TS, bool VEC, int BLOCK>
__device__ __forceinline__ void gcs_store(T* __restrict__ base, const T* lds, uint32_t R, uint32_t C,
                                          Div32 dntj, uint32_t m, uint32_t g, uint32_t k0,
                                          uint32_t n) {
    using G = GcsTile<T, TS, VEC>;
    using V = typename G::V;
    constexpr uint32_t W = 4;
    const uint32_t total = n * G::VPT;
    for (uint32_t i0 = 0; i0 < total; i0 += W * BLOCK) {
        union { V v; T e[G::PER]; } u[W];
        uint64_t off[W];
        bool ok[W];
#pragma unroll
        for (uint32_t w = 0; w < W; w++) {
            const uint32_t idx = i0 + w * BLOCK + threadIdx.x;
            ok[w] = false;
            if (idx < total) {
                const uint32_t k = idx / G::VPT, x = idx % G::VPT;
                const uint32_t t = m + (k0 + k) * g;
                const uint32_t ti = dntj.div(t), tj = t - ti * dntj.d;
                const uint32_t cc = x / G::VPR, rr = (x % G::VPR) * G::PER;  // output row cc, cols rr..
                const uint32_t orow = tj * TS + cc, ocol = ti * TS + rr;
                ok[w] = orow < C && ocol < R;
                off[w] = (uint64_t)orow * R + ocol;
                const T* tile = lds + k * G::TILE;
#pragma unroll
                for (uint32_t q = 0; q < G::PER; q++) u[w].e[q] = tile[(rr + q) * G::TP + cc];
            }
        }
#pragma unroll
        for (uint32_t w = 0; w < W; w++)
            if (ok[w]) *reinterpret_cast<V*>(base + off[w]) = u[w].v;
    }
}

// GCS kernel.  Member m of slice S owns tiles t = m + k*g, k < nk; the first
// kl of them are staged in LDS, the remaining nk - kl (at most RV * BLOCK /
// VPT) stay in registers across the group barrier and go through LDS after
// the LDS-staged tiles have been written.
template <typename T, int TS, bool VEC, int BLOCK, int RV>
__global__ void __launch_bounds__(BLOCK)
k_gcs(T* __restrict__ data, uint32_t R, uint32_t C, uint64_t nslices, uint32_t g, uint32_t K,
      uint32_t KL, Div32 dntj, uint32_t ntiles, uint32_t* __restrict__ ctl,
      uint32_t* __restrict__ cnt) {
    using G = GcsTile<T, TS, VEC>;
    using V = typename G::V;
    constexpr uint32_t U = 16;  // LDS-bound vectors in flight per thread
    IPT_DYN_LDS(T, lds);
    const uint64_t total = nslices * g;
    const uint64_t slice_elems = (uint64_t)R * C;
    for (uint64_t slot = blockIdx.x; slot < total; slot += gridDim.x) {
        const uint64_t S = slot / g;
        const uint32_t m = (uint32_t)(slot - S * g);
        T* base = data + S * slice_elems;
        const uint32_t nk = m < ntiles ? min(K, (ntiles - m + g - 1) / g) : 0;
        const uint32_t kl = min(nk, KL), kr = nk - kl;

        // register-held tiles: issued first, kept across the barrier
        V rreg[RV];
        const uint32_t rtotal = kr * G::VPT;
#pragma unroll
        for (uint32_t u = 0; u < RV; u++) {
            const uint32_t idx = u * BLOCK + threadIdx.x;
            if (idx < rtotal) {
                uint64_t off;
                if (gcs_src<T, TS, VEC>(m + (kl + idx / G::VPT) * g, idx % G::VPT, R, C, dntj, &off))
                    rreg[u] = *reinterpret_cast<const V*>(base + off);
            }
        }
        // LDS-staged tiles, U vectors per thread in flight per batch
        const uint32_t ltotal = kl * G::VPT;
        for (uint32_t b0 = 0; b0 < ltotal; b0 += U * BLOCK) {
            V reg[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t idx = b0 + u * BLOCK + threadIdx.x;
                if (idx < ltotal) {
                    uint64_t off;
                    if (gcs_src<T, TS, VEC>(m + (idx / G::VPT) * g, idx % G::VPT, R, C, dntj, &off))
                        reg[u] = *reinterpret_cast<const V*>(base + off);
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t idx = b0 + u * BLOCK + threadIdx.x;
                if (idx < ltotal) gcs_put<T, TS, VEC>(lds + (idx / G::VPT) * G::TILE, idx % G::VPT, reg[u]);
            }
        }
        __syncthreads();
        if (g > 1) {
            if (threadIdx.x == 0) {
                atomic_add_agent(&cnt[32 * S], 1u);
                if (!spin_until(&cnt[32 * S], g)) atomic_store_agent(&ctl[2], 1u);
            }
            __syncthreads();
        }
        gcs_store<T, TS, VEC, BLOCK>(base, lds, R, C, dntj, m, g, 0, kl);
        if (kr) {
            __syncthreads();
#pragma unroll
            for (uint32_t u = 0; u < RV; u++) {
                const uint32_t idx = u * BLOCK + threadIdx.x;
                if (idx < rtotal) gcs_put<T, TS, VEC>(lds + (idx / G::VPT) * G::TILE, idx % G::VPT, rreg[u]);
            }
            __syncthreads();
            gcs_store<T, TS, VEC, BLOCK>(base, lds, R, C, dntj, m, g, kl, kr);
        }
        if (g > 1 && threadIdx.x == 0) {
            if (atomic_add_agent(&cnt[32 * S + 16], 1u) == g - 1) {
                atomic_store_agent(&cnt[32 * S], 0u);
                atomic_store_agent(&cnt[32 * S + 16], 0u);
            }
        }
        __syncthreads();  // LDS is reused by the next slot
    }
}

}  // namespace ipt
