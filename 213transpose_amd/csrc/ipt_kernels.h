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
// phase modulo 16 as `g`; the aligned body then moves 16 bytes per lane,
// U vectors per thread in flight (lanes past the end repeat the last vector:
// same address, same value, no branches around the loads).
template <typename T>
__device__ __forceinline__ void copy_in(const T* __restrict__ g, T* l, uint32_t n) {
    constexpr uint32_t PER = 16 / sizeof(T);
    constexpr uint32_t U = 8;
    const uint32_t tid = threadIdx.x, nthr = blockDim.x;
    uint32_t head = (uint32_t)(((16u - ((uintptr_t)g & 15u)) & 15u) / sizeof(T));
    if (head > n) head = n;
    for (uint32_t e = tid; e < head; e += nthr) l[e] = g[e];
    const uint32_t nv = (n - head) / PER;
    const uint4* gv = reinterpret_cast<const uint4*>(g + head);
    uint4* lv = reinterpret_cast<uint4*>(l + head);
    for (uint32_t v0 = 0; v0 < nv; v0 += U * nthr) {
        uint4 r[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) r[u] = gv[min(v0 + u * nthr + tid, nv - 1)];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) lv[min(v0 + u * nthr + tid, nv - 1)] = r[u];
    }
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
// Whole slices in LDS: one sweep.  `ns` consecutive slices of R x C starting
// at `base` are read once into LDS with an odd row pitch P (so a column walk
// touches distinct banks) and written back transposed.
template <typename T>
__device__ __forceinline__ void slice_body(T* __restrict__ base, uint32_t R, uint32_t C, uint32_t P,
                                           uint32_t ns, Div32 dS, Div32 dC, Div32 dR, T* lds) {
    const uint32_t S = R * C;
    const uint32_t SP = R * P;
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
}

// Vector form of slice_body for slices of a multiple of 16 bytes at a
// 16-byte aligned base: 16-byte loads (U per thread in flight; lanes past
// the end repeat the last vector, so the loads carry no branches), scattered
// into the odd-pitch LDS image with one division per vector, then 16-byte
// stores assembled from a column walk of the image.  Row r of the staged
// slices (all ns of them stacked) starts at LDS r * P.
template <typename T, int BLOCK>
__device__ __forceinline__ void slice_body_vec(T* __restrict__ base, uint32_t R, uint32_t C,
                                               uint32_t P, uint32_t ns, Div32 dS, Div32 dC,
                                               Div32 dR, T* lds) {
    constexpr uint32_t PER = 16 / sizeof(T);
    constexpr uint32_t U = 8;
    using V = union { uint4 v; T e[PER]; };
    const uint32_t S = R * C;
    const uint32_t nv = ns * S / PER;
    const uint4* src = reinterpret_cast<const uint4*>(base);
    for (uint32_t v0 = 0; v0 < nv; v0 += U * BLOCK) {
        V x[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) x[u].v = src[min(v0 + u * BLOCK + threadIdx.x, nv - 1)];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t e = min(v0 + u * BLOCK + threadIdx.x, nv - 1) * PER;
            uint32_t r = dC.div(e), j = e - r * C;  // stacked row, column
#pragma unroll
            for (uint32_t q = 0; q < PER; q++) {
                lds[r * P + j] = x[u].e[q];
                const bool wrap = ++j == C;
                j = wrap ? 0 : j;
                r += wrap ? 1 : 0;
            }
        }
    }
    __syncthreads();
    uint4* dst = reinterpret_cast<uint4*>(base);
    for (uint32_t v = threadIdx.x; v < nv; v += BLOCK) {
        const uint32_t o = v * PER;
        const uint32_t sl = dS.div(o), rem = o - sl * S;
        uint32_t jj = dR.div(rem), ii = rem - jj * R;  // output row jj (input column), col ii
        uint32_t row0 = sl * R;                         // stacked row of input row 0 of slice sl
        V y;
#pragma unroll
        for (uint32_t q = 0; q < PER; q++) {
            y.e[q] = lds[(row0 + ii) * P + jj];
            const bool wrap = ++ii == R;
            ii = wrap ? 0 : ii;
            const bool next = wrap && ++jj == C;
            jj = next ? 0 : jj;
            row0 += next ? R : 0;
        }
        dst[v] = y.v;
    }
}

// Each work item is `per` consecutive slices.  VEC: slices of a multiple of
// 16 bytes at a 16-byte aligned base (slice_body_vec).
template <typename T, bool VEC>
__global__ void __launch_bounds__(1024)
k_slice_lds(T* __restrict__ data, uint32_t R, uint32_t C, uint32_t P, uint64_t nslices,
            uint32_t per, Div32 dS, Div32 dC, Div32 dR) {
    IPT_DYN_LDS(T, lds);
    for (uint64_t s0 = (uint64_t)blockIdx.x * per; s0 < nslices; s0 += (uint64_t)gridDim.x * per) {
        const uint32_t ns = (uint32_t)umin64(per, nslices - s0);
        if (VEC) slice_body_vec<T, 256>(data + s0 * R * C, R, C, P, ns, dS, dC, dR, lds);
        else slice_body(data + s0 * R * C, R, C, P, ns, dS, dC, dR, lds);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Square slices N x N: swap-transpose tile pairs (I,J) / (J,I), I <= J, of an
// nt x nt tile grid.  Pair t in column order: t = J(J+1)/2 + I, so the
// workgroups running at one time take neighbouring tiles of a few tile
// columns, whose (I,J) tiles lie TS rows apart in one column.  In diagonal
// order (dord) the pairs with J - I = d are consecutive (d = 0, 1, ...; S(d) =
// d*nt - d(d-1)/2 pairs come before diagonal d), so both tiles of
// neighbouring pairs move along a row and a column at once.  Where the row
// pitch is a power of two of 16 KB or more, the column-order tiles' row
// segments fall on a few memory channels (f32, GB/s, column -> diagonal:
// 16384^2 4706 -> 4924, 8192^2 x 4 4755 -> 4985, 4096^2 x 16 4855 -> 4931);
// other pitches are 0.5-1.5 % slower in diagonal order (2048^2 x 64,
// 12288^2 x 2, 10240^2 x 3, 6144^2 x 8, 14336^2 x 2, 5120^2 x 12).
__device__ __forceinline__ void swap_pair(uint64_t t, uint64_t nt, bool dord, uint64_t* I, uint64_t* J) {
    if (!dord) {
        uint64_t j = (uint64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while (j * (j + 1) / 2 > t) j--;
        while ((j + 1) * (j + 2) / 2 <= t) j++;
        *I = t - j * (j + 1) / 2;
        *J = j;
        return;
    }
    const double b = 2.0 * (double)nt + 1.0;
    const double disc = b * b - 8.0 * (double)t;
    int64_t d = (int64_t)((b - sqrt(disc > 0.0 ? disc : 0.0)) * 0.5);
    auto S = [nt](int64_t e) { return (uint64_t)(e * (int64_t)nt - e * (e - 1) / 2); };
    if (d < 0) d = 0;
    while (d > 0 && S(d) > t) d--;
    while (S(d + 1) <= t) d++;
    *I = t - S(d);
    *J = *I + (uint64_t)d;
}

// One sweep; both tiles are staged in LDS (2 * TS * (TS+1) elements at lds)
// before either is written.  VEC: 16-byte global loads / stores (N*esz a
// multiple of 16 and a 16-byte aligned base); both tiles' loads are in flight
// before the first LDS store.
template <typename T, int TS, int BLOCK, bool VEC>
__device__ __forceinline__ void swap_body(T* __restrict__ base, uint32_t N, uint64_t t, T* lds, bool dord) {
    constexpr uint32_t TP = TS + 1;
    constexpr uint32_t PER = VEC ? 16 / sizeof(T) : 1;
    constexpr uint32_t VPR = TS / PER;           // vectors per tile row
    constexpr uint32_t VPT = TS * VPR;           // vectors per tile
    constexpr uint32_t UPT = (VPT + BLOCK - 1) / BLOCK;
    using V = typename std::conditional<VEC, uint4, T>::type;
    T* A = lds;
    T* B = lds + TS * TP;
    uint64_t I, J;
    swap_pair(t, (N + TS - 1) / TS, dord, &I, &J);
    const uint32_t r0 = (uint32_t)I * TS, c0 = (uint32_t)J * TS;
    const bool diag = I == J;
    V ra[UPT], rb[UPT];
#pragma unroll
    for (uint32_t u = 0; u < UPT; u++) {
        const uint32_t v = u * BLOCK + threadIdx.x;
        const uint32_t y = v / VPR, x = (v % VPR) * PER;
        if (v < VPT) {
            if (r0 + y < N && c0 + x < N) ra[u] = *reinterpret_cast<const V*>(base + (uint64_t)(r0 + y) * N + c0 + x);
            if (!diag && c0 + y < N && r0 + x < N)
                rb[u] = *reinterpret_cast<const V*>(base + (uint64_t)(c0 + y) * N + r0 + x);
        }
    }
#pragma unroll
    for (uint32_t u = 0; u < UPT; u++) {
        const uint32_t v = u * BLOCK + threadIdx.x;
        const uint32_t y = v / VPR, x = (v % VPR) * PER;
        if (v < VPT) {
            const T* ea = reinterpret_cast<const T*>(&ra[u]);
            const T* eb = reinterpret_cast<const T*>(&rb[u]);
#pragma unroll
            for (uint32_t q = 0; q < PER; q++) {
                A[y * TP + x + q] = ea[q];
                if (!diag) B[y * TP + x + q] = eb[q];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < UPT; u++) {
        const uint32_t v = u * BLOCK + threadIdx.x;
        const uint32_t y = v / VPR, x = (v % VPR) * PER;  // output row y, cols x..x+PER
        if (v < VPT) {
            // tile (I,J) <- transpose of tile (J,I) (of itself on the diagonal)
            union { V v; T e[PER]; } o;
            const T* src = diag ? A : B;
#pragma unroll
            for (uint32_t q = 0; q < PER; q++) o.e[q] = src[(x + q) * TP + y];
            if (r0 + y < N && c0 + x < N) *reinterpret_cast<V*>(base + (uint64_t)(r0 + y) * N + c0 + x) = o.v;
            if (!diag) {
#pragma unroll
                for (uint32_t q = 0; q < PER; q++) o.e[q] = A[(x + q) * TP + y];
                if (c0 + y < N && r0 + x < N)
                    *reinterpret_cast<V*>(base + (uint64_t)(c0 + y) * N + r0 + x) = o.v;
            }
        }
    }
}

constexpr int kSwapBlock = 256;

template <typename T, int TS, bool VEC>
__global__ void __launch_bounds__(kSwapBlock)
k_square_swap(T* __restrict__ data, uint32_t N, uint64_t pairs, uint64_t d3, bool dord) {
    __shared__ T tiles[2 * TS * (TS + 1)];
    const uint64_t total = pairs * d3;
    for (uint64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const uint64_t sl = w / pairs;
        swap_body<T, TS, kSwapBlock, VEC>(data + sl * (uint64_t)N * N, N, w - sl * pairs, tiles, dord);
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
        {
            // U rows per thread in flight; lanes past the strip's last column
            // or row repeat the last element (same address, same value).
            constexpr uint32_t U = 8;
            const uint32_t jc = min(jj, wn - 1);
            for (uint32_t i0 = threadIdx.x / W; i0 < R; i0 += U * rstep) {
                T r[U];
#pragma unroll
                for (uint32_t u = 0; u < U; u++) r[u] = base[(uint64_t)min(i0 + u * rstep, R - 1) * C + jc];
#pragma unroll
                for (uint32_t u = 0; u < U; u++) lds[min(i0 + u * rstep, R - 1) * W + jc] = r[u];
            }
        }
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

// Scratch of a chunk of n strips w in [first, first + n) is row-major across
// the chunk: scratch[i][w - first][jj] holds new[i][j] of strip w (element
// e = (i * n + w - first) * W + jj), so a scratch row is a run of n * W
// columns and the copy back writes each data row's run contiguously (with
// 256-byte strips strided by the row pitch the copy back ran at 3.3 TB/s).
// k_col_gather_to_scratch: scratch[i][w - first][jj] = old[g(i,j)][j].
// Element e of the chunk (total < 2^31, so 32-bit reduced division by the
// scratch row n * W): consecutive lanes walk one row's run of columns, so a
// map whose source row is constant across runs (a rotation by a per-run
// constant) reads contiguous bytes.  U elements per thread in flight; lanes
// past the end repeat the last element, and columns past C (a partial last
// strip) read column C - 1 into scratch slots the copy back never reads.
template <typename T, typename Ix, typename Map>
__global__ void __launch_bounds__(256)
k_col_gather_to_scratch(const T* __restrict__ data, T* __restrict__ scratch, Plan<Ix> p,
                        uint32_t wshift, Div64 dstrips, uint64_t first, uint32_t total,
                        Div32 drow) {
    constexpr uint32_t U = 8;
    const uint64_t C = p.C;
    const uint32_t W = 1u << wshift;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t e0 = blockIdx.x * blockDim.x + threadIdx.x; e0 < total; e0 += stride * U) {
        T r[U];
        uint32_t at[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t e = min(e0 + u * stride, total - 1);
            const uint32_t i = drow.div(e), rem = e - i * drow.d;
            const uint32_t wl = rem >> wshift, jj = rem & (W - 1);
            const uint64_t w = first + wl, sl = dstrips.div(w);
            const uint64_t j = umin64((w - sl * dstrips.d) * W + jj, C - 1);
            const T* base = data + sl * (uint64_t)p.R * C;
            at[u] = e;
            r[u] = base[(uint64_t)Map::src(p, (Ix)i, (Ix)j) * C + j];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) scratch[at[u]] = r[u];
    }
}

// The two diagonal column maps through scratch.  Both factor through p =
// r2c_prepermute (the C2R post row term h(i) = (i*C - i/a) mod R is p's
// inverse):
//   C2R post  new[i][j] = old[(h(i) + j) mod R][j]: for k = h(i), row p(k)
//             of the strip takes old[(k + j) mod R][j];
//   R2C pre   new[i][j] = old[p((i - j) mod R)][j]: source row p(k) goes to
//             rows (k + j) mod R.
// A unit is kb consecutive k (from k0) of one strip of W columns j0 + t.
// Its diagonal side is the band of rows (k0 + j0 + r) mod R, r < nk + W - 1,
// element (k, t) in band row (k - k0) + t, which the workgroup stages in LDS
// as whole W-wide row segments: C2R post loads the band and writes row p(k)
// of the strip for each k, R2C pre loads rows p(k) and writes the band.
// Both sides then move whole row segments (reading the diagonal directly
// from memory took one 4-byte access per row: f32 R = 28876 C2R post gather
// 1345 us, 1.6 TB/s per sweep).  Row segments are loaded as vectors V (16,
// 8 or 4 bytes: the widest the row pitch C * esz and the data's alignment
// allow), 32 KB in flight per workgroup.  The band's extra W - 1 rows are
// read twice (once by each neighbouring unit).  In the last strip, vectors
// past the row's end repeat its last vector: those lanes carry columns no
// output keeps (the copy back skips them).  512 threads, so that eight waves
// hide the LDS latency of the write side (unrolled by UW).  R * W < 2^31.
constexpr uint32_t kDiagBlock = 512;
template <typename T, typename V, typename Ix, bool C2R_POST>
__global__ void __launch_bounds__(kDiagBlock)
k_col_diag_to_scratch(const T* __restrict__ data, T* __restrict__ scratch, Plan<Ix> p,
                      uint32_t wshift, Div64 dstrips, uint64_t first, uint32_t n, uint32_t kb,
                      Div32 dnkb) {
    IPT_DYN_LDS(T, band);
    constexpr uint32_t E = sizeof(V) / sizeof(T);  // elements per vector
    constexpr uint32_t U = 4 * 16 / sizeof(V);     // vector loads per thread in flight
    constexpr uint32_t UW = 4;                     // band elements per thread in flight
    const uint32_t R = (uint32_t)p.R;
    const uint64_t C = p.C;
    const uint32_t W = 1u << wshift;
    const uint32_t VW = W / E;                   // vectors per band row
    const uint32_t vrows = kDiagBlock / VW;      // band rows one load pass covers
    const uint32_t q = threadIdx.x % VW, vr = threadIdx.x / VW;
    const uint32_t rows_per_pass = kDiagBlock >> wshift;  // element passes
    const uint32_t t = threadIdx.x & (W - 1), kr = threadIdx.x >> wshift;
    const uint64_t row_vecs = C / E;             // whole vectors per row (C * esz % sizeof(V) == 0)
    const uint32_t units = n * dnkb.d;
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const uint32_t wl = dnkb.div(u), kblk = u - wl * dnkb.d;
        const uint64_t w = first + wl, sl = dstrips.div(w);
        const uint64_t j0 = (w - sl * dstrips.d) * W;
        const V* vbase = reinterpret_cast<const V*>(data + sl * (uint64_t)R * C) +
                         umin64(j0 / E + q, row_vecs - 1);
        T* out = scratch + (uint64_t)wl * W + t;  // row i at out[i * n * W]
        const uint64_t pitch = (uint64_t)n * W;
        const uint32_t k0 = kblk * kb, nk = min(kb, R - k0);
        const uint32_t b0 = (uint32_t)p.dR.mod((Ix)(k0 + (uint32_t)(j0 % R)));
        const uint32_t nb = nk + W - 1;
        const uint32_t nload = C2R_POST ? nb : nk;  // band rows loaded
        for (uint32_t r0 = vr; r0 < nload; r0 += vrows * U) {
            V v[U];
            uint32_t at[U];
#pragma unroll
            for (uint32_t x = 0; x < U; x++) {
                const uint32_t r = min(r0 + x * vrows, nload - 1);  // past the end: the last again
                const uint32_t row = C2R_POST ? (uint32_t)p.dR.mod((Ix)(b0 + r))
                                              : (uint32_t)r2c_prepermute(p, (Ix)(k0 + r));
                at[x] = r * VW + q;
                v[x] = vbase[(uint64_t)row * row_vecs];
            }
#pragma unroll
            for (uint32_t x = 0; x < U; x++) reinterpret_cast<V*>(band)[at[x]] = v[x];
        }
        __syncthreads();
        if (C2R_POST) {
            for (uint32_t kk0 = kr; kk0 < nk; kk0 += rows_per_pass * UW) {
                T v[UW];
                uint64_t o[UW];
#pragma unroll
                for (uint32_t x = 0; x < UW; x++) {
                    const uint32_t kk = min(kk0 + x * rows_per_pass, nk - 1);  // past the end: the last again
                    o[x] = (uint64_t)r2c_prepermute(p, (Ix)(k0 + kk)) * pitch;
                    v[x] = band[(kk + t) * W + t];
                }
#pragma unroll
                for (uint32_t x = 0; x < UW; x++) out[o[x]] = v[x];
            }
        } else {
            for (uint32_t r0 = kr; r0 < nb; r0 += rows_per_pass * UW) {
                T v[UW];
                uint64_t o[UW];
                bool ok[UW];
#pragma unroll
                for (uint32_t x = 0; x < UW; x++) {
                    const uint32_t r = r0 + x * rows_per_pass;
                    ok[x] = r < nb && r >= t && r - t < nk;
                    const uint32_t rr = ok[x] ? r : t;  // a valid band row either way
                    o[x] = (uint64_t)p.dR.mod((Ix)(b0 + rr)) * pitch;
                    v[x] = band[(rr - t) * W + t];
                }
#pragma unroll
                for (uint32_t x = 0; x < UW; x++)
                    if (ok[x]) out[o[x]] = v[x];
            }
        }
        __syncthreads();  // the band is reused by the next unit
    }
}

// data[i][j] = scratch[i][w - first][jj] over the chunk (element e as in
// k_col_gather_to_scratch).
template <typename T>
__global__ void __launch_bounds__(256)
k_col_scratch_back(T* __restrict__ data, const T* __restrict__ scratch, uint64_t R, uint64_t C,
                   uint32_t wshift, Div64 dstrips, uint64_t first, uint32_t total, Div32 drow) {
    constexpr uint32_t U = 8;  // elements per thread in flight
    const uint32_t W = 1u << wshift;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t e0 = blockIdx.x * blockDim.x + threadIdx.x; e0 < total; e0 += stride * U) {
        T r[U];
        uint64_t dst[U];
        bool ok[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t e = min(e0 + u * stride, total - 1);
            const uint32_t i = drow.div(e), rem = e - i * drow.d;
            const uint32_t wl = rem >> wshift, jj = rem & (W - 1);
            const uint64_t w = first + wl, sl = dstrips.div(w);
            const uint64_t j = (w - sl * dstrips.d) * W + jj;
            ok[u] = j < C;
            dst[u] = sl * R * C + i * C + j;
            r[u] = scratch[e];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++)
            if (ok[u]) data[dst[u]] = r[u];
    }
}

// scratch[r][j] = old[g0+r][s_i(j)], i = (g0 + r) mod R, for rows [g0, g0+n).
template <typename T, typename Ix, typename Map>
__global__ void __launch_bounds__(256)
k_row_gather_to_scratch(const T* __restrict__ data, T* __restrict__ scratch, Plan<Ix> p,
                        uint64_t g0, uint64_t n) {
    constexpr uint32_t U = 4;  // elements per thread in flight; past the end: the last again
    const uint64_t C = p.C;
    const uint64_t total = n * C;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e0 < total; e0 += stride * U) {
        T r[U];
        uint64_t at[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint64_t e = umin64(e0 + u * stride, total - 1);
            const uint64_t rr = e / C, j = e - rr * C;
            const uint64_t g = g0 + rr;
            const Ix i = (Ix)(g % (uint64_t)p.R);
            at[u] = e;
            r[u] = data[g * C + (uint64_t)Map::src(p, i, (Ix)j)];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) scratch[at[u]] = r[u];
    }
}

}  // namespace ipt

namespace ipt {

// ---------------------------------------------------------------------------
// Group-cooperative staged transpose (GCS): one sweep for slices larger than
// one workgroup's LDS.  Slot s is member s % g of the group that owns slice
// s / g; workgroup b takes slots b, b + grid, b + 2*grid, ... (static
// round-robin, no global counter).  Each member stages its share of the
// slice's tiles (LDS plus registers), arrives at the slice's group barrier,
// waits for the members whose staged tiles its stores may overwrite (the
// whole group, or in the batch kernel the range gcs_deps gives), then writes
// its tiles transposed.  Every element is read before any write to it and
// slices never overlap, so no workgroup ever reads data another one wrote:
// the barrier words carry no data and relaxed agent-scope atomics suffice.
//
// Sync state (device memory, zero at rest; each launch restores it):
//   ctl[2] watchdog flag (sticky); ctl[6..7] the host watchdog word's address
//   cnt[bar_words(g) * S ...] the barrier state of slice S (see
//   group_arrive / group_wait); no two slices share a line.
// The spin is bounded: after ~2^26 polls the member records ctl[2] and
// proceeds, so a broken invariant shows up as an error, never as a hang.
// ctl[3] (a poll limit in place of 2^26) and ctl[4] (arrivals to wait for
// beyond the group) are zero in use; a test sets them, through
// ipt_debug_set_watchdog, to trip the watchdog on purpose.
//
// Deadlock freedom: the grid never exceeds the resident workgroups and
// g <= grid.  A member of slice S waits only for other members of S, and
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
__device__ __forceinline__ bool spin_until(uint32_t* ctr, uint32_t target, uint32_t limit) {
    uint32_t polls = 0;
    while (atomic_load_agent(ctr) < target) {
        __builtin_amdgcn_s_sleep(1);
        if (++polls > limit) return false;
    }
    return true;
}

// Wait until every vector-memory load this wave issued has returned its data.
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Group barrier of the g members of one slice, in one of two forms (state:
// bar_words(g) words, zero at rest; the last member to depart zeroes every
// word it used).
//
// Counters (single launches, where every member starts at once): up to
// kBarLines - 1 members, one arrival counter (line 0, word 0), each member
// adds 1 and waits for g.  More members: line 1 + q, word 0 counts the
// arrivals of sub-group q (members m with m % nsub == q, nsub = kBarLines -
// 1); the last arriver of a sub-group bumps line 0 word 0 and members wait
// for it to reach nsub.  Line 0 word 16 counts departures.
//
// Flags (the persistent batch kernel, whose members start whenever a
// workgroup frees up): word 32 + m is member m's arrival flag, and a member
// waits only for the members whose staged tiles its stores may overwrite
// (gcs_deps: a contiguous range lo..hi), one wave polling up to 64 flags per
// load; word 16 counts departures.  A member no longer waits for the last
// of its whole group to be taken from the queue: uneven 4096-matrix batch,
// k_batch 3.71 -> 3.36 ms.  In single launches, where all members wait at
// once, the polling of many flags costs more than the wait it saves (f32
// 3000 x 1000, 256 members: 1457 -> 1105 GB/s; 4099 x 4093 4142 -> 3716).
constexpr uint32_t kBarLines = 17;

// Words of barrier state per slice for a group of g (either form).
IPT_HD uint64_t bar_words(uint32_t g) {
    const uint64_t cnt = 32ull * (1 + (g < kBarLines - 1 ? g : kBarLines - 1));
    const uint64_t flg = 32ull + ((uint64_t)g + 31) / 32 * 32;
    return cnt > flg ? cnt : flg;
}

// Counters: arrive, then wait until every member has arrived (thread 0;
// false on watchdog).  A member may do work of its own in between, as long
// as it needs nothing from the others.
__device__ __forceinline__ void group_arrive(uint32_t* bar, uint32_t m, uint32_t g) {
    if (g <= kBarLines - 1) {
        atomic_add_agent(&bar[0], 1u);
        return;
    }
    const uint32_t nsub = kBarLines - 1;
    const uint32_t q = m % nsub;
    const uint32_t members = g / nsub + (q < g % nsub ? 1u : 0u);
    if (atomic_add_agent(&bar[32 * (1 + q)], 1u) == members - 1) atomic_add_agent(&bar[0], 1u);
}

__device__ __forceinline__ bool group_wait(uint32_t* bar, uint32_t g, uint32_t extra, uint32_t limit) {
    return spin_until(&bar[0], (g <= kBarLines - 1 ? g : kBarLines - 1) + extra, limit);
}

__device__ __forceinline__ void group_depart(uint32_t* bar, uint32_t g) {
    if (atomic_add_agent(&bar[16], 1u) == g - 1) {
        if (g > kBarLines - 1)
            for (uint32_t q = 0; q < kBarLines - 1; q++) atomic_store_agent(&bar[32 * (1 + q)], 0u);
        atomic_store_agent(&bar[0], 0u);
        atomic_store_agent(&bar[16], 0u);
    }
}

// Flags: arrive (thread 0), then wait (wave 0, every lane) until members
// lo..hi have arrived (false on watchdog).
__device__ __forceinline__ void flag_arrive(uint32_t* bar, uint32_t m) { atomic_store_agent(&bar[32 + m], 1u); }

__device__ __forceinline__ bool flag_wait(uint32_t* bar, uint32_t lo, uint32_t hi, uint32_t extra,
                                          uint32_t limit, uint32_t lane) {
    uint32_t polls = 0;
    for (uint32_t b = lo; b <= hi; b += 64) {
        const uint32_t i = b + lane;
        bool seen = i > hi;
        for (;;) {
            if (!seen) seen = atomic_load_agent(&bar[32 + i]) != 0;
            if (__all(seen)) break;
            __builtin_amdgcn_s_sleep(1);
            if (++polls > limit) return false;
        }
    }
    if (extra) {  // test hook: wait for arrivals that never come
        while (++polls <= limit) __builtin_amdgcn_s_sleep(1);
        return false;
    }
    return true;
}

__device__ __forceinline__ void flag_depart(uint32_t* bar, uint32_t g) {
    if (atomic_add_agent(&bar[16], 1u) == g - 1) {
        for (uint32_t m = 0; m < g; m++) atomic_store_agent(&bar[32 + m], 0u);
        atomic_store_agent(&bar[16], 0u);
    }
}

// Staging shared by the members of one slice: after every wave has its
// loads back, thread 0 arrives, and before the first store it waits for
// the group (group_arrive / group_wait); `bar` is null when the slice has
// one member.
// The watchdog settings ctl[3], ctl[4], read by thread 0 before the member's
// loads are issued (so the read overlaps them instead of delaying the
// arrival; nothing writes them while a kernel runs).
struct Watchdog {
    uint32_t limit = kSpinLimit, extra = 0;
    __device__ __forceinline__ void read(const uint32_t* ctl, uint32_t g) {
        if (g > 1 && threadIdx.x < 64) {
            const uint32_t l = ctl[3];
            limit = l ? l : kSpinLimit;
            extra = ctl[4];
        }
    }
};

// A barrier gave up: set the sticky device flag ctl[2] and, when ctl[6..7]
// (low, high half; written once when the control block is set up, never
// while a kernel runs) hold the address of the owner's host-mapped watchdog
// word, set that too with a system-scope store, so the host sees it on its
// next call without synchronising.
__device__ __forceinline__ void report_watchdog(uint32_t* ctl) {
    atomic_store_agent(&ctl[2], 1u);
    const uint64_t a = ((uint64_t)ctl[7] << 32) | ctl[6];
    if (a) __hip_atomic_store(reinterpret_cast<uint32_t*>(a), 1u, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---------------------------------------------------------------------------
// Wave-tile path (any row length and element alignment, slices below 2^31
// bytes).
//
// Global accesses go through a raw buffer resource over the slice: a load at
// an offset outside it returns zero without touching memory and a store
// outside it is dropped.  Lanes with nothing to move use kOutside, so the
// code has no branches around its memory operations and every load of a
// wave is in flight at once.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __amdgpu_buffer_rsrc_t Rsrc;

constexpr uint32_t kOutside = 0x80000000u;  // past every resource built here (< 2^31 bytes)

// Cache policy of the transposing stores: sc1 (device scope) writes through
// the XCD's L2, so a launch ends with no dirty lines for the end-of-kernel
// write-back and the next launch starts about 4 us sooner (measured on the
// [512][384][64] f32 permute: 24.4 -> 23.0 us per back-to-back launch).
// Launches over larger slices store write-back instead (policy 0), which
// lets the L2 merge the partial lines that neighbouring tiles share.
constexpr int kStoreDevice = 0x10;

// Raw buffer resource over [base, base + bytes); word 3 = 0x00020000 is the
// gfx9 raw-buffer format (32-bit data, no swizzle, range-checked).
__device__ __forceinline__ Rsrc make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             0x00020000);
}

// Wave tile: one wavefront moves a TS x TS tile, TS = 8 * PER with PER =
// 16 / sizeof(T) (32 x 32 for 4-byte, 16 x 16 for 8-byte elements), as 64
// lane micro-tiles of PER x PER.  Lane (a, b) = (lane / 8, lane % 8) loads
// rows PER*a + k (k < PER) at columns PER*b .. PER*b + PER - 1, so eight
// lanes cover one 128-byte row segment; it transposes its micro-tile in
// registers and stores output rows PER*b + q at columns PER*a .., eight lanes
// again covering one 128-byte segment.
// Any row length works: 16-byte buffer accesses need only 4-byte alignment
// on gfx950, and the range check is per dword (a load past the end of the
// slice returns zeros there, a store there is dropped).  A micro-row that
// runs past the end of its matrix row therefore loads harmless extra values
// (the next row's, or zeros), which land in output rows >= cols or output
// columns >= rows.  Output rows >= cols are not stored; an output micro-row
// that runs past `rows` (only in tiles of the last row band) is stored
// element by element, up to `rows`.
// The first n (1 .. PER - 1) elements of the 16-byte vector x to byte
// offset o (lanes with live false store nothing): 4-byte elements as one
// 8-byte store of elements 0-1 when n >= 2 and one 4-byte store of element
// n - 1 when n is odd, 8-byte elements as one 8-byte store.
template <typename T, int SP>
__device__ __forceinline__ void store_head(u32x4 x, Rsrc rs, uint32_t o, uint32_t n, bool live) {
    __builtin_amdgcn_raw_buffer_store_b64(u32x2{x[0], x[1]}, rs,
                                          live && (sizeof(T) == 8 || n >= 2) ? o : kOutside, 0, SP);
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(n == 1 ? x[0] : x[2], rs, live && (n & 1) ? o + (n - 1) * 4 : kOutside,
                                              0, SP);
}

template <typename T>
struct WaveTile {
    static constexpr uint32_t PER = 16 / sizeof(T);
    static constexpr uint32_t TS = 8 * PER;
    u32x4 v[PER];

    // Tile with top-left (r0, c0) of a rows x cols row-major slice; `live`
    // false loads nothing (zeros).
    __device__ __forceinline__ void load(Rsrc rs, uint32_t r0, uint32_t c0, uint32_t rows,
                                         uint32_t cols, uint32_t lane, bool live) {
        const uint32_t r = r0 + PER * (lane >> 3), c = c0 + PER * (lane & 7);
        const bool ok = live && r < rows && c < cols;
        const uint32_t o = (r * cols + c) * (uint32_t)sizeof(T);
#pragma unroll
        for (uint32_t k = 0; k < PER; k++)
            v[k] = __builtin_amdgcn_raw_buffer_load_b128(
                rs, ok ? o + k * cols * (uint32_t)sizeof(T) : kOutside, 0, 0);
    }
    // Line-aligned form of load(): each lane loads 16-byte chunk lane % 8 of
    // the 128-byte line that holds the start of its row's segment, through
    // `rl`, a resource over the slice from the line below its base (`mis` =
    // the base's offset in that line, bytes).  realign() then moves the bytes
    // into place, given the lines of the right-hand neighbour tile.
    __device__ __forceinline__ void load_lines(Rsrc rl, uint32_t mis, uint32_t r0, uint32_t c0,
                                               uint32_t rows, uint32_t cols, uint32_t lane, bool live) {
        const uint32_t r = r0 + PER * (lane >> 3), chunk = 16 * (lane & 7);
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            const uint32_t a = mis + ((r + k) * cols + c0) * (uint32_t)sizeof(T);
            v[k] = __builtin_amdgcn_raw_buffer_load_b128(
                rl, live && r + k < rows ? (a & ~127u) + chunk : kOutside, 0, 0);
        }
    }
    // x rotated down by t dwords (t uniform): dword q of the result is
    // x[(q + t) mod 4].
    static __device__ __forceinline__ u32x4 rotate_dwords(u32x4 x, uint32_t t) {
        if (t & 1) x = u32x4{x[1], x[2], x[3], x[0]};
        if (t & 2) x = u32x4{x[2], x[3], x[0], x[1]};
        return x;
    }
    // After load_lines() of this tile and of `nx` (the tile at (r0, c0 + TS),
    // whose lines are one line further along each row), put into v[] what
    // load() would have loaded: row bytes [d, d + 128) of the two lines, d
    // the segment's offset in its line (values past the row end or the
    // slice end differ from load()'s; such positions are never stored).
    // Lane (a, b) wants chunk b + d / 16 of the 16 (this tile's 8, then
    // nx's), shifted by (d / 4) mod 4 dwords.  That dword shift is the same
    // for every lane of one load (its rows are PER apart and PER * cols * E
    // is a multiple of 16), so only the source lane varies: one ds_bpermute
    // per dword, from lanes that offer nx's dword when their chunk index is
    // below d / 16 plus the carry into the next chunk.
    __device__ __forceinline__ void realign(const WaveTile& nx, uint32_t mis, uint32_t r0, uint32_t c0,
                                            uint32_t cols, uint32_t lane) {
        constexpr uint32_t E = sizeof(T);
        const uint32_t a = lane >> 3, b = lane & 7;
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            const uint32_t d = (mis + ((r0 + PER * a + k) * cols + c0) * E) & 127u;
            const uint32_t s = d >> 4;
            const uint32_t t = __builtin_amdgcn_readfirstlane((d >> 2) & 3u);
            const u32x4 cur = rotate_dwords(v[k], t), nxt = rotate_dwords(nx.v[k], t);
#pragma unroll
            for (uint32_t q = 0; q < 4; q++) {
                const uint32_t e = q + t >= 4 ? 1u : 0u;  // carried into the next chunk
                const uint32_t src = (lane & ~7u) | ((b + s + e) & 7u);
                const uint32_t y = b < s + e ? nxt[q] : cur[q];
                v[k][q] = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)y);
            }
        }
    }
    __device__ __forceinline__ void transpose() {
        union U { u32x4 v; T e[PER]; };
        U in[PER], out[PER];
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) in[k].v = v[k];
#pragma unroll
        for (uint32_t q = 0; q < PER; q++)
#pragma unroll
            for (uint32_t k = 0; k < PER; k++) out[q].e[k] = in[k].e[q];
#pragma unroll
        for (uint32_t q = 0; q < PER; q++) v[q] = out[q].v;
    }
    // After transpose(): the tile loaded at (r0, c0) of a rows x cols slice
    // goes to output rows c0 + .., columns r0 + .. (the output has `rows`
    // columns); `live` false stores nothing.  SP: the stores' cache policy.
    template <int SP>
    __device__ __forceinline__ void store_t(Rsrc rs, uint32_t r0, uint32_t c0, uint32_t rows,
                                            uint32_t cols, uint32_t lane, bool live) const {
        constexpr uint32_t E = sizeof(T);
        const uint32_t orow = c0 + PER * (lane & 7), ocol = r0 + PER * (lane >> 3);
        const bool ok = live && ocol < rows;
        const bool full = ocol + PER <= rows;
        const uint32_t o = (orow * rows + ocol) * E;
#pragma unroll
        for (uint32_t q = 0; q < PER; q++)
            __builtin_amdgcn_raw_buffer_store_b128(
                v[q], rs, ok && full && orow + q < cols ? o + q * rows * E : kOutside, 0, SP);
        if (r0 + TS > rows && __any(ok && !full)) {  // last row band, some lane at a partial row end
#pragma unroll
            for (uint32_t q = 0; q < PER; q++)
                store_head<T, SP>(v[q], rs, o + q * rows * E, rows - ocol, ok && !full && orow + q < cols);
        }
    }
};

// Square swap through a buffer resource (slices below 2^31 bytes, any side
// and element-aligned base): the same tile-pair exchange as swap_body with
// branch-free 16-byte loads (lanes outside the slice load nothing) and
// device-scope stores.  Measured on a 16384^2 f32 matrix (HBM): 4705 -> 4819
// GB/s.  Sides that are not a multiple of 16 bytes work as in WaveTile:
// buffer accesses need only element alignment on gfx950 and are range
// checked per dword, so a micro-row that runs past its matrix row loads
// harmless values (they land in positions past the row, never stored), and
// in the last tile column the stores of such a micro-row go element by
// element up to the row end.  (Only the last tile column has them: I <= J,
// so a pair whose row band is the last one is a diagonal pair.)
// LP, SP: the loads' and stores' cache policies.
template <typename T, int TS, int BLOCK, int LP = 0, int SP = kStoreDevice>
__device__ __forceinline__ void swap_body_rs(T* __restrict__ base, uint32_t N, uint64_t t, T* lds, bool dord) {
    constexpr uint32_t TP = TS + 1, PER = 16 / sizeof(T), VPR = TS / PER, VPT = TS * VPR;
    constexpr uint32_t UPT = (VPT + BLOCK - 1) / BLOCK;
    constexpr uint32_t E = sizeof(T);
    const Rsrc rs = make_rsrc(base, N * N * E);
    T* A = lds;
    T* B = lds + TS * TP;
    uint64_t I, J;
    swap_pair(t, (N + TS - 1) / TS, dord, &I, &J);
    const uint32_t r0 = (uint32_t)I * TS, c0 = (uint32_t)J * TS;
    const bool diag = I == J;
    u32x4 ra[UPT], rb[UPT];
#pragma unroll
    for (uint32_t u = 0; u < UPT; u++) {
        const uint32_t v = u * BLOCK + threadIdx.x, y = v / VPR, x = (v % VPR) * PER;
        const bool oka = v < VPT && r0 + y < N && c0 + x < N;
        const bool okb = v < VPT && !diag && c0 + y < N && r0 + x < N;
        ra[u] = __builtin_amdgcn_raw_buffer_load_b128(rs, oka ? ((r0 + y) * N + c0 + x) * E : kOutside, 0, LP);
        rb[u] = __builtin_amdgcn_raw_buffer_load_b128(rs, okb ? ((c0 + y) * N + r0 + x) * E : kOutside, 0, LP);
    }
#pragma unroll
    for (uint32_t u = 0; u < UPT; u++) {
        const uint32_t v = u * BLOCK + threadIdx.x, y = v / VPR, x = (v % VPR) * PER;
        if (v < VPT) {
            union { u32x4 v; T e[PER]; } a, b;
            a.v = ra[u];
            b.v = rb[u];
#pragma unroll
            for (uint32_t q = 0; q < PER; q++) {
                A[y * TP + x + q] = a.e[q];
                B[y * TP + x + q] = b.e[q];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < UPT; u++) {
        const uint32_t v = u * BLOCK + threadIdx.x, y = v / VPR, x = (v % VPR) * PER;
        if (v < VPT) {
            // tile (I,J) <- transpose of tile (J,I) (of itself on the diagonal)
            union { u32x4 v; T e[PER]; } o, p;
            const T* src = diag ? A : B;
#pragma unroll
            for (uint32_t q = 0; q < PER; q++) {
                o.e[q] = src[(x + q) * TP + y];
                p.e[q] = A[(x + q) * TP + y];
            }
            const bool oka = r0 + y < N && c0 + x < N;
            const bool okb = !diag && c0 + y < N && r0 + x < N;
            const bool fa = c0 + x + PER <= N, fb = r0 + x + PER <= N;  // whole micro-row inside
            const uint32_t offa = ((r0 + y) * N + c0 + x) * E, offb = ((c0 + y) * N + r0 + x) * E;
            __builtin_amdgcn_raw_buffer_store_b128(o.v, rs, oka && fa ? offa : kOutside, 0, SP);
            __builtin_amdgcn_raw_buffer_store_b128(p.v, rs, okb && fb ? offb : kOutside, 0, SP);
            if (c0 + TS > N) {  // the last tile column: partial micro-rows (uniform per pair)
#pragma unroll
                for (uint32_t q = 0; q < PER; q++) {
                    const bool wa = oka && !fa && c0 + x + q < N;
                    const bool wb = okb && !fb && r0 + x + q < N;
                    if constexpr (sizeof(T) == 4) {
                        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)o.e[q], rs, wa ? offa + q * E : kOutside,
                                                              0, SP);
                        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)p.e[q], rs, wb ? offb + q * E : kOutside,
                                                              0, SP);
                    } else {
                        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, (uint64_t)o.e[q]), rs,
                                                              wa ? offa + q * E : kOutside, 0, SP);
                        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, (uint64_t)p.e[q]), rs,
                                                              wb ? offb + q * E : kOutside, 0, SP);
                    }
                }
            }
        }
    }
}

template <typename T, int TS>
__global__ void __launch_bounds__(kSwapBlock)
k_square_swap_rs(T* __restrict__ data, uint32_t N, uint64_t pairs, uint64_t d3, bool dord) {
    __shared__ T tiles[2 * TS * (TS + 1)];
    const uint64_t total = pairs * d3;
    for (uint64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const uint64_t sl = w / pairs;
        swap_body_rs<T, TS, kSwapBlock>(data + sl * (uint64_t)N * N, N, w - sl * pairs, tiles, dord);
        __syncthreads();
    }
}

// Large tiles for squares whose row pitch is a power of two of 64 KB or more
// (512-byte row segments, 1024 threads, one workgroup and its 2 * TS * (TS +
// 1) elements of LDS per CU), pairs in diagonal order, non-temporal loads and
// stores.
constexpr int kSwapBigBlock = 1024;
template <typename T, int TS>
__global__ void __launch_bounds__(kSwapBigBlock)
k_square_swap_big(T* __restrict__ data, uint32_t N, uint64_t pairs, uint64_t d3) {
    IPT_DYN_LDS(T, tiles);
    const uint64_t total = pairs * d3;
    for (uint64_t w = blockIdx.x; w < total; w += gridDim.x) {
        const uint64_t sl = w / pairs;
        swap_body_rs<T, TS, kSwapBigBlock, 2, 2>(data + sl * (uint64_t)N * N, N, w - sl * pairs, tiles, true);
        __syncthreads();
    }
}

// Slot numbering for the cooperative kernels.  Workgroups are dispatched
// round-robin over the device's XCDs (workgroup i on XCD i % nxcd), and each
// XCD has its own L2.  A member that writes lines another member read just
// before is about twice as slow when that member ran on another XCD
// (measured on the [512][384][64] f32 permute: 24.1 -> 16.5 us per launch
// with XCD-local groups), so slots are numbered to give each XCD one
// contiguous range.  When a group fits an XCD's share (h = (n / nxcd) / g
// >= 1 groups per XCD), the range holds whole groups: workgroup i < h * g *
// nxcd takes slot x * h * g + i / nxcd (x = i % nxcd), so each of those
// groups runs on one XCD and shares its L2, and the remaining workgroups
// keep slot i (a group among them spans XCDs).  Groups of g that do not
// divide an XCD's share then no longer straddle two XCDs (f64 [256][1000]
// [17]: 15 members per slice, 255 workgroups; 16 of the 17 groups on one
// XCD instead of 10).  Larger groups take consecutive ranges of XCDs (each
// XCD one contiguous range of n / nxcd slots).  A bijection on [0, n): the
// set of slots run, and the deadlock-freedom argument (which is over slot
// order), do not depend on it.
__device__ __forceinline__ uint32_t xcd_slot(uint32_t i, uint32_t n, uint32_t nxcd, uint32_t g) {
    if (nxcd <= 1) return i;
    const uint32_t x = i % nxcd;
    const uint32_t span = (n / nxcd / g) * g;  // slots per XCD in whole groups
    if (span == 0) {
        const uint32_t q = n / nxcd, rem = n % nxcd;
        return x * q + min(x, rem) + i / nxcd;
    }
    if (i >= span * nxcd) return i;
    return x * span + i / nxcd;
}

// Wave tiles a wave stages in LDS (JL) and holds in registers (JR) on the
// cooperative kernel, per element size: 40 / 80 LDS tiles fill 160 KB, and the
// register tiles hold up to 144 KB per 512-thread workgroup.
constexpr int gcs_wave_lds_tiles(int esz) { return esz == 4 ? 5 : 10; }
constexpr int kBatchJrClass = 1;
// Register tiles per wave are a template argument: class 0, 1, 2 or 3 holds
// 1, 4, 8 or 9 tiles (f32) and 1, 8, 16 or 18 (f64): the smallest class that
// holds a member's tiles beyond the LDS ones (unused register tiles still
// cost load / store instructions and registers).  Class 3 (256 / 250
// VGPRs, no spills) only raises the capacity for slices that class 2
// cannot hold in one group of 256 members: up to 116 MB instead of 109 /
// 106 MB (f32 5400 x 5300: 2342 GB/s on the block-cycle path, 5080 in one
// sweep; f64 3751 x 3723: 2239 -> 4927); as the class of every slice past
// class 1 it cost up to 3.6 % (f32 2326 x 4023 x 28).  All fit the two waves
// per SIMD that LDS allows; the batch kernel, which also holds the other
// item kinds, stops at class 1.
constexpr int gcs_jr_of(int esz, int cls) {
    return cls == 0 ? 1 : cls == 1 ? (esz == 4 ? 4 : 8) : cls == 2 ? (esz == 4 ? 8 : 16) : (esz == 4 ? 9 : 18);
}
// tail: classes 0 and 1 keep one register tile per wave free (gcs_geom).
inline int gcs_jr_class(int esz, uint32_t K, uint32_t KL, uint32_t nw, bool tail = false) {
    const uint32_t extra = K - KL, t = tail ? 1u : 0u;
    if (extra <= ((uint32_t)gcs_jr_of(esz, 0) - t) * nw) return 0;
    if (extra <= ((uint32_t)gcs_jr_of(esz, 1) - t) * nw) return 1;
    if (extra <= (uint32_t)gcs_jr_of(esz, 2) * nw) return 2;
    return 3;
}

// Which tiles a member of a cooperative group owns.  Strided (bh == 0): member
// m takes tiles m, m + g, m + 2g, ... (k < nk).  Banded (bh > 0): tiles are
// walked band by band, a band being bh tile rows (the last band may be
// shorter) walked column by column: position p = b * bh * ntj + tj * h + r is
// tile (b * bh + r, tj), h the band's height; member m owns positions
// m * K .. m * K + nk - 1, a block of adjacent tiles about bh tall.
// Adjacent tiles share the 128-byte lines that a row not aligned to 128
// bytes splits, on the input side (along a tile row) and on the output side
// (along a tile column); in a block those stay with one member, instead of
// being fetched and partly written by several (on several XCDs when the
// group spans them).  Both orders are bijections onto the tiles.
struct TileOrder {
    uint32_t bh = 0;    // band height in tiles; 0: strided
    uint32_t full = 0;  // bands of height bh
    Div32 dband{};      // bh * ntj: positions per full band
    Div32 dbh{};        // bh
    Div32 dlast{};      // the last, shorter band's height (unused when none)
};

// Banded order over a slice of nti x ntj tiles for members of nw waves:
// bands nw tiles tall.  Wave w of a member takes its tiles k = w + j * nw
// (gcs_body), which in a band of nw rows are one tile row, consecutive
// columns: the line two horizontally adjacent tiles share is then requested
// by back-to-back loads of one wave and fetched once, where loads from
// different waves fetch it twice (read probe on misaligned rows: 1.03-1.06x
// vs 1.12-1.22x the bytes).  Host side.
inline TileOrder banded_order(uint32_t nw, uint32_t nti, uint32_t ntj) {
    TileOrder o;
    const uint32_t bh = nw;
    o.bh = bh < nti ? bh : nti;
    o.full = nti / o.bh;
    o.dband = Div32::make(o.bh * ntj);
    o.dbh = Div32::make(o.bh);
    o.dlast = Div32::make(nti - o.full * o.bh);
    return o;
}

// The members whose staged tiles member m's stores may overwrite, lo..hi
// (empty when m owns no tile).  Strided order: the whole group.  Banded
// order: m's tiles lie in tile columns tja..tjb (those of one band; every
// column when its positions run into the next band), so it writes only
// output rows tja*TS .. (tjb+1)*TS - 1 of R elements each, slice elements
// e0 .. e1 - 1, which hold input rows e0 / C .. (e1 - 1) / C; every member
// owning a position in those rows' bands is waited for.
template <int TS>
IPT_HD void gcs_deps(uint32_t m, uint32_t nk, uint32_t K, uint32_t ntiles, uint32_t g, uint32_t R,
                     uint32_t C, uint32_t ntj, const TileOrder& ord, uint32_t* lo, uint32_t* hi) {
    if (ord.bh == 0) {
        *lo = 0;
        *hi = g - 1;
        return;
    }
    if (nk == 0) {
        *lo = 1;
        *hi = 0;
        return;
    }
    const uint32_t bandpos = ord.dband.d;  // bh * ntj positions per full band
    auto band = [&](uint32_t p) {
        const uint32_t b = ord.dband.div(p);
        return b < ord.full ? b : ord.full;
    };
    const uint32_t p0 = m * K, p1 = m * K + nk - 1;
    const uint32_t b0 = band(p0), b1 = band(p1);
    uint32_t tja = 0, tjb = ntj - 1;
    if (b0 == b1) {
        const Div32& dh = b0 < ord.full ? ord.dbh : ord.dlast;
        tja = dh.div(p0 - b0 * bandpos);
        tjb = dh.div(p1 - b0 * bandpos);
    }
    const uint32_t rows_end = (tjb + 1) * (uint32_t)TS < C ? (tjb + 1) * (uint32_t)TS : C;
    const uint64_t e0 = (uint64_t)tja * TS * R, e1 = (uint64_t)rows_end * R;
    const uint32_t tr0 = (uint32_t)(e0 / C) / TS, tr1 = (uint32_t)((e1 - 1) / C) / TS;
    const uint32_t B0 = band(tr0 * ntj), B1 = band(tr1 * ntj);  // the bands of tile rows tr0, tr1
    const uint32_t q1 = B1 < ord.full ? (B1 + 1) * bandpos : ntiles;
    *lo = B0 * bandpos / K;
    *hi = (q1 - 1) / K < g - 1 ? (q1 - 1) / K : g - 1;
}

// One GCS slot: member m of the group of g that transposes the
// slice at `base`.  Member m owns tiles t = m + k*g, k < nk; wave w of the
// workgroup takes k = w, w + NW, ...  Tiles k < kl are staged in LDS
// (already transposed, lane-linear: conflict-free 16-byte accesses, and
// each wave reads back only what it wrote); tiles kl <= k < nk stay in
// registers across the group barrier.
//
// Rows whose pitch (or base) is off the 128-byte line grid split every row
// segment of a tile over two lines; loading them that way costs 14-16 % of
// the batch kernel's time on such shapes.  With AL (the batch kernel), a
// wave whose tiles are one run of neighbours along a tile row (banded
// order, inside one full band) loads whole lines instead
// (WaveTile::load_lines, plus the line after its last tile into a register
// tile the planner leaves free: gcs_geom's `tail`) and moves the bytes into
// place in registers (WaveTile::realign) after its member has arrived at
// the group barrier, so the shuffling overlaps the wait for the other
// members.  (Single launches keep plain loads: every member loads at once
// there and the wait is short.)
struct NoHook {
    __device__ void operator()() const {}
};
// `released`: called by every thread once the group barrier has released
// the member, before its stores.
template <typename T, int BLOCK, int JR, int SP, bool AL = true, bool DEP = true, typename HOOK = NoHook>
__device__ __forceinline__ void gcs_body(T* __restrict__ base, uint32_t R, uint32_t C, uint32_t g,
                                          uint32_t K, uint32_t KL, Div32 dntj, uint32_t ntiles,
                                          uint32_t m, uint32_t* bar, uint32_t* ctl, u32x4* lds,
                                          const TileOrder& ord = TileOrder{}, HOOK released = HOOK{}) {
    using W = WaveTile<T>;
    constexpr uint32_t NW = BLOCK / 64;
    constexpr int JL = gcs_wave_lds_tiles(sizeof(T));
    constexpr uint32_t TV = 64 * W::PER;  // 16-byte vectors per staged tile
    // LDS unit of 16-byte chunk c of output row ro of staged tile k: rows 8
    // units (128 bytes) apart, the chunk XORed with the row's PER-row group
    // and the row's low bit flipped by the parity of c + k.  A transposing
    // write (lanes 8b + a: row PER * a + q, chunk b) and a run read (lanes
    // 8r + c: tile r, chunk c) then put any 8 consecutive lanes on distinct
    // 16-byte slots of 128 bytes and any 16 on distinct slots of 256 bytes;
    // unswizzled, the eight lanes of a write shared one slot (k_batch: 2/3 of
    // its LDS cycles were bank conflicts).
    auto unit = [](uint32_t k, uint32_t ro, uint32_t c) {
        return k * TV + (ro ^ ((c ^ k) & 1u)) * 8 + (c ^ ((ro / W::PER) & 7u));
    };
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nk = ord.bh == 0 ? (m < ntiles ? min(K, (ntiles - m + g - 1) / g) : 0)
                                    : (m * K < ntiles ? min(K, ntiles - m * K) : 0);
    const uint32_t kl = min(nk, KL);
    const Rsrc rs = make_rsrc(base, R * C * (uint32_t)sizeof(T));
    Watchdog wd;
    wd.read(ctl, g);
    auto origin = [&](uint32_t k, uint32_t* r0, uint32_t* c0) {
        uint32_t ti, tj;
        if (ord.bh == 0) {
            const uint32_t t = m + k * g;
            ti = dntj.div(t);
            tj = t - ti * dntj.d;
        } else {
            const uint32_t p = m * K + k;
            const uint32_t b = ord.dband.div(p), rem = p - b * ord.dband.d;
            const Div32& dh = b < ord.full ? ord.dbh : ord.dlast;  // the band's height
            tj = dh.div(rem);
            ti = b * ord.dbh.d + (rem - tj * dh.d);
        }
        *r0 = ti * W::TS;
        *c0 = tj * W::TS;
    };
    // Line-aligned loads (AL; `sw`, the same for the whole member): rows off
    // the 128-byte line grid in the banded order.  Then wave w's tiles are
    // k = w + NW * i, the first jl of them (k < kl) staged in LDS and the
    // rest in register tiles j = i - jl, so each wave's tiles are one
    // sequence (a wave still holds at most ceil((nk - kl) / NW) <= JR
    // register tiles); otherwise register tile j of wave w is k = kl + w +
    // NW * j.  A wave goes line-aligned (al) when its live tiles i < n lie in
    // one full band, so they are the neighbours (ar0, ac0 + i * TS), and a
    // register tile is left after them (n - jl < JR): it takes the line
    // after the run, the tail.
    constexpr uint32_t E = sizeof(T);
    const uint32_t mis = (uint32_t)reinterpret_cast<uintptr_t>(base) & 127u;
    const bool sw = AL && ord.bh == NW && (mis != 0 || ((C * E) & 127u) != 0) &&
                    (uint64_t)R * C * E + 256 < ((uint64_t)1 << 31);
    const uint32_t jl = kl > wave ? (kl - wave + NW - 1) / NW : 0u;
    auto rk = [&](int j) { return sw ? wave + NW * (jl + (uint32_t)j) : kl + wave + NW * (uint32_t)j; };
    uint32_t n = 0, ar0 = 0, ac0 = 0;
    bool al = false;
    if constexpr (AL) {
        n = wave < nk ? (nk - wave + NW - 1) / NW : 0u;
        if (sw && n > 0 && n < jl + JR) {
            const uint32_t p0 = m * K + wave;
            const uint32_t b0 = ord.dband.div(p0), b1 = ord.dband.div(p0 + NW * (n - 1));
            al = b0 == b1 && b0 < ord.full;
            if (al) origin(wave, &ar0, &ac0);
        }
    }
    W rt[JR], lt[JL];
    if (al) {
        const Rsrc rl = make_rsrc(reinterpret_cast<const char*>(base) - mis, mis + R * C * E);
#pragma unroll
        for (int j = 0; j < JR; j++)
            rt[j].load_lines(rl, mis, ar0, ac0 + (jl + j) * W::TS, R, C, lane, jl + j <= n);
#pragma unroll
        for (int j = 0; j < JL; j++)
            lt[j].load_lines(rl, mis, ar0, ac0 + j * W::TS, R, C, lane, (uint32_t)j < jl);
    } else {
#pragma unroll
        for (int j = 0; j < JR; j++) {
            const uint32_t k = rk(j);
            uint32_t r0, c0;
            origin(k, &r0, &c0);
            rt[j].load(rs, r0, c0, R, C, lane, k < nk);
        }
#pragma unroll
        for (int j = 0; j < JL; j++) {
            const uint32_t k = wave + j * NW;
            uint32_t r0, c0;
            origin(k, &r0, &c0);
            lt[j].load(rs, r0, c0, R, C, lane, k < kl);
        }
    }
    // Without line-aligned loads in the member, waves transpose their tiles
    // and stage their LDS tiles as the loads come back (overlapping the later
    // loads), before the member arrives at the barrier.  With them, the
    // member arrives first and then its line-aligned waves realign each tile
    // with tile i + 1's lines (still raw: tiles go in order; after the last
    // live one, the tail), so the shuffling overlaps the wait for the other
    // members; an LDS tile goes to LDS as soon as it is in place, which frees
    // its registers.
    if (sw) {
        wait_vmem();
        __syncthreads();
        if (g > 1 && threadIdx.x == 0) {
            if constexpr (DEP) flag_arrive(bar, m);
            else group_arrive(bar, m, g);
        }
    }
#pragma unroll
    for (int j = 0; j < JL; j++) {
        const uint32_t k = wave + j * NW;
        if (al && (uint32_t)j < jl) {
            const W& lnext = lt[j + 1 < JL ? j + 1 : j];
            W nx;
#pragma unroll
            for (uint32_t q = 0; q < W::PER; q++) nx.v[q] = (uint32_t)j + 1 < jl ? lnext.v[q] : rt[0].v[q];
            lt[j].realign(nx, mis, ar0, ac0 + j * W::TS, C, lane);
        }
        if (k < kl) {
            lt[j].transpose();
#pragma unroll
            for (uint32_t q = 0; q < W::PER; q++) lds[unit(k, W::PER * (lane & 7) + q, lane >> 3)] = lt[j].v[q];
        }
    }
#pragma unroll
    for (int j = 0; j < JR; j++) {
        if (j + 1 < JR && al && jl + j < n)
            rt[j].realign(rt[j + 1 < JR ? j + 1 : j], mis, ar0, ac0 + (jl + j) * W::TS, C, lane);
        rt[j].transpose();
    }
    if (!sw) {
        wait_vmem();
        __syncthreads();
        if (g > 1 && threadIdx.x == 0) {
            if constexpr (DEP) flag_arrive(bar, m);
            else group_arrive(bar, m, g);
        }
    }
    if constexpr (DEP) {
        if (g > 1 && threadIdx.x < 64) {
            uint32_t lo, hi;
            gcs_deps<W::TS>(m, nk, K, ntiles, g, R, C, dntj.d, ord, &lo, &hi);
            if (!flag_wait(bar, lo, hi, wd.extra, wd.limit, lane) && lane == 0) report_watchdog(ctl);
        }
    } else {
        if (g > 1 && threadIdx.x == 0 && !group_wait(bar, g, wd.extra, wd.limit)) report_watchdog(ctl);
    }
    __syncthreads();
    released();
    // Staged tiles: LDS unit(k, R_o, c) holds 16-byte chunk c of output row
    // R_o of tile k.  In the batch kernel, when the member's staged
    // tiles lie in full bands, each store writes one output row across the
    // band's eight tiles of one tile column (1 KB runs, lane r * 8 + c: tile
    // row r, chunk c) instead of each wave writing its own tiles' 128-byte
    // segments: uneven batch +0.3-0.6 % (4268-4279 -> 4280-4298 GB/s over
    // two A/B runs of three).
    bool runs = false;
    if constexpr (AL) runs = ord.bh == NW && kl > 0 && m * K + kl <= ord.full * ord.dband.d;
    // Positions P0 .. P1 - 1 of the member's tiles, staged in LDS as tiles
    // p - P0, stored as runs.
    auto store_runs = [&](uint32_t P0, uint32_t P1) {
        constexpr uint32_t E_ = sizeof(T);
        const uint32_t G0 = P0 / NW * NW;
        const uint32_t ng = (P1 + NW - 1) / NW - P0 / NW;
        static_assert(W::TS % NW == 0, "whole rows per wave");
        const uint32_t r = lane >> 3, c = lane & 7;
        // wave w stores rows w, w + NW, .. of each tile column in turn
        for (uint32_t gi = 0; gi < ng; gi++) {
            const uint32_t gp = G0 + NW * gi;  // first position of the tile column
            const uint32_t p = gp + r;
            const bool valid = p >= P0 && p < P1;
            const uint32_t kk = valid ? p - P0 : 0u;
            const uint32_t b = ord.dband.div(gp), tj = (gp - b * ord.dband.d) / NW;
            const uint32_t ocol = (b * NW + r) * W::TS + W::PER * c;
            const bool vcol = valid && ocol < R;
            const bool full = ocol + W::PER <= R;
            const bool part = (b + 1) * NW * W::TS > R;  // the band holds the last, partial tile row
#pragma unroll
            for (uint32_t i = 0; i < W::TS / NW; i++) {
                const uint32_t ro = wave + i * NW;
                const u32x4 x = lds[unit(kk, ro, c)];
                const uint32_t orow = tj * W::TS + ro;
                const bool ok = vcol && orow < C;
                const uint32_t o = (orow * R + ocol) * E_;
                __builtin_amdgcn_raw_buffer_store_b128(x, rs, ok && full ? o : kOutside, 0, SP);
                if (part && __any(ok && !full))  // some lane holds a row's partial end
                    store_head<T, SP>(x, rs, o, R - ocol, ok && !full);
            }
        }
    };
    if (runs) {
        store_runs(m * K, m * K + kl);
    } else {
#pragma unroll
        for (int j = 0; j < JL; j++) {
            const uint32_t k = wave + j * NW;
            if (k < kl) {
                W x;
#pragma unroll
                for (uint32_t q = 0; q < W::PER; q++)
                    x.v[q] = lds[unit(k, W::PER * (lane & 7) + q, lane >> 3)];
                uint32_t r0, c0;
                origin(k, &r0, &c0);
                x.template store_t<SP>(rs, r0, c0, R, C, lane, true);
            }
        }
    }
    // Register tiles: each wave its own (restaging them through LDS to store
    // them as runs too cost two more workgroup barriers: uneven batch 4268
    // -> 3948 GB/s).
#pragma unroll
    for (int j = 0; j < JR; j++) {
        const uint32_t k = rk(j);
        uint32_t r0, c0;
        origin(k, &r0, &c0);
        rt[j].template store_t<SP>(rs, r0, c0, R, C, lane, k < nk);
    }
    if (g > 1 && threadIdx.x == 0) {
        if constexpr (DEP) flag_depart(bar, g);
        else group_depart(bar, g);
    }
}

// GCS kernel: slot s (numbered by xcd_slot) is member s % g of
// slice s / g.  SP: the stores' cache policy (kStoreDevice or 0, write-back);
// ord: the members' tile order.
template <typename T, int BLOCK, int JR, int SP>
__global__ void __launch_bounds__(BLOCK)
k_gcs(T* __restrict__ data, uint32_t R, uint32_t C, uint64_t nslices, uint32_t g, uint32_t K,
       uint32_t KL, Div32 dntj, uint32_t ntiles, uint32_t* __restrict__ ctl,
       uint32_t* __restrict__ cnt, uint32_t nxcd, TileOrder ord) {
    IPT_DYN_LDS(u32x4, lds);
    const uint64_t total = nslices * g;
    const uint64_t slice_elems = (uint64_t)R * C;
    for (uint64_t slot = xcd_slot(blockIdx.x, gridDim.x, nxcd, g); slot < total; slot += gridDim.x) {
        const uint64_t S = slot / g;
        const uint32_t m = (uint32_t)(slot - S * g);
        // plain loads and the counter barrier: in a single launch every
        // member loads at once and the wait at the barrier is short, so
        // realigning after the arrival would add to the critical path (f32
        // 4093 x 4099: 4151 -> 3600 GB/s), and polling per-member flags
        // costs more than it saves
        gcs_body<T, BLOCK, JR, SP, false, false>(data + S * slice_elems, R, C, g, K, KL, dntj, ntiles, m,
                                          g > 1 ? cnt + bar_words(g) * S : nullptr, ctl, lds, ord);
        __syncthreads();  // LDS is reused by the next slot
    }
}

}  // namespace ipt

namespace ipt {

// ---------------------------------------------------------------------------
// Batched uneven transposes: one persistent launch over a table of work
// items.  Each instance (d3 slices of R x C, 4- or 8-byte elements) is cut
// into items of one kind:
//   BK_SLICE  `sub` consecutive slices staged whole in LDS (slice_body)
//   BK_GCS    member `sub` of the group of g that owns one slice (gcs_body);
//             the g members of a slice occupy consecutive items
//   BK_SWAP   tile pair `sub` of one square slice (swap_body)
// Workgroups take items from a queue in increasing order, one at a time, and
// a workgroup never holds an item it is not running.  Deadlock freedom: the
// grid never exceeds the resident workgroups and every g <= grid; each group
// before the queue's frontier has all its members taken, hence running, so
// the earliest unfinished group completes; the frontier group gets its
// remaining members from workgroups freed as earlier groups complete (all
// grid workgroups would suffice).  Only group members ever wait.
enum BatchKind : uint32_t { BK_SLICE = 0, BK_GCS = 1, BK_SWAP = 2, BK_SWAP_VEC = 3, BK_SWAP_RS = 4 };

struct BatchInst {
    uint64_t data;        // device address of slice 0
    uint64_t bar;         // word offset of slice 0's barrier state (BK_GCS)
    uint32_t R, C, esz, kind;
    uint32_t g, K, KL, ntiles;   // BK_GCS
    Div32 dntj;                  // BK_GCS
    uint32_t jr;                 // BK_GCS: register-tile class (gcs_jr_class)
    TileOrder ord;               // BK_GCS: which tiles each member owns
    uint32_t P;                  // BK_SLICE: LDS row pitch
    Div32 dS, dC, dR;            // BK_SLICE
};

struct BatchItem {
    uint32_t inst;
    uint32_t sub;    // slices in the item / member / tile pair
    uint64_t slice;  // first slice
};

constexpr int kBatchBlock = 512;

template <typename T, typename HOOK>
__device__ __forceinline__ void batch_item(const BatchInst& I, const BatchItem& it, uint32_t* ctl,
                                           uint32_t* bars, T* lds, HOOK released) {
    constexpr int TS_SWAP = sizeof(T) == 4 ? 64 : 32;
    T* base = reinterpret_cast<T*>(I.data) + it.slice * (uint64_t)I.R * I.C;
    switch (I.kind) {
    case BK_SLICE:
        slice_body<T>(base, I.R, I.C, I.P, it.sub, I.dS, I.dC, I.dR, lds);
        break;
    case BK_GCS: {
        // Plain (write-back) stores: a long-running persistent kernel does not
        // pay for the end-of-kernel write-back that device-scope stores avoid
        // for single launches, and L2 merges the partial lines of unaligned
        // rows (uneven 4096-matrix batch, its cooperative groups: 3.74 ->
        // 3.33 ms).
        uint32_t* bar = bars + I.bar + bar_words(I.g) * it.slice;
        u32x4* l = reinterpret_cast<u32x4*>(lds);
        static_assert(kBatchJrClass == 1, "one body per register-tile class");
        if (I.jr == 0)
            gcs_body<T, kBatchBlock, gcs_jr_of(sizeof(T), 0), 0, true, true, HOOK>(
                base, I.R, I.C, I.g, I.K, I.KL, I.dntj, I.ntiles, it.sub, bar, ctl, l, I.ord, released);
        else
            gcs_body<T, kBatchBlock, gcs_jr_of(sizeof(T), 1), 0, true, true, HOOK>(
                base, I.R, I.C, I.g, I.K, I.KL, I.dntj, I.ntiles, it.sub, bar, ctl, l, I.ord, released);
        break;
    }
    case BK_SWAP_RS:  // buffer-resource form: slices below 2^31 bytes, any side
        swap_body_rs<T, TS_SWAP, kBatchBlock>(base, I.R, it.sub, lds, false);
        break;
    case BK_SWAP_VEC:
        swap_body<T, TS_SWAP, kBatchBlock, true>(base, I.R, it.sub, lds, false);
        break;
    default:
        swap_body<T, TS_SWAP, kBatchBlock, false>(base, I.R, it.sub, lds, false);
        break;
    }
}

}  // namespace ipt
