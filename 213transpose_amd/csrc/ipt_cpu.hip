// ipt_cpu.hip — the in-place transpose on the host CPU.
//
// Replaces inplace::openmp::transpose(bool row_major, T* data, int m, int n,
// T* tmp) (include/openmp.h:7-8), which the reference declares but never
// implements.  Host code only (no kernels): the library's own index maps
// (ipt_plan.h, the code the GPU kernels run) drive three cache-blocked passes
// on host threads:
//   square slices   tile pairs (I,J) / (J,I) swapped through a per-thread
//                   buffer: one sweep;
//   otherwise       the C2R decomposition (src/transpose.cu:30-47) when the
//                   rows are at least as long as the columns, else R2C on the
//                   swapped view (src/transpose.cu:70-89), so the view always
//                   has R <= C: column gathers over strips of W columns x all
//                   R rows (each strip copied into a per-thread buffer, then
//                   written back gathered), row gathers over whole rows (each
//                   row copied into a buffer, then written back gathered; a
//                   row longer than kCpuLongRow bytes is copied and gathered by
//                   all threads together).
// Index arithmetic uses the 64-bit plan (exact below 2^63, the same bound
// the GPU planner enforces).  Payloads are moved as raw bits.
#include "ipt_internal.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace ipt {

namespace {

constexpr int64_t kCpuStripBytes = 256 * 1024;   // one strip buffer: R x W elements
constexpr int64_t kCpuLongRow = 4 << 20;         // rows above this are shared by all threads
constexpr int64_t kCpuMinStrip = 16;             // columns per strip, at least

// Run fn(lo, hi) over [0, n) split into `threads` contiguous ranges; the
// calling thread takes the first range.
template <typename F>
void parallel_for(int64_t n, int threads, F fn) {
    if (n <= 0) return;
    const int64_t t = std::max<int64_t>(1, std::min<int64_t>(threads, n));
    if (t == 1) {
        fn(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve((size_t)t - 1);
    for (int64_t w = 1; w < t; w++)
        pool.emplace_back([=] { fn(n * w / t, n * (w + 1) / t, (int)w); });
    fn(0, n / t, 0);
    for (auto& th : pool) th.join();
}

// How a problem runs and the scratch it needs (elements).  Many slices:
// each worker takes whole slices, one at a time, with private buffers
// (a strip buffer and a row buffer).  Few slices: every pass of a slice is
// split over all workers; each has a strip buffer and a row buffer, except
// that rows above kCpuLongRow bytes (or fewer rows than workers) use one
// shared row buffer that all workers copy and gather together.
struct CpuMode {
    bool square = false;
    bool by_slice = false;     // workers take whole slices
    int64_t strip_w = 0;       // columns per strip
    int64_t per_thread = 0;    // private elements per worker
    int64_t shared_row = 0;    // elements of the shared long-row buffer
    int64_t total(int threads) const { return per_thread * threads + shared_row; }
};

template <typename T>
constexpr int64_t swap_tile() { return 64 / (int64_t)sizeof(T) * 4; }  // 64 (4-byte) / 32 (8-byte)

CpuMode cpu_mode(int esz, int64_t d1, int64_t d2, int64_t d3, int threads) {
    CpuMode m;
    const int64_t slice_bytes = d1 * d2 * esz;
    m.by_slice = d3 >= 2 * threads || (d3 > 1 && slice_bytes <= ((int64_t)8 << 20));
    if (d1 == d2) {
        m.square = true;
        const int64_t ts = esz == 4 ? swap_tile<uint32_t>() : swap_tile<uint64_t>();
        m.per_thread = 2 * ts * ts;
        return m;
    }
    const int64_t R = std::min(d1, d2), C = std::max(d1, d2);
    m.strip_w = std::max<int64_t>(kCpuMinStrip, kCpuStripBytes / (R * esz));
    m.strip_w = std::min<int64_t>(m.strip_w, C);
    // rows shared by all workers: long ones, or fewer rows than workers
    const bool long_row = !m.by_slice && (C * esz > kCpuLongRow || R < threads);
    m.per_thread = R * m.strip_w + (long_row ? 0 : C);
    m.shared_row = long_row ? C : 0;
    return m;
}

// new[i][j] = old[Map::src(i, j)][j] over columns [0, C) of one R x C slice,
// strips of W columns, strips split over the threads.
template <typename T, typename Ix, typename Map>
void col_pass(T* d, const Plan<Ix>& p, int64_t W, T* scr, int64_t per_thread, int threads) {
    const int64_t R = (int64_t)p.R, C = (int64_t)p.C;
    const int64_t nstrips = (C + W - 1) / W;
    parallel_for(nstrips, threads, [&](int64_t lo, int64_t hi, int tid) {
        T* buf = scr + per_thread * tid;
        for (int64_t s = lo; s < hi; s++) {
            const int64_t j0 = s * W, w = std::min(W, C - j0);
            for (int64_t i = 0; i < R; i++) std::memcpy(buf + i * W, d + i * C + j0, (size_t)w * sizeof(T));
            for (int64_t i = 0; i < R; i++) {
                T* out = d + i * C + j0;
                for (int64_t jj = 0; jj < w; jj++)
                    out[jj] = buf[(int64_t)Map::src(p, (Ix)i, (Ix)(j0 + jj)) * W + jj];
            }
        }
    });
}

// new[i][j] = old[i][Map::src(i, j)] over the R rows of one slice.
template <typename T, typename Ix, typename Map>
void row_pass(T* d, const Plan<Ix>& p, int64_t strip_elems, T* scr, int64_t per_thread,
              T* shared_row, int threads) {
    const int64_t R = (int64_t)p.R, C = (int64_t)p.C;
    if (shared_row) {  // long rows: every thread copies and gathers a share of each row
        for (int64_t i = 0; i < R; i++) {
            T* row = d + i * C;
            parallel_for(C, threads, [&](int64_t lo, int64_t hi, int) {
                std::memcpy(shared_row + lo, row + lo, (size_t)(hi - lo) * sizeof(T));
            });
            parallel_for(C, threads, [&](int64_t lo, int64_t hi, int) {
                for (int64_t j = lo; j < hi; j++)
                    row[j] = shared_row[(int64_t)Map::src(p, (Ix)i, (Ix)j)];
            });
        }
        return;
    }
    parallel_for(R, threads, [&](int64_t lo, int64_t hi, int tid) {
        T* buf = scr + per_thread * tid + strip_elems;
        for (int64_t i = lo; i < hi; i++) {
            T* row = d + i * C;
            std::memcpy(buf, row, (size_t)C * sizeof(T));
            for (int64_t j = 0; j < C; j++) row[j] = buf[(int64_t)Map::src(p, (Ix)i, (Ix)j)];
        }
    });
}

// Square N x N slice: tile pairs swapped (tile side TS), pairs split over the
// threads; each pair is staged in a per-thread buffer of 2 * TS * TS.
template <typename T>
void square_swap(T* d, int64_t N, T* scr, int64_t per_thread, int threads) {
    constexpr int64_t TS = swap_tile<T>();
    const int64_t nt = (N + TS - 1) / TS;
    const int64_t pairs = nt * (nt + 1) / 2;
    parallel_for(pairs, threads, [&](int64_t lo, int64_t hi, int tid) {
        T* A = scr + per_thread * tid;
        T* B = A + TS * TS;
        int64_t J = 0;
        while ((J + 1) * (J + 2) / 2 <= lo) J++;
        int64_t I = lo - J * (J + 1) / 2;
        for (int64_t t = lo; t < hi; t++) {
            const int64_t r0 = I * TS, c0 = J * TS;
            const int64_t h = std::min(TS, N - r0), w = std::min(TS, N - c0);  // tile (I,J) is h x w
            for (int64_t y = 0; y < h; y++) std::memcpy(A + y * TS, d + (r0 + y) * N + c0, (size_t)w * sizeof(T));
            if (I != J)
                for (int64_t y = 0; y < w; y++) std::memcpy(B + y * TS, d + (c0 + y) * N + r0, (size_t)h * sizeof(T));
            const T* src = I == J ? A : B;  // tile (I,J) <- transpose of tile (J,I)
            for (int64_t y = 0; y < h; y++)
                for (int64_t x = 0; x < w; x++) d[(r0 + y) * N + c0 + x] = src[x * TS + y];
            if (I != J)
                for (int64_t y = 0; y < w; y++)
                    for (int64_t x = 0; x < h; x++) d[(c0 + y) * N + r0 + x] = A[x * TS + y];
            if (++I > J) {
                J++;
                I = 0;
            }
        }
    });
}

// The three passes of the C2R (c2r) or R2C decomposition of the view h, in
// index type Ix (uint32_t where the plan fits it: cheaper divisions).
template <typename T, typename Ix>
void decompose(T* d, const HostPlan& h, bool c2r, T* scr, T* shared_row, const CpuMode& m,
               int threads) {
    const Plan<Ix> p = Plan<Ix>::make(h);
    const int64_t W = m.strip_w, strip_elems = h.R * W, pt = m.per_thread;
    if (c2r) {  // src/transpose.cu:30-47: pre-rotate (c > 1), row shuffle, column op
        if (h.c > 1) col_pass<T, Ix, MapC2RPre>(d, p, W, scr, pt, threads);
        row_pass<T, Ix, MapC2RShuffle>(d, p, strip_elems, scr, pt, shared_row, threads);
        col_pass<T, Ix, MapC2RPost>(d, p, W, scr, pt, threads);
    } else {  // src/transpose.cu:70-89: column op, row shuffle, post-rotate (c > 1)
        col_pass<T, Ix, MapR2CPre>(d, p, W, scr, pt, threads);
        row_pass<T, Ix, MapR2CShuffle>(d, p, strip_elems, scr, pt, shared_row, threads);
        if (h.c > 1) col_pass<T, Ix, MapR2CPost>(d, p, W, scr, pt, threads);
    }
}

// One slice of d2 rows x d1 columns on `threads` workers; scr holds each
// worker's private buffers (per_thread elements apart).
template <typename T>
void slice_transpose(T* d, int64_t d1, int64_t d2, T* scr, T* shared_row, const CpuMode& m,
                     int threads) {
    if (d1 == 1 || d2 == 1) return;
    if (m.square) {
        square_swap<T>(d, d1, scr, m.per_thread, threads);
        return;
    }
    const bool c2r = d1 >= d2;
    const HostPlan h = c2r ? HostPlan::make(d2, d1) : HostPlan::make(d1, d2);
    if (h.fits_index32()) decompose<T, uint32_t>(d, h, c2r, scr, shared_row, m, threads);
    else decompose<T, uint64_t>(d, h, c2r, scr, shared_row, m, threads);
}

int resolve_threads(int threads) {
    if (threads > 0) return std::min(threads, 1024);
    const unsigned hc = std::thread::hardware_concurrency();
    return hc ? (int)std::min(hc, 1024u) : 1;
}

// The view slice_transpose picks must have an exact 64-bit plan.
bool cpu_plan_ok(int64_t d1, int64_t d2) {
    if (d1 == 1 || d2 == 1 || d1 == d2) return true;
    const HostPlan h = d1 >= d2 ? HostPlan::make(d2, d1) : HostPlan::make(d1, d2);
    return h.fits_index64();
}

template <typename T>
void transpose_all(T* d, int64_t d1, int64_t d2, int64_t d3, T* scr, const CpuMode& m,
                   int threads) {
    const int64_t S = d1 * d2;
    if (m.by_slice) {
        parallel_for(d3, threads, [&](int64_t lo, int64_t hi, int tid) {
            for (int64_t s = lo; s < hi; s++)
                slice_transpose<T>(d + s * S, d1, d2, scr + m.per_thread * tid, nullptr, m, 1);
        });
        return;
    }
    T* shared_row = m.shared_row ? scr + m.per_thread * threads : nullptr;
    for (int64_t s = 0; s < d3; s++) slice_transpose<T>(d + s * S, d1, d2, scr, shared_row, m, threads);
}

}  // namespace

static ipt_status cpu_transpose_213(void* data, int esz, int64_t d1, int64_t d2, int64_t d3,
                                    void* tmp, int threads) {
    ipt_status st = validate(data, esz, d1, d2, d3);
    if (st != IPT_OK) return st;
    if (!cpu_plan_ok(d1, d2)) return IPT_ERR_TOO_LARGE;
    // Device memory is not the host's to touch; everything else (pageable,
    // pinned, managed) is.  Without a usable device every pointer is host
    // memory.
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, data) == hipSuccess) {
        if (at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeArray) return IPT_ERR_INVALID_ARG;
    } else {
        (void)hipGetLastError();
    }
    const int t = resolve_threads(threads);
    const CpuMode m = cpu_mode(esz, d1, d2, d3, t);
    const int64_t need = (d1 == 1 || d2 == 1) ? 0 : m.total(t);
    void* scr = tmp;
    if (!scr && need > 0) {
        scr = std::malloc((size_t)need * (size_t)esz);
        if (!scr) return IPT_ERR_OUT_OF_MEMORY;
    }
    if (esz == 4) transpose_all<uint32_t>((uint32_t*)data, d1, d2, d3, (uint32_t*)scr, m, t);
    else transpose_all<uint64_t>((uint64_t*)data, d1, d2, d3, (uint64_t*)scr, m, t);
    if (scr != tmp) std::free(scr);
    return IPT_OK;
}

}  // namespace ipt

using namespace ipt;

extern "C" {

ipt_status ipt_cpu_transpose_213(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                                 void* tmp, int threads) {
    return cpu_transpose_213(data, elem_bytes, d1, d2, d3, tmp, threads);
}

ipt_status ipt_cpu_transpose(int row_major, void* data, int elem_bytes, int64_t m, int64_t n,
                             void* tmp, int threads) {
    if (row_major) return cpu_transpose_213(data, elem_bytes, n, m, 1, tmp, threads);
    return cpu_transpose_213(data, elem_bytes, m, n, 1, tmp, threads);
}

int64_t ipt_cpu_scratch_bytes(int elem_bytes, int64_t d1, int64_t d2, int64_t d3, int threads) {
    if ((elem_bytes != 4 && elem_bytes != 8) || d1 <= 0 || d2 <= 0 || d3 <= 0) return -1;
    if (d1 == 1 || d2 == 1) return 0;
    const int t = resolve_threads(threads);
    return cpu_mode(elem_bytes, d1, d2, d3, t).total(t) * elem_bytes;
}

}  // extern "C"
