/*
 * inplace_transpose.h — C ABI of the MI355X (gfx950) in-place transpose engine
 * (libinplace_hip.so).
 *
 * Semantics follow the reference's oracle TensorUtil::seq_tt with permutation
 * {1,0,2} (src/tensor_util.cpp:80-92): dims are listed fastest-first, so the
 * buffer holds d3 contiguous slices, each a row-major matrix of d2 rows and d1
 * columns, and each slice is replaced in place by its row-major d1 x d2
 * transpose:
 *
 *     out[x1 + x0*d2 + x2*d1*d2] = in[x0 + x1*d1 + x2*d1*d2]
 *
 * Payloads are moved as raw bits (elem_bytes 4 or 8: the reference's int,
 * long long, float and double instantiations, src/transpose.cu:126-129).
 *
 * Conventions (all entry points):
 *   - `data` is memory the current device can reach, owned by the caller:
 *     device memory of the current device (hipMalloc), managed memory
 *     (hipMallocManaged) or pinned / registered host memory.  Managed memory
 *     is first prefetched to the current device on `stream`, as the reference
 *     does (src/transpose.cu:100-104), except while `stream` is being
 *     captured; where the system keeps managed memory on the host the
 *     prefetch moves nothing and, like pinned host memory, the data is
 *     transposed in place over the host link.  Pageable host memory (malloc,
 *     stack) and device memory of another device are refused with
 *     IPT_ERR_INVALID_ARG before anything is enqueued (ipt_memory_kind).
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *     Work is stream-ordered; the calls return once it is enqueued.
 *   - Arguments are validated before any HIP call.  Nothing prints, throws or
 *     exits; every failure is a returned ipt_status.
 *   - Calls may come from several host threads; per-device state is guarded.
 */
#ifndef INPLACE_TRANSPOSE_H
#define INPLACE_TRANSPOSE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    IPT_OK = 0,
    IPT_ERR_INVALID_ARG = 1,  /* NULL pointer, non-positive dim, bad elem_bytes, ... */
    IPT_ERR_UNSUPPORTED = 2,  /* a forced variant that cannot run this shape */
    IPT_ERR_TOO_LARGE = 3,    /* sizes beyond the 64-bit index plan */
    IPT_ERR_HIP = 4,          /* a HIP runtime call failed */
    IPT_ERR_NO_DEVICE = 5,    /* no usable GPU */
    IPT_ERR_OUT_OF_MEMORY = 6,/* library scratch could not be allocated */
    IPT_ERR_IO = 7,           /* selector / tuning file missing or corrupt */
    IPT_ERR_NOT_READY = 8     /* object used before it was built / trained */
} ipt_status;

/* Human-readable name of a status code (static storage). */
const char* ipt_status_string(ipt_status s);

/* Detail of the last IPT_ERR_HIP / IPT_ERR_OUT_OF_MEMORY on the calling
 * thread (thread-local storage; empty when none). */
const char* ipt_last_error(void);

/* Where a pointer lives, as the entry points classify it
 * (hipPointerGetAttributes; the values are hipMemoryType's). */
typedef enum {
    IPT_MEM_PAGEABLE = 0, /* host memory unknown to HIP: refused unless the device
                             can access pageable memory */
    IPT_MEM_HOST = 1,     /* pinned / registered host memory: used in place */
    IPT_MEM_DEVICE = 2,   /* device memory (hipMalloc) */
    IPT_MEM_MANAGED = 3   /* managed memory: prefetched to the current device */
} ipt_memory_kind_t;

/* Kind of the memory at `data` (an ipt_memory_kind_t); *device, when not
 * NULL, receives the owning device of device memory, else -1.  Enqueues
 * nothing.  Replaces the pointer requirement implied by the reference's
 * cudaMemPrefetchAsync call (src/transpose.cu:104), which exits on anything
 * but managed memory. */
int ipt_memory_kind(const void* data, int* device);

/* ---- transpose entry points ------------------------------------------- */

/* Replaces inplace::transpose<T>(T* data, int d1, int d2, int d3)
 * (include/transpose.h:17, src/transpose.cu:99-124): d3 slices of d2 rows x
 * d1 columns, each transposed in place.  The variant is chosen by the
 * dispatcher (learned selector when one is loaded, else the heuristic). */
ipt_status ipt_transpose_213(void* data, int elem_bytes, int64_t d1, int64_t d2,
                             int64_t d3, void* stream);

/* Replaces inplace::c2r::transpose<T> (include/transpose.h:7,
 * src/transpose.cu:30-47): forces the C2R decomposition; slices of d2 rows x
 * d1 columns (kernel view R = d2, C = d1). */
ipt_status ipt_c2r_transpose(void* data, int elem_bytes, int64_t d1, int64_t d2,
                             int64_t d3, void* stream);

/* Replaces inplace::r2c::transpose<T> (include/transpose.h:13,
 * src/transpose.cu:70-89): forces the R2C decomposition with the reference's
 * argument order: slices of d1 rows x d2 columns (kernel view R = d2, C = d1)
 * are transposed to d2 rows x d1 columns. */
ipt_status ipt_r2c_transpose(void* data, int elem_bytes, int64_t d1, int64_t d2,
                             int64_t d3, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation): the same result as ipt_transpose_213, calling no hipMalloc,
 * hipFree or other allocating routine.  The cooperative kernel's barrier
 * state and the decomposition's scratch staging live in one static 32 MiB
 * device array (1 MiB barrier state, 31 MiB staging), which every device the
 * library runs on holds until the process ends.  Routes that need
 * allocations run differently: the block-cycle path (move lists) and
 * cooperative groups whose barrier state exceeds 1 MiB take the
 * decomposition, whose scratch tiers stage through the array in pieces;
 * IPT_ERR_TOO_LARGE, before anything is enqueued, when one of those pieces
 * (a row of the R x C view, or a 256-byte strip of all R rows) exceeds
 * 31 MiB.  Every call on a device shares that device's array and leaves its
 * barrier state at zero: calls that overlap on one device (several streams
 * or host threads at once) race.  A barrier that gives up is reported by
 * ipt_device_check. */
ipt_status ipt_transpose_213_noalloc(void* data, int elem_bytes, int64_t d1, int64_t d2,
                                     int64_t d3, void* stream);

/* The row-/column-major interface of inplace::openmp::transpose(bool
 * row_major, T* data, int m, int n, T* tmp) (include/openmp.h:7-8) on the GPU:
 * row-major m x n maps to (d1 = n, d2 = m, d3 = 1), column-major to
 * (d1 = m, d2 = n, d3 = 1); `tmp` is not needed (scratch is library-owned).
 * The host form, which the reference's declaration implies, is
 * ipt_cpu_transpose below. */
ipt_status ipt_transpose(int row_major, void* data, int elem_bytes, int64_t m,
                         int64_t n, void* stream);

/* ---- host CPU path ------------------------------------------------------ */

/* Replaces inplace::openmp::transpose(bool row_major, T* data, int m, int n,
 * T* tmp) (include/openmp.h:7-8; declared but never implemented in the
 * reference): the in-place transpose of an m x n matrix in host memory, on
 * `threads` host threads (<= 0: one per hardware thread).  Row-major m x n
 * maps to (d1 = n, d2 = m), column-major to (d1 = m, d2 = n), as
 * ipt_transpose.  The same decomposition as the GPU path (C2R when the rows
 * are at least as long as the columns, else R2C; square matrices by tile-pair
 * swaps), driven by the library's own index maps; bit-exact.  `tmp`: NULL
 * (the call allocates its scratch and frees it before returning) or caller
 * scratch of at least ipt_cpu_scratch_bytes(elem_bytes, d1, d2, 1, threads)
 * bytes.  `data` must be host-accessible (pageable, pinned or managed
 * memory); device memory is refused with IPT_ERR_INVALID_ARG.  Blocking:
 * returns when the transpose is done.  IPT_ERR_TOO_LARGE where the 64-bit
 * index plan is not exact (a view with b = C / gcd(R, C) above ~3e9). */
ipt_status ipt_cpu_transpose(int row_major, void* data, int elem_bytes, int64_t m, int64_t n,
                             void* tmp, int threads);

/* The 213 form of ipt_cpu_transpose (d3 slices of d2 rows x d1 columns, each
 * transposed in place; inplace::transpose<T>'s semantics on the host). */
ipt_status ipt_cpu_transpose_213(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                                 void* tmp, int threads);

/* Bytes of caller scratch ipt_cpu_transpose_213 needs for this problem and
 * thread count (0 when nothing moves; -1 for invalid arguments). */
int64_t ipt_cpu_scratch_bytes(int elem_bytes, int64_t d1, int64_t d2, int64_t d3, int threads);

/* Any permutation of a rank-3 tensor in place, in the oracle's convention
 * (TensorUtil::seq_tt, src/tensor_util.cpp:80-92): dims d1, d2, d3 fastest
 * first, and output dim j is source dim perm[j]; {1,0,2} is the 213
 * transpose above.  The next step the reference's oracle points to (it
 * handles any rank and permutation); no in-place counterpart in the
 * reference.  One in-place pass for every permutation but the reversal
 * {2,1,0}, which takes two: 213 transposes of the dims or of merged dims,
 * and the exchange of the outer two dims with the innermost kept
 * contiguous (a transpose of a matrix of innermost-dim chunks, of any size;
 * IPT_ERR_UNSUPPORTED past 2^23 chunks per outer slice or when the saved
 * cycle heads of one slice exceed the library's 256 MB scratch). */
ipt_status ipt_permute3(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                        const int* perm, void* stream);

/* Any permutation of a tensor of rank 1..6 in place (same convention: dims
 * fastest first, output dim j is source dim perm[j]).  Each pass exchanges
 * two adjacent groups of dims, the dims before them kept innermost: a 213
 * transpose (of elements or of 4-/8-byte super-elements) or a transpose of
 * matrices of inner chunks (chunks of any size, limits as ipt_permute3);
 * the fewest passes are chosen (dims of size 1 cost nothing). */
ipt_status ipt_permute(void* data, int elem_bytes, int rank, const int64_t* dims, const int* perm,
                       void* stream);

/* ---- variants (used by the autotuner and the selector) ---------------- */

typedef enum {
    IPT_VARIANT_AUTO = 0,        /* dispatcher's choice */
    IPT_VARIANT_SLICE_LDS = 1,   /* whole slice staged in LDS: one sweep */
    IPT_VARIANT_SQUARE_SWAP = 2, /* square slices: tile-pair swap, one sweep */
    IPT_VARIANT_C2R = 3,         /* C2R, view R = d2, C = d1 */
    IPT_VARIANT_R2C = 4,         /* R2C, view R = d1, C = d2 */
    IPT_VARIANT_GCS = 5,         /* group-cooperative staged: a group of workgroups
                                    holds a slice in LDS and registers, then
                                    writes it transposed; one sweep for slices
                                    up to ~109 MB (f32) or ~117 MB (f64), any row length and element
                                    alignment */
    IPT_VARIANT_BLOCK_CYCLE = 6, /* one side cut into chunks of T elements:
                                    the chunks moved by segmented cycle
                                    following, blocks of T x (other side)
                                    transposed by a one-sweep kernel; about two
                                    sweeps.  With a short side and LDS-sized
                                    blocks it is the skinny path of
                                    src/skinny.cu:313-363, rebuilt; it also
                                    takes large rectangles.  Any element
                                    alignment (chunks move as 16-, 8- or 4-byte
                                    vectors) */
    IPT_VARIANT_COUNT = 7
} ipt_variant;

/* Tuning knobs; zero fields mean "library default". */
typedef struct {
    int32_t lds_bytes;    /* LDS cap per workgroup; smaller values push passes to
                             the scratch tiers (tests use this) */
    int32_t block;        /* threads per workgroup (power of two, 64..1024) */
    int32_t strip;        /* column-strip width for LDS column passes (elements) */
    int32_t grid_mult;    /* workgroups per CU for grid-stride kernels */
    int32_t chunk_min;    /* block-cycle path: chunk sizes (bytes) the plan may    */
    int32_t chunk_max;    /*   pick from (default 512 .. 32768); its price model
                             ranks the divisors T of the chunked side inside them */
    int32_t tile;         /* square swap: tile side in elements (64 / 32 for 4- /
                             8-byte elements by default; half that is the other
                             compiled size) */
} ipt_tune;

/* Same semantics as ipt_transpose_213 with a forced variant and optional
 * knobs (tune may be NULL).  IPT_ERR_UNSUPPORTED when the variant cannot run
 * the shape (e.g. SQUARE_SWAP with d1 != d2). */
ipt_status ipt_transpose_variant(void* data, int elem_bytes, int64_t d1, int64_t d2,
                                 int64_t d3, int variant, const ipt_tune* tune,
                                 void* stream);

/* Whether `variant` can run the shape on the current device (1) or not (0).
 * Host-only query: makes no HIP call apart from reading device properties. */
int ipt_variant_feasible(int variant, int elem_bytes, int64_t d1, int64_t d2, int64_t d3);

/* Plan inspection: the decomposition parameters the library uses for a view of
 * R rows x C columns (c = gcd, a = R/c, b = C/c, k = a^-1 mod b,
 * q = b^-1 mod a), and the variant the heuristic picks for (d1, d2, d3). */
typedef struct {
    int64_t R, C, c, a, b, k, q;
    int32_t heuristic_variant;
    int32_t index_bits;   /* 32 or 64: width of the in-slice index plan; 0 when no
                             exact plan exists (a view with b = C / c above ~3e9:
                             the decomposition then returns IPT_ERR_TOO_LARGE) */
} ipt_plan_info;
ipt_status ipt_plan_query(int elem_bytes, int64_t d1, int64_t d2, int64_t d3,
                          ipt_plan_info* out);

/* Host evaluation of the library's own index maps (the code the kernels run),
 * for checking them against the oracle without a GPU.  map_id: 0 c2r pre,
 * 1 c2r post, 2 r2c pre, 3 r2c post (column gathers: source row), 4 c2r
 * shuffle, 5 r2c shuffle (row gathers: source column).  index_bits selects the
 * 32- or 64-bit plan (32 only where the plan allows it).  Writes n results
 * for the points (i[t], j[t]).  Pure host code.  IPT_ERR_TOO_LARGE at 64 bits
 * for a view without an exact plan (index_bits 0 in ipt_plan_query). */
ipt_status ipt_debug_eval_map(int map_id, int64_t R, int64_t C, int index_bits,
                              const int64_t* i, const int64_t* j, int64_t n,
                              int64_t* out);

/* Run one pass of the decomposition on d3 slices viewed as R rows x C
 * columns (map_id as for ipt_debug_eval_map: column gathers 0-3, row gathers
 * 4-5).  Used for per-pass parity tests and per-kernel timing.  The
 * reference's pass dispatchers this exposes: detail::rotate /
 * detail::col_shuffle_fn (src/rotate.cu:237-270, src/col_shuffle.cu:30-42)
 * and detail::shuffle_fn (src/shuffle.cu:66-76). */
ipt_status ipt_debug_run_pass(void* data, int elem_bytes, int64_t R, int64_t C, int64_t d3,
                              int map_id, const ipt_tune* tune, void* stream);

/* Plain device copy of `bytes` (16-byte aligned pointers and size): the
 * calibration reference for bandwidth and for the HBM traffic counters. */
ipt_status ipt_debug_copy(void* dst, const void* src, size_t bytes, void* stream);

/* Out-of-place transpose: dst receives the 213 transpose of src (same
 * dims and layout conventions as ipt_transpose_213; the two buffers must not
 * overlap).  The comparator the reference times with cuTT's out-of-place
 * plan (src/test_cutt.cu:18-41): it moves each byte once in each direction
 * with no in-place constraint, the practical ceiling for the in-place rates.
 * Tiled through LDS, coalesced on both sides. */
ipt_status ipt_transpose_oop(void* dst, const void* src, int elem_bytes, int64_t d1, int64_t d2,
                             int64_t d3, void* stream);

/* Test hook for the cooperative kernels' barrier watchdog on the current
 * device: every group barrier of later standalone launches gives up after
 * `polls` polls (0: the default, ~2^26) and waits for `extra` arrivals beyond
 * its group (0 in use), so a test can trip the watchdog without a real
 * fault; ipt_device_check then reports it and clears both settings.
 * Synchronises the device. */
ipt_status ipt_debug_set_watchdog(uint32_t polls, uint32_t extra);

/* Test hooks for the library's per-device arenas (which: 0 barrier state,
 * 1 scratch): the current buffer, its size and how many older buffers are
 * retired (kept allocated because a stream capture used them, so graphs
 * captured earlier stay valid after the arena grows); and a forced growth to
 * at least `bytes` (outside stream capture; synchronises the device). */
ipt_status ipt_debug_arena_info(int which, void** ptr, size_t* bytes, int64_t* retired);
ipt_status ipt_debug_arena_grow(int which, size_t bytes);

/* Test hook for the chunk step's move lists (the cycles of the rc x cc
 * chunk-matrix transpose, src/permute.cu:18-49 in the reference, which walks
 * them on the host): builds them with segments of at most `ell` moves
 * (1..128; rc * cc < 2^23) by the library's device builder (builder 1) or
 * by the host walk (builder 0), and copies them out: items as (first move,
 * moves) pairs, the moves' sources (a saved head's slot with the top bit
 * set) and destinations, the heads' positions.  counts = (items, moves,
 * heads); IPT_ERR_TOO_LARGE, with counts set, when one exceeds cap. */
ipt_status ipt_debug_chunk_moves(int64_t rc, int64_t cc, int64_t ell, int builder,
                                 int64_t counts[3], uint32_t* items, uint32_t* msrc,
                                 uint32_t* mdst, uint32_t* hpos, int64_t cap);

/* Test hook for the cooperative kernels' group barrier: for a d2 x d1 slice
 * of elem_bytes elements held by g members of K tiles each, in the banded
 * tile order (banded != 0, bands as tall as a workgroup's 8 waves) or the
 * strided one, writes for every member m < g the range lo[m]..hi[m] of
 * members it waits for before its stores (those whose staged tiles its
 * stores may overwrite; lo > hi when it owns no tile).  Pure host code (the
 * function the kernels run).  No counterpart in the reference, whose kernels
 * synchronise the whole grid. */
ipt_status ipt_debug_gcs_deps(int elem_bytes, int64_t d1, int64_t d2, uint32_t g, uint32_t K,
                              int banded, uint32_t* lo, uint32_t* hi);

/* ---- autotuner and learned selector ----------------------------------- */
/* No counterpart in the reference: its tier thresholds are magic numbers
 * (src/shuffle.cu:15-50, src/rotate.cu:241-258).  Here IPT_NUM_CANDIDATES
 * fixed configurations (variant + knobs, listed by ipt_candidate) are timed by
 * the autotuner; a small MLP (IPT_MLP_IN -> IPT_MLP_HIDDEN -> IPT_MLP_HIDDEN ->
 * IPT_NUM_CANDIDATES, fp32) trained on those timings predicts the log2 time
 * (microseconds) of each candidate, and the dispatcher runs the feasible
 * candidate with the lowest prediction, for problems up to 1 GiB of payload
 * (the domain of its training sweep).  Without a loaded selector, and above
 * that size, the plain heuristic decides. */
#define IPT_NUM_CANDIDATES 13
#define IPT_MLP_IN 8
#define IPT_MLP_HIDDEN 64
/* Number of fp32 parameters: W1, b1, W2, b2, W3, b3 (row-major, out x in). */
#define IPT_MLP_PARAMS (IPT_MLP_HIDDEN * IPT_MLP_IN + IPT_MLP_HIDDEN + \
                        IPT_MLP_HIDDEN * IPT_MLP_HIDDEN + IPT_MLP_HIDDEN + \
                        IPT_NUM_CANDIDATES * IPT_MLP_HIDDEN + IPT_NUM_CANDIDATES)

/* Candidate k: its variant and knobs. */
ipt_status ipt_candidate(int k, int* variant, ipt_tune* tune);
/* Features of a shape (IPT_MLP_IN floats; host only). */
ipt_status ipt_shape_features(int elem_bytes, int64_t d1, int64_t d2, int64_t d3, float* out);
/* The candidate the plain heuristic picks for a shape (host only). */
int ipt_heuristic_candidate(int elem_bytes, int64_t d1, int64_t d2, int64_t d3);

/* Autotuner: time every candidate on `data` (contents are permuted; any
 * values will do), one warm-up run then `reps` timed runs each, HIP events
 * on an internal stream, after a device synchronise (blocking call).  times_us[k] is the mean time in microseconds, or a
 * negative value when candidate k is infeasible for the shape.  *best is the
 * fastest candidate. */
ipt_status ipt_autotune(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3, int reps,
                        float* times_us, int* best);

/* Measured choices: every successful ipt_autotune records the fastest
 * candidate for its exact shape and for its shape class (element size,
 * half-octave buckets of the rows d2 and columns d1, octave bucket of d3,
 * square or not); IPT_VARIANT_AUTO and batches use a recorded choice before
 * the selector, the exact shape's first, else the class's where that
 * candidate can run the shape.  The table can be saved and loaded as text
 * ("elem_bytes d1 d2 d3 candidate" and "class elem_bytes rows_bucket
 * cols_bucket d3_bucket square candidate" lines; IPT_ERR_IO on a bad file,
 * which then changes nothing). */
ipt_status ipt_tuning_clear(void);
ipt_status ipt_tuning_save(const char* path);
ipt_status ipt_tuning_load(const char* path);

/* Train the MLP on the GPU with the library's HIP kernels (full-batch Adam on
 * masked squared error of log2 times).  x: n x IPT_MLP_IN features; y: n x
 * IPT_NUM_CANDIDATES log2 times; mask: n x IPT_NUM_CANDIDATES, 1 where y is
 * valid.  params: IPT_MLP_PARAMS floats, initial values in, trained values out.
 * loss_out (may be NULL): final masked MSE. */
ipt_status ipt_selector_train(const float* x, const float* y, const float* mask, int64_t n,
                              int epochs, float lr, float* params, float* loss_out);
/* One forward + backward pass on the GPU (for gradient checks): the masked
 * MSE loss and its gradient with respect to every parameter. */
ipt_status ipt_selector_grad(const float* x, const float* y, const float* mask, int64_t n,
                             const float* params, float* loss_out, float* grad_out);
/* GPU inference kernel: y = MLP(x) for n rows (n x IPT_NUM_CANDIDATES out). */
ipt_status ipt_selector_infer(const float* params, const float* x, int64_t n, float* y);
/* The same inference on the host (bit-identical to the GPU kernel: both
 * accumulate with fused multiply-adds in the same order). */
ipt_status ipt_selector_infer_host(const float* params, const float* x, int64_t n, float* y);
/* Install parameters for dispatch (copied; NULL uninstalls and restores the
 * heuristic).  Affects IPT_VARIANT_AUTO in every entry point and batches
 * created afterwards.  Until this is called, the library uses the trained
 * weights shipped with it (selector/weights.f32 next to the shared object,
 * or the file named by $IPT_SELECTOR_WEIGHTS); IPT_SELECTOR=heuristic
 * disables them, and a missing or corrupt file leaves the heuristic. */
ipt_status ipt_selector_load(const float* params);
/* 1 when a selector decides IPT_VARIANT_AUTO, 0 when the heuristic does. */
int ipt_selector_active(void);
/* The candidate the dispatcher would run for a shape now (selector if
 * loaded, else heuristic); host only. */
int ipt_select_candidate(int elem_bytes, int64_t d1, int64_t d2, int64_t d3);

/* ---- batches of uneven transposes ------------------------------------- */
/* No counterpart in the reference (one tensor per call, src/transpose.cu:
 * 99-124).  A batch is planned once: every instance becomes work items of one
 * persistent kernel (whole-slice LDS, cooperative group, or square tile-pair
 * items).  An instance too large for those takes the block-cycle path when
 * it can: its blocks become items of the persistent kernel, and its chunk
 * steps run just before (chunked columns) or after (chunked rows) that kernel,
 * on move lists and a head buffer the batch owns.  The rest run through the
 * per-call C2R / R2C launchers (only their LDS tiers: a batch never uses
 * library scratch).
 * ipt_batch_run captures everything once into a hipGraph on the first call
 * and replays it afterwards.  Instance buffers must stay valid and must not
 * overlap; a batch must not run concurrently with itself, nor with another
 * batch or a cooperative (GCS) transpose on another stream: the persistent
 * kernel's groups need every workgroup of its grid resident, and only
 * stream-ordered calls into this library are ordered against it (an
 * overlap is caught by the barrier watchdog, ipt_batch_check). */
typedef struct {
    void* data;          /* device buffer of d1*d2*d3 elements */
    int32_t elem_bytes;  /* 4 or 8 */
    int32_t variant;     /* IPT_VARIANT_AUTO or a forced variant */
    int64_t d1, d2, d3;  /* as ipt_transpose_213 */
} ipt_matrix_desc;

typedef struct {
    /* n_big: instances the per-call launchers run after the persistent
     * kernel (IPT_ROUTE_GCS_LAUNCH and IPT_ROUTE_SEPARATE) */
    int64_t count, skipped, n_instances, n_big;
    int64_t n_slice_items, n_gcs_items, n_swap_items;
    int64_t grid, lds_bytes;
    uint64_t bytes;      /* payload bytes of the transposed instances */
    int64_t n_block_cycle; /* instances on the block-cycle path */
} ipt_batch_info;

typedef struct ipt_batch_s ipt_batch;

/* tune may be NULL; tune->lds_bytes caps the persistent kernel's LDS. */
ipt_status ipt_batch_create(const ipt_matrix_desc* descs, int64_t count, const ipt_tune* tune,
                            ipt_batch** out);
/* One graph launch on `stream` (captured on the first call). */
ipt_status ipt_batch_run(ipt_batch* b, void* stream);
/* The same work launched directly, without a graph. */
ipt_status ipt_batch_run_eager(ipt_batch* b, void* stream);
ipt_status ipt_batch_get_info(const ipt_batch* b, ipt_batch_info* info);

/* Route each descriptor of the batch took (descriptor order), for
 * breaking a batch's time down by route: */
typedef enum {
    IPT_ROUTE_SKIP = 0,        /* one row or one column: nothing to move */
    IPT_ROUTE_SLICE = 1,       /* whole slices in LDS (persistent kernel) */
    IPT_ROUTE_GCS = 2,         /* cooperative groups (persistent kernel) */
    IPT_ROUTE_SWAP = 3,        /* square tile pairs (persistent kernel) */
    IPT_ROUTE_BLOCK_CYCLE = 4, /* blocks in the persistent kernel + chunk step */
    IPT_ROUTE_SEPARATE = 5,    /* per-call launchers (C2R / R2C / others) */
    IPT_ROUTE_GCS_LAUNCH = 6   /* one separate cooperative launch (the per-call
                                  kernel: too large for one group of the
                                  persistent kernel, not for the per-call one) */
} ipt_batch_route;
/* routes: `count` int32 values (the batch's descriptor count). */
ipt_status ipt_batch_get_routes(const ipt_batch* b, int32_t* routes);
/* Measurement: with timing on, every ipt_batch_run_eager has its persistent
 * kernel's launch record two events around it on the run's stream;
 * ipt_batch_kernel_ms waits for the last such run and gives that kernel's
 * duration in milliseconds (IPT_ERR_NOT_READY before any timed run).  Graph
 * runs are never timed this way. */
ipt_status ipt_batch_set_timing(ipt_batch* b, int on);
ipt_status ipt_batch_kernel_ms(ipt_batch* b, float* ms);
/* Synchronises the device and reports (IPT_ERR_HIP) whether a group barrier
 * of this batch gave up waiting; the barrier state is then reset. */
ipt_status ipt_batch_check(ipt_batch* b);
/* Test hook for this batch's group barriers (as ipt_debug_set_watchdog for
 * standalone launches): in later runs every member gives up after `polls`
 * polls (0: the default, ~2^26) and also waits for `extra` arrivals that
 * never come (0 in use), so a test can trip the watchdog without a real
 * fault; ipt_batch_check then reports it and clears both settings.
 * Synchronises the device. */
ipt_status ipt_batch_debug_set_watchdog(ipt_batch* b, uint32_t polls, uint32_t extra);
/* Synchronises the device, then frees the batch. */
ipt_status ipt_batch_destroy(ipt_batch* b);

/* Synchronises the current device and reports whether any group barrier of
 * the cooperative kernels gave up waiting (IPT_ERR_HIP, with detail in
 * ipt_last_error; the barrier state is then reset).  IPT_OK otherwise. */
ipt_status ipt_device_check(void);

/* Library build / device description, e.g. "gfx950 lds=163840 cus=256". */
const char* ipt_version(void);

#ifdef __cplusplus
}
#endif
#endif
