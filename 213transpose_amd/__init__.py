"""MI355X (gfx950) in-place transpose engine — Python host mirror.

The compute path is libinplace_hip.so (hand-written HIP kernels behind a C ABI,
declared in include/inplace_transpose.h).  This module binds that ABI with
ctypes and mirrors the reference's interface (include/transpose.h):

    transpose(data, d1, d2, d3)        inplace::transpose<T>      (src/transpose.cu:99-124)
    c2r.transpose(data, d1, d2, d3)    inplace::c2r::transpose<T> (src/transpose.cu:30-47)
    r2c.transpose(data, d1, d2, d3)    inplace::r2c::transpose<T> (src/transpose.cu:70-89)
    openmp_style(row_major, data, m, n) inplace::openmp::transpose (include/openmp.h:7-8), GPU
    cpu_transpose(row_major, a, m, n)   inplace::openmp::transpose on the host CPU

`data` is a contiguous CUDA (HIP) tensor of 4- or 8-byte elements holding d3
slices of d2 rows x d1 columns; every slice is replaced by its transpose.  Work
is enqueued on torch's current stream.  A pinned host tensor is transposed in
place by the GPU over the host link (synchronise before reading it); managed
memory, reached through the C ABI, is prefetched to the device first, as the
reference does (src/transpose.cu:100-104).  There is no CPU fallback: if the
library is missing or a call fails, a TransposeError is raised.

The package directory name starts with a digit, so import it with
importlib.import_module("213transpose_amd").
"""
import ctypes
import os
import types

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libinplace_hip.so")

VARIANT_AUTO, VARIANT_SLICE_LDS, VARIANT_SQUARE_SWAP, VARIANT_C2R, VARIANT_R2C, VARIANT_GCS, \
    VARIANT_BLOCK_CYCLE = range(7)
VARIANT_NAMES = {VARIANT_AUTO: "auto", VARIANT_SLICE_LDS: "slice_lds",
                 VARIANT_SQUARE_SWAP: "square_swap", VARIANT_C2R: "c2r", VARIANT_R2C: "r2c",
                 VARIANT_GCS: "gcs", VARIANT_BLOCK_CYCLE: "block_cycle"}
N_VARIANTS = 7

# Map ids of ipt_debug_eval_map.
MAP_C2R_PRE, MAP_C2R_POST, MAP_R2C_PRE, MAP_R2C_POST, MAP_C2R_SHUFFLE, MAP_R2C_SHUFFLE = range(6)

STATUS = {0: "IPT_OK", 1: "IPT_ERR_INVALID_ARG", 2: "IPT_ERR_UNSUPPORTED", 3: "IPT_ERR_TOO_LARGE",
          4: "IPT_ERR_HIP", 5: "IPT_ERR_NO_DEVICE", 6: "IPT_ERR_OUT_OF_MEMORY", 7: "IPT_ERR_IO",
          8: "IPT_ERR_NOT_READY"}


class TransposeError(RuntimeError):
    def __init__(self, status, what, detail=""):
        self.status = status
        msg = f"{what}: {STATUS.get(status, status)}"
        if detail:
            msg += f" ({detail})"
        super().__init__(msg)


class Tune(ctypes.Structure):
    """ipt_tune: zero fields mean library default."""
    _fields_ = [("lds_bytes", ctypes.c_int32), ("block", ctypes.c_int32),
                ("strip", ctypes.c_int32), ("grid_mult", ctypes.c_int32),
                ("chunk_min", ctypes.c_int32), ("chunk_max", ctypes.c_int32),
                ("tile", ctypes.c_int32)]


class PlanInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("R", "C", "c", "a", "b", "k", "q")] + \
               [("heuristic_variant", ctypes.c_int32), ("index_bits", ctypes.c_int32)]


class MatrixDesc(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("elem_bytes", ctypes.c_int32),
                ("variant", ctypes.c_int32), ("d1", ctypes.c_int64), ("d2", ctypes.c_int64),
                ("d3", ctypes.c_int64)]


class BatchInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("count", "skipped", "n_instances", "n_big",
                                               "n_slice_items", "n_gcs_items", "n_swap_items",
                                               "grid", "lds_bytes")] + [
        ("bytes", ctypes.c_uint64), ("n_block_cycle", ctypes.c_int64)]


_i64, _vp, _int = ctypes.c_int64, ctypes.c_void_p, ctypes.c_int
_lib = None


def lib():
    """Load libinplace_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    # torch wheels bundle their own HIP runtime under the same soname as the
    # system one (libamdhip64.so.7); whichever is loaded first serves the
    # whole process.  Load torch's first so a process that uses both has one
    # runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise TransposeError(-1, "libinplace_hip.so not built",
                             f"expected {LIB_PATH}; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    L.ipt_status_string.restype = ctypes.c_char_p
    L.ipt_status_string.argtypes = [_int]
    L.ipt_last_error.restype = ctypes.c_char_p
    L.ipt_last_error.argtypes = []
    for name in ("ipt_transpose_213", "ipt_c2r_transpose", "ipt_r2c_transpose",
                 "ipt_transpose_213_noalloc"):
        f = getattr(L, name)
        f.restype = _int
        f.argtypes = [_vp, _int, _i64, _i64, _i64, _vp]
    L.ipt_transpose.restype = _int
    L.ipt_transpose.argtypes = [_int, _vp, _int, _i64, _i64, _vp]
    L.ipt_transpose_variant.restype = _int
    L.ipt_transpose_variant.argtypes = [_vp, _int, _i64, _i64, _i64, _int,
                                        ctypes.POINTER(Tune), _vp]
    L.ipt_variant_feasible.restype = _int
    L.ipt_variant_feasible.argtypes = [_int, _int, _i64, _i64, _i64]
    L.ipt_plan_query.restype = _int
    L.ipt_plan_query.argtypes = [_int, _i64, _i64, _i64, ctypes.POINTER(PlanInfo)]
    L.ipt_debug_eval_map.restype = _int
    L.ipt_debug_eval_map.argtypes = [_int, _i64, _i64, _int, _vp, _vp, _i64, _vp]
    L.ipt_debug_run_pass.restype = _int
    L.ipt_debug_run_pass.argtypes = [_vp, _int, _i64, _i64, _i64, _int, ctypes.POINTER(Tune), _vp]
    L.ipt_debug_copy.restype = _int
    L.ipt_permute.restype = _int
    L.ipt_permute.argtypes = [_vp, _int, _int, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int), _vp]
    L.ipt_permute3.restype = _int
    L.ipt_permute3.argtypes = [_vp, _int, _i64, _i64, _i64, ctypes.POINTER(ctypes.c_int), _vp]
    L.ipt_debug_set_watchdog.restype = _int
    L.ipt_debug_set_watchdog.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    L.ipt_debug_copy.argtypes = [_vp, _vp, ctypes.c_size_t, _vp]
    L.ipt_device_check.restype = _int
    L.ipt_device_check.argtypes = []
    L.ipt_batch_create.restype = _int
    L.ipt_batch_create.argtypes = [ctypes.POINTER(MatrixDesc), _i64, ctypes.POINTER(Tune),
                                   ctypes.POINTER(_vp)]
    for name in ("ipt_batch_run", "ipt_batch_run_eager"):
        getattr(L, name).restype = _int
        getattr(L, name).argtypes = [_vp, _vp]
    L.ipt_batch_get_info.restype = _int
    L.ipt_batch_get_info.argtypes = [_vp, ctypes.POINTER(BatchInfo)]
    L.ipt_batch_set_timing.restype = _int
    L.ipt_batch_set_timing.argtypes = [_vp, _int]
    L.ipt_batch_kernel_ms.restype = _int
    L.ipt_batch_kernel_ms.argtypes = [_vp, ctypes.POINTER(ctypes.c_float)]
    L.ipt_batch_get_routes.restype = _int
    L.ipt_batch_get_routes.argtypes = [_vp, _vp]
    L.ipt_batch_check.restype = _int
    L.ipt_batch_check.argtypes = [_vp]
    L.ipt_batch_debug_set_watchdog.restype = _int
    L.ipt_batch_debug_set_watchdog.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32]
    L.ipt_batch_destroy.restype = _int
    L.ipt_batch_destroy.argtypes = [_vp]
    L.ipt_candidate.restype = _int
    L.ipt_candidate.argtypes = [_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(Tune)]
    L.ipt_shape_features.restype = _int
    L.ipt_shape_features.argtypes = [_int, _i64, _i64, _i64, _vp]
    for name in ("ipt_heuristic_candidate", "ipt_select_candidate"):
        getattr(L, name).restype = _int
        getattr(L, name).argtypes = [_int, _i64, _i64, _i64]
    L.ipt_autotune.restype = _int
    L.ipt_autotune.argtypes = [_vp, _int, _i64, _i64, _i64, _int, _vp, ctypes.POINTER(ctypes.c_int)]
    L.ipt_selector_train.restype = _int
    L.ipt_selector_train.argtypes = [_vp, _vp, _vp, _i64, _int, ctypes.c_float, _vp, _vp]
    L.ipt_selector_grad.restype = _int
    L.ipt_selector_grad.argtypes = [_vp, _vp, _vp, _i64, _vp, _vp, _vp]
    L.ipt_selector_infer.restype = _int
    L.ipt_selector_infer.argtypes = [_vp, _vp, _i64, _vp]
    L.ipt_selector_infer_host.restype = _int
    L.ipt_selector_infer_host.argtypes = [_vp, _vp, _i64, _vp]
    for name in ("ipt_tuning_save", "ipt_tuning_load"):
        getattr(L, name).restype = _int
        getattr(L, name).argtypes = [ctypes.c_char_p]
    L.ipt_tuning_clear.restype = _int
    L.ipt_tuning_clear.argtypes = []
    L.ipt_selector_active.restype = _int
    L.ipt_selector_active.argtypes = []
    L.ipt_selector_load.restype = _int
    L.ipt_selector_load.argtypes = [_vp]
    L.ipt_cpu_transpose.restype = _int
    L.ipt_cpu_transpose.argtypes = [_int, _vp, _int, _i64, _i64, _vp, _int]
    L.ipt_cpu_transpose_213.restype = _int
    L.ipt_cpu_transpose_213.argtypes = [_vp, _int, _i64, _i64, _i64, _vp, _int]
    L.ipt_cpu_scratch_bytes.restype = _i64
    L.ipt_cpu_scratch_bytes.argtypes = [_int, _i64, _i64, _i64, _int]
    L.ipt_debug_arena_info.restype = _int
    L.ipt_debug_arena_info.argtypes = [_int, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(_i64)]
    L.ipt_debug_arena_grow.restype = _int
    L.ipt_debug_arena_grow.argtypes = [_int, ctypes.c_size_t]
    L.ipt_debug_chunk_moves.restype = _int
    L.ipt_debug_chunk_moves.argtypes = [_i64, _i64, _i64, _int, _vp, _vp, _vp, _vp, _vp, _i64]
    L.ipt_transpose_oop.restype = _int
    L.ipt_transpose_oop.argtypes = [_vp, _vp, _int, _i64, _i64, _i64, _vp]
    L.ipt_debug_gcs_deps.restype = _int
    L.ipt_debug_gcs_deps.argtypes = [_int, _i64, _i64, ctypes.c_uint32, ctypes.c_uint32, _int, _vp, _vp]
    L.ipt_memory_kind.restype = _int
    L.ipt_memory_kind.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
    L.ipt_version.restype = ctypes.c_char_p
    L.ipt_version.argtypes = []
    _lib = L
    return L


def _check(status, what):
    if status != 0:
        detail = lib().ipt_last_error().decode(errors="replace")
        raise TransposeError(status, what, detail)


def _stream_ptr(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _args(data, d1, d2, d3):
    if not (data.is_cuda or data.is_pinned()):
        raise TransposeError(1, "data must be a GPU tensor or pinned host memory (no CPU path)")
    if not data.is_contiguous():
        raise TransposeError(1, "data must be contiguous")
    esz = data.element_size()
    if esz not in (4, 8):
        raise TransposeError(1, f"element size {esz} not supported (4 or 8 bytes)")
    if d1 * d2 * d3 != data.numel():
        raise TransposeError(1, f"d1*d2*d3 = {d1 * d2 * d3} != numel {data.numel()}")
    return ctypes.c_void_p(data.data_ptr()), esz


# ipt_memory_kind_t: where a pointer lives (hipMemoryType's values).
MEM_PAGEABLE, MEM_HOST, MEM_DEVICE, MEM_MANAGED = 0, 1, 2, 3
MEM_NAMES = {MEM_PAGEABLE: "pageable", MEM_HOST: "host", MEM_DEVICE: "device", MEM_MANAGED: "managed"}


def memory_kind(ptr):
    """ipt_memory_kind: (kind, owning device or -1) of a pointer or tensor."""
    if hasattr(ptr, "data_ptr"):
        ptr = ptr.data_ptr()
    dev = ctypes.c_int(-1)
    k = lib().ipt_memory_kind(ctypes.c_void_p(ptr), ctypes.byref(dev))
    return k, dev.value


def transpose(data, d1, d2, d3=1, stream=None):
    """inplace::transpose<T>: transpose each of d3 slices (d2 rows x d1 cols) in place."""
    ptr, esz = _args(data, d1, d2, d3)
    _check(lib().ipt_transpose_213(ptr, esz, d1, d2, d3, _stream_ptr(stream)), "ipt_transpose_213")
    return data


def transpose_noalloc(data, d1, d2, d3=1, stream=None):
    """ipt_transpose_213_noalloc: as transpose(), allocating nothing (scratch
    in the library's static device array; overlapping calls on one device race)."""
    ptr, esz = _args(data, d1, d2, d3)
    _check(lib().ipt_transpose_213_noalloc(ptr, esz, d1, d2, d3, _stream_ptr(stream)),
           "ipt_transpose_213_noalloc")
    return data


def _c2r(data, d1, d2, d3=1, stream=None):
    ptr, esz = _args(data, d1, d2, d3)
    _check(lib().ipt_c2r_transpose(ptr, esz, d1, d2, d3, _stream_ptr(stream)), "ipt_c2r_transpose")
    return data


def _r2c(data, d1, d2, d3=1, stream=None):
    ptr, esz = _args(data, d1, d2, d3)
    _check(lib().ipt_r2c_transpose(ptr, esz, d1, d2, d3, _stream_ptr(stream)), "ipt_r2c_transpose")
    return data


# inplace::c2r::transpose: slices of d2 rows x d1 cols.
c2r = types.SimpleNamespace(transpose=_c2r)
# inplace::r2c::transpose: slices of d1 rows x d2 cols (reference argument order).
r2c = types.SimpleNamespace(transpose=_r2c)


def permute3(data, dims, perm, stream=None):
    """ipt_permute3: any permutation of a rank-3 tensor in place (dims fastest
    first; output dim j is source dim perm[j], as the oracle's seq_tt)."""
    d1, d2, d3 = dims
    ptr, esz = _args(data, d1, d2, d3)
    p = (ctypes.c_int * 3)(*perm)
    _check(lib().ipt_permute3(ptr, esz, d1, d2, d3, p, _stream_ptr(stream)), "ipt_permute3")
    return data


def permute(data, dims, perm, stream=None):
    """ipt_permute: any permutation of a tensor of rank <= 6 in place (dims
    fastest first; output dim j is source dim perm[j], as the oracle's seq_tt)."""
    total = 1
    for d in dims:
        total *= d
    ptr, esz = _args(data, total, 1, 1)
    d = (ctypes.c_int64 * len(dims))(*dims)
    p = (ctypes.c_int * len(perm))(*perm)
    _check(lib().ipt_permute(ptr, esz, len(dims), d, p, _stream_ptr(stream)), "ipt_permute")
    return data


def openmp_style(row_major, data, m, n, stream=None):
    """inplace::openmp::transpose(row_major, data, m, n, tmp) on the GPU."""
    ptr, esz = _args(data, m, n, 1)
    _check(lib().ipt_transpose(int(bool(row_major)), ptr, esz, m, n, _stream_ptr(stream)),
           "ipt_transpose")
    return data


def _host_args(a, n):
    import numpy as np
    if not isinstance(a, np.ndarray) or not a.flags["C_CONTIGUOUS"]:
        raise TransposeError(1, "cpu path: data must be a contiguous numpy array")
    if a.itemsize not in (4, 8):
        raise TransposeError(1, f"element size {a.itemsize} not supported (4 or 8 bytes)")
    if a.size != n:
        raise TransposeError(1, f"dims give {n} elements, array has {a.size}")
    return ctypes.c_void_p(a.ctypes.data), a.itemsize


def cpu_transpose(row_major, a, m, n, threads=0, tmp=None):
    """inplace::openmp::transpose(row_major, data, m, n, tmp) on the host
    (ipt_cpu_transpose): `a` is a numpy array of m * n 4- or 8-byte
    elements, transposed in place; `tmp` optional caller scratch (a numpy
    array of at least cpu_scratch_bytes(...) bytes)."""
    ptr, esz = _host_args(a, m * n)
    t = ctypes.c_void_p(tmp.ctypes.data) if tmp is not None else None
    _check(lib().ipt_cpu_transpose(int(bool(row_major)), ptr, esz, m, n, t, threads),
           "ipt_cpu_transpose")
    return a


def cpu_transpose_213(a, d1, d2, d3=1, threads=0, tmp=None):
    """The 213 transpose on the host (ipt_cpu_transpose_213)."""
    ptr, esz = _host_args(a, d1 * d2 * d3)
    t = ctypes.c_void_p(tmp.ctypes.data) if tmp is not None else None
    _check(lib().ipt_cpu_transpose_213(ptr, esz, d1, d2, d3, t, threads), "ipt_cpu_transpose_213")
    return a


def cpu_scratch_bytes(esz, d1, d2, d3=1, threads=0):
    return lib().ipt_cpu_scratch_bytes(esz, d1, d2, d3, threads)


def transpose_variant(data, d1, d2, d3, variant, tune=None, stream=None):
    """Force one variant (and optional Tune knobs)."""
    ptr, esz = _args(data, d1, d2, d3)
    t = ctypes.byref(tune) if tune is not None else None
    _check(lib().ipt_transpose_variant(ptr, esz, d1, d2, d3, variant, t, _stream_ptr(stream)),
           f"ipt_transpose_variant({VARIANT_NAMES.get(variant, variant)})")
    return data


def feasible(variant, esz, d1, d2, d3=1):
    return bool(lib().ipt_variant_feasible(variant, esz, d1, d2, d3))


def plan_query(esz, d1, d2, d3=1):
    info = PlanInfo()
    _check(lib().ipt_plan_query(esz, d1, d2, d3, ctypes.byref(info)), "ipt_plan_query")
    return {f: getattr(info, f) for f, _ in PlanInfo._fields_}


def eval_map(map_id, R, C, index_bits, i, j):
    """Host evaluation of the library's index map (numpy int64 arrays in/out)."""
    import numpy as np
    i = np.ascontiguousarray(i, dtype=np.int64)
    j = np.ascontiguousarray(j, dtype=np.int64)
    out = np.empty_like(i)
    _check(lib().ipt_debug_eval_map(map_id, R, C, index_bits, i.ctypes.data, j.ctypes.data,
                                    i.size, out.ctypes.data), "ipt_debug_eval_map")
    return out


def run_pass(data, R, C, d3, map_id, tune=None, stream=None):
    """One gather pass on d3 slices viewed as R x C (per-pass parity / timing)."""
    ptr, esz = _args(data, C, R, d3)
    t = ctypes.byref(tune) if tune is not None else None
    _check(lib().ipt_debug_run_pass(ptr, esz, R, C, d3, map_id, t, _stream_ptr(stream)),
           f"ipt_debug_run_pass({map_id})")
    return data


def copy(dst, src, stream=None):
    """Plain device copy src -> dst (calibration reference)."""
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() < nbytes:
        raise TransposeError(1, "copy: destination too small")
    _check(lib().ipt_debug_copy(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()),
                                nbytes, _stream_ptr(stream)), "ipt_debug_copy")
    return dst


# ipt_batch_route: how a batch runs each item.
(ROUTE_SKIP, ROUTE_SLICE, ROUTE_GCS, ROUTE_SWAP, ROUTE_BLOCK_CYCLE, ROUTE_SEPARATE,
 ROUTE_GCS_LAUNCH) = range(7)
ROUTE_NAMES = {ROUTE_SKIP: "skip", ROUTE_SLICE: "slice", ROUTE_GCS: "gcs", ROUTE_SWAP: "swap",
               ROUTE_BLOCK_CYCLE: "block_cycle", ROUTE_SEPARATE: "separate",
               ROUTE_GCS_LAUNCH: "gcs_launch"}


def transpose_oop(dst, src, d1, d2, d3=1, stream=None):
    """Out-of-place comparator (ipt_transpose_oop): dst = 213 transpose of src."""
    sptr, esz = _args(src, d1, d2, d3)
    dptr, esz2 = _args(dst, d1, d2, d3)
    if esz != esz2:
        raise TransposeError(1, "transpose_oop: element sizes differ")
    _check(lib().ipt_transpose_oop(dptr, sptr, esz, d1, d2, d3, _stream_ptr(stream)),
           "ipt_transpose_oop")
    return dst


class Batch:
    """A planned batch of uneven in-place transposes (ipt_batch_*).

    items: iterable of (tensor, d1, d2, d3) or (tensor, d1, d2, d3, variant);
    every tensor holds d3 slices of d2 x d1 and must stay alive while the
    batch exists.  run() replays one captured hipGraph; run(graph=False)
    launches the same work directly.
    """

    def __init__(self, items, tune=None):
        items = list(items)
        self._tensors = [it[0] for it in items]
        descs = (MatrixDesc * max(1, len(items)))()
        for k, it in enumerate(items):
            t, d1, d2, d3 = it[:4]
            ptr, esz = _args(t, d1, d2, d3)
            descs[k] = MatrixDesc(ptr.value, esz, it[4] if len(it) > 4 else VARIANT_AUTO, d1, d2, d3)
        self._h = ctypes.c_void_p()
        self._count = len(items)
        tptr = ctypes.byref(tune) if tune is not None else None
        _check(lib().ipt_batch_create(descs, len(items), tptr, ctypes.byref(self._h)),
               "ipt_batch_create")

    def run(self, graph=True, stream=None):
        f = lib().ipt_batch_run if graph else lib().ipt_batch_run_eager
        _check(f(self._h, _stream_ptr(stream)), "ipt_batch_run" if graph else "ipt_batch_run_eager")

    def info(self):
        inf = BatchInfo()
        _check(lib().ipt_batch_get_info(self._h, ctypes.byref(inf)), "ipt_batch_get_info")
        return {f: getattr(inf, f) for f, _ in BatchInfo._fields_}

    def set_timing(self, on=True):
        """Eager runs record events around the persistent kernel."""
        _check(lib().ipt_batch_set_timing(self._h, int(bool(on))), "ipt_batch_set_timing")

    def kernel_ms(self):
        """Duration of the persistent kernel in the last timed eager run (ms)."""
        ms = ctypes.c_float()
        _check(lib().ipt_batch_kernel_ms(self._h, ctypes.byref(ms)), "ipt_batch_kernel_ms")
        return ms.value

    def routes(self):
        """Route of each item (ROUTE_* values, item order)."""
        import numpy as np
        out = np.zeros(max(1, self._count), dtype=np.int32)
        _check(lib().ipt_batch_get_routes(self._h, out.ctypes.data), "ipt_batch_get_routes")
        return out[:self._count]

    def check(self):
        """Synchronise and raise if a group barrier of this batch gave up waiting."""
        _check(lib().ipt_batch_check(self._h), "ipt_batch_check")

    def set_watchdog(self, polls, extra):
        """Test hook: later runs' group barriers give up after `polls` polls and
        wait for `extra` arrivals that never come (0, 0: normal)."""
        _check(lib().ipt_batch_debug_set_watchdog(self._h, polls, extra),
               "ipt_batch_debug_set_watchdog")

    def close(self):
        if self._h:
            _check(lib().ipt_batch_destroy(self._h), "ipt_batch_destroy")
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uneven_shapes(count, lo=2, hi=4096, seed=0):
    """The uneven-batch workload: `count` row-major matrices with m, n drawn
    log-uniform in [lo, hi] and 4- or 8-byte elements (p = 1/2), seeded.
    Returns [(m, n, esz)]; as a 213 problem each is (d1 = n, d2 = m, d3 = 1)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    lm = rng.uniform(np.log(lo), np.log(hi + 1), size=(count, 2))
    mn = np.minimum(np.floor(np.exp(lm)).astype(np.int64), hi)
    esz = np.where(rng.random(count) < 0.5, 4, 8)
    return [(int(m), int(n), int(e)) for (m, n), e in zip(mn, esz)]


# ---------------------------------------------------------------- selector

N_CANDIDATES = 13
MLP_IN, MLP_HIDDEN = 8, 64
MLP_PARAMS = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + \
    N_CANDIDATES * MLP_HIDDEN + N_CANDIDATES
SELECTOR_WEIGHTS = os.path.join(_HERE, "selector", "weights.f32")


def candidate(k):
    """(variant, Tune) of candidate k."""
    v, t = ctypes.c_int(), Tune()
    _check(lib().ipt_candidate(k, ctypes.byref(v), ctypes.byref(t)), "ipt_candidate")
    return v.value, t


def shape_features(esz, d1, d2, d3=1):
    import numpy as np
    out = np.empty(MLP_IN, dtype=np.float32)
    _check(lib().ipt_shape_features(esz, d1, d2, d3, out.ctypes.data), "ipt_shape_features")
    return out


def heuristic_candidate(esz, d1, d2, d3=1):
    return lib().ipt_heuristic_candidate(esz, d1, d2, d3)


def select_candidate(esz, d1, d2, d3=1):
    return lib().ipt_select_candidate(esz, d1, d2, d3)


def autotune(data, d1, d2, d3, reps=3):
    """Time every candidate on `data` (permuted in the process): (times_us, best)."""
    import numpy as np
    ptr, esz = _args(data, d1, d2, d3)
    times = np.empty(N_CANDIDATES, dtype=np.float32)
    best = ctypes.c_int()
    _check(lib().ipt_autotune(ptr, esz, d1, d2, d3, reps, times.ctypes.data, ctypes.byref(best)),
           "ipt_autotune")
    return times, best.value


def _f32(a, shape):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise TransposeError(1, f"expected shape {shape}, got {a.shape}")
    return a


def selector_train(x, y, mask, params, epochs=2000, lr=3e-3):
    """Train on the GPU (the library's HIP kernels); returns (params, final loss)."""
    n = len(x)
    x = _f32(x, (n, MLP_IN))
    y = _f32(y, (n, N_CANDIDATES))
    mask = _f32(mask, (n, N_CANDIDATES))
    p = _f32(params, (MLP_PARAMS,)).copy()
    loss = ctypes.c_float()
    _check(lib().ipt_selector_train(x.ctypes.data, y.ctypes.data, mask.ctypes.data, n, epochs,
                                    lr, p.ctypes.data, ctypes.byref(loss)), "ipt_selector_train")
    return p, loss.value


def selector_grad(x, y, mask, params):
    """Masked-MSE loss and its gradient, computed by the GPU kernels."""
    import numpy as np
    n = len(x)
    x = _f32(x, (n, MLP_IN))
    y = _f32(y, (n, N_CANDIDATES))
    mask = _f32(mask, (n, N_CANDIDATES))
    p = _f32(params, (MLP_PARAMS,))
    g = np.empty(MLP_PARAMS, dtype=np.float32)
    loss = ctypes.c_float()
    _check(lib().ipt_selector_grad(x.ctypes.data, y.ctypes.data, mask.ctypes.data, n,
                                   p.ctypes.data, ctypes.byref(loss), g.ctypes.data),
           "ipt_selector_grad")
    return loss.value, g


def selector_infer(params, x):
    """GPU inference kernel: predicted log2 microseconds per candidate."""
    import numpy as np
    n = len(x)
    x = _f32(x, (n, MLP_IN))
    p = _f32(params, (MLP_PARAMS,))
    y = np.empty((n, N_CANDIDATES), dtype=np.float32)
    _check(lib().ipt_selector_infer(p.ctypes.data, x.ctypes.data, n, y.ctypes.data),
           "ipt_selector_infer")
    return y


def selector_infer_host(params, x):
    """Host inference (bit-identical to selector_infer)."""
    import numpy as np
    n = len(x)
    x = _f32(x, (n, MLP_IN))
    p = _f32(params, (MLP_PARAMS,))
    y = np.empty((n, N_CANDIDATES), dtype=np.float32)
    _check(lib().ipt_selector_infer_host(p.ctypes.data, x.ctypes.data, n, y.ctypes.data),
           "ipt_selector_infer_host")
    return y


def selector_load(params):
    """Install MLP parameters for dispatch (None restores the heuristic)."""
    if params is None:
        _check(lib().ipt_selector_load(None), "ipt_selector_load")
        return
    p = _f32(params, (MLP_PARAMS,))
    _check(lib().ipt_selector_load(p.ctypes.data), "ipt_selector_load")


def tuning_clear():
    _check(lib().ipt_tuning_clear(), "ipt_tuning_clear")


def tuning_save(path):
    _check(lib().ipt_tuning_save(os.fsencode(path)), "ipt_tuning_save")


def tuning_load(path):
    _check(lib().ipt_tuning_load(os.fsencode(path)), "ipt_tuning_load")


def selector_active():
    """True when the learned selector decides VARIANT_AUTO (the library loads
    the shipped weights by default)."""
    return bool(lib().ipt_selector_active())


def load_default_selector():
    """Install the trained weights shipped with the package (if present).
    Returns True when installed; a missing or corrupt file leaves the
    heuristic in place."""
    import numpy as np
    try:
        p = np.fromfile(SELECTOR_WEIGHTS, dtype=np.float32)
    except OSError:
        return False
    if p.size != MLP_PARAMS or not np.isfinite(p).all():
        return False
    selector_load(p)
    return True


def device_check():
    """Synchronise and raise if a cooperative group barrier gave up waiting."""
    _check(lib().ipt_device_check(), "ipt_device_check")


def set_watchdog(polls=0, extra=0):
    """Test hook (ipt_debug_set_watchdog): group barriers give up after
    `polls` polls and wait for `extra` arrivals beyond their group."""
    _check(lib().ipt_debug_set_watchdog(polls, extra), "ipt_debug_set_watchdog")


ARENA_SYNC, ARENA_SCRATCH = 0, 1


def arena_info(which):
    """(pointer, bytes, retired buffers) of a per-device arena (test hook)."""
    p, n, r = ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_int64()
    _check(lib().ipt_debug_arena_info(which, ctypes.byref(p), ctypes.byref(n), ctypes.byref(r)),
           "ipt_debug_arena_info")
    return p.value, n.value, r.value


def arena_grow(which, nbytes):
    """Force a per-device arena to grow to at least nbytes (test hook)."""
    _check(lib().ipt_debug_arena_grow(which, nbytes), "ipt_debug_arena_grow")


def chunk_moves(rc, cc, ell, device=True):
    """The chunk step's move lists of the rc x cc chunk transpose with
    segments of at most ell moves, built by the device builder (or the host
    walk): (items (n, 2) of (first move, moves), msrc, mdst, hpos), uint32
    numpy arrays (test hook)."""
    import numpy as np
    cap = rc * cc
    counts = np.zeros(3, dtype=np.int64)
    items = np.zeros((max(cap, 1), 2), dtype=np.uint32)
    msrc = np.zeros(max(cap, 1), dtype=np.uint32)
    mdst = np.zeros(max(cap, 1), dtype=np.uint32)
    hpos = np.zeros(max(cap, 1), dtype=np.uint32)
    _check(lib().ipt_debug_chunk_moves(rc, cc, ell, 1 if device else 0, counts.ctypes.data,
                                       items.ctypes.data, msrc.ctypes.data, mdst.ctypes.data,
                                       hpos.ctypes.data, cap), "ipt_debug_chunk_moves")
    ni, nm, nh = (int(c) for c in counts)
    return items[:ni].copy(), msrc[:nm].copy(), mdst[:nm].copy(), hpos[:nh].copy()


def gcs_deps(esz, d1, d2, g, K, banded=True):
    """For a d2 x d1 slice held by g members of K tiles: the range of members
    each member waits for at the group barrier, as two uint32 numpy arrays
    (lo, hi); lo > hi for a member that owns no tile (test hook, host code)."""
    import numpy as np
    lo = np.zeros(g, dtype=np.uint32)
    hi = np.zeros(g, dtype=np.uint32)
    _check(lib().ipt_debug_gcs_deps(esz, d1, d2, g, K, 1 if banded else 0, lo.ctypes.data,
                                    hi.ctypes.data), "ipt_debug_gcs_deps")
    return lo, hi


def version():
    return lib().ipt_version().decode()
