"""ctypes binding of oracle/liboracle.so for the test suite.

Test infrastructure only: the oracle is the checker, never the thing measured.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p

MAP_C2R_PRE, MAP_C2R_POST, MAP_R2C_PRE, MAP_R2C_POST, MAP_C2R_SHUFFLE, MAP_R2C_SHUFFLE = range(6)
ROT_C2R_PRE, ROT_C2R_POST, ROT_R2C_PRE, ROT_R2C_POST = range(4)
PERM_C2R_POST, PERM_R2C_PRE = range(2)


class Plan(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("R", "C", "c", "a", "b", "k", "q")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        L.ipt_oracle_fill_iota.argtypes = [_vp, ctypes.c_int, ctypes.c_uint64]
        L.ipt_oracle_fill_random.argtypes = [_vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.ipt_oracle_seq_tt_213.argtypes = [_vp, _vp, ctypes.c_int, _i64, _i64, _i64]
        L.ipt_oracle_seq_tt.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
        L.ipt_oracle_extended_gcd.argtypes = [_i64, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
        L.ipt_oracle_find_divisor32.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                                ctypes.POINTER(ctypes.c_uint32)]
        L.ipt_oracle_rdiv32.argtypes = [ctypes.c_uint32] * 4
        L.ipt_oracle_rdiv32.restype = ctypes.c_uint32
        L.ipt_oracle_find_divisor64.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                                ctypes.POINTER(ctypes.c_uint32)]
        L.ipt_oracle_rdiv64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]
        L.ipt_oracle_rdiv64.restype = ctypes.c_uint64
        L.ipt_oracle_make_plan.argtypes = [_i64, _i64, ctypes.POINTER(Plan)]
        L.ipt_oracle_map.argtypes = [ctypes.c_int, ctypes.POINTER(Plan), _i64, _i64]
        L.ipt_oracle_map.restype = _i64
        L.ipt_oracle_rotator.argtypes = [ctypes.c_int, ctypes.POINTER(Plan), _i64]
        L.ipt_oracle_rotator.restype = _i64
        L.ipt_oracle_permuter.argtypes = [ctypes.c_int, ctypes.POINTER(Plan), _i64]
        L.ipt_oracle_permuter.restype = _i64
        for name in ("ipt_oracle_c2r", "ipt_oracle_r2c"):
            getattr(L, name).argtypes = [_vp, ctypes.c_int, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int]
        L.ipt_oracle_transpose_213.argtypes = [_vp, ctypes.c_int, _i64, _i64, _i64, ctypes.c_int]
        L.ipt_oracle_openmp_transpose.argtypes = [ctypes.c_int, _vp, ctypes.c_int, _i64, _i64, _vp,
                                                  ctypes.c_int]
        L.ipt_oracle_apply_col_gather.argtypes = [_vp, ctypes.c_int, _i64, _i64, _i64, ctypes.c_int,
                                                  ctypes.c_int]
        L.ipt_oracle_apply_row_gather.argtypes = L.ipt_oracle_apply_col_gather.argtypes
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_vp)


def dtype_of(esz):
    return np.uint32 if esz == 4 else np.uint64


def iota(n, esz):
    a = np.empty(n, dtype=dtype_of(esz))
    lib().ipt_oracle_fill_iota(_ptr(a), esz, n)
    return a


def random_bits(n, esz, seed):
    a = np.empty(n, dtype=dtype_of(esz))
    lib().ipt_oracle_fill_random(_ptr(a), esz, n, seed)
    return a


def seq_tt_213(a, d1, d2, d3):
    """Closed-form oracle: each of the d3 slices (d2 rows x d1 cols) transposed."""
    out = np.empty_like(a)
    lib().ipt_oracle_seq_tt_213(_ptr(a), _ptr(out), a.itemsize, d1, d2, d3)
    return out


def seq_tt(a, dims, perm):
    out = np.empty_like(a)
    d = (ctypes.c_int64 * len(dims))(*dims)
    p = (ctypes.c_int * len(perm))(*perm)
    rc = lib().ipt_oracle_seq_tt(_ptr(a), _ptr(out), a.itemsize, len(dims), d, p)
    assert rc == 0
    return out


def plan(R, C):
    p = Plan()
    lib().ipt_oracle_make_plan(R, C, ctypes.byref(p))
    return p


def extended_gcd(a, b):
    g, m = _i64(), _i64()
    lib().ipt_oracle_extended_gcd(a, b, ctypes.byref(g), ctypes.byref(m))
    return g.value, m.value


def c2r(a, d1, d2, d3, fused=True, threads=0):
    rc = lib().ipt_oracle_c2r(_ptr(a), a.itemsize, d1, d2, d3, int(fused), threads)
    assert rc == 0, rc


def r2c(a, d1, d2, d3, fused=True, threads=0):
    rc = lib().ipt_oracle_r2c(_ptr(a), a.itemsize, d1, d2, d3, int(fused), threads)
    assert rc == 0, rc


def transpose_213(a, d1, d2, d3, threads=0):
    rc = lib().ipt_oracle_transpose_213(_ptr(a), a.itemsize, d1, d2, d3, threads)
    assert rc == 0, rc


def openmp_transpose(row_major, a, m, n, threads=0):
    return lib().ipt_oracle_openmp_transpose(int(row_major), _ptr(a), a.itemsize, m, n, None, threads)


def apply_col_gather(a, R, C, d3, which, threads=0):
    rc = lib().ipt_oracle_apply_col_gather(_ptr(a), a.itemsize, R, C, d3, which, threads)
    assert rc == 0, rc


def apply_row_gather(a, R, C, d3, which, threads=0):
    rc = lib().ipt_oracle_apply_row_gather(_ptr(a), a.itemsize, R, C, d3, which, threads)
    assert rc == 0, rc
