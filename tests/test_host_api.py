"""CPU-only checks of the product library: it loads, exports every symbol the
C header declares, validates arguments before any HIP call, and its own index
maps (the code its kernels run, evaluated on the host) agree with the oracle's
restatement of the reference's maps.  No compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(ipt_[a-z0-9_]+)\s*\(",
                            text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(ipt):
    L = ipt.lib()
    names = header_functions()
    assert "ipt_transpose_213" in names and "ipt_r2c_transpose" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings(ipt):
    L = ipt.lib()
    for s in range(9):
        assert L.ipt_status_string(s)
    assert L.ipt_status_string(99) == b"unknown status"


def test_validation_before_any_hip_call(ipt):
    L = ipt.lib()
    buf = ctypes.create_string_buffer(64)
    p = ctypes.cast(buf, ctypes.c_void_p)
    assert L.ipt_transpose_213(None, 4, 2, 2, 1, None) == 1          # NULL data
    assert L.ipt_transpose_213(p, 3, 2, 2, 1, None) == 1             # bad element size
    assert L.ipt_transpose_213(p, 4, 0, 2, 1, None) == 1             # zero dim
    assert L.ipt_transpose_213(p, 4, 2, -1, 1, None) == 1            # negative dim
    assert L.ipt_transpose_213(p, 8, 1 << 21, 1 << 21, 1, None) == 3  # slice beyond 2^40
    assert L.ipt_c2r_transpose(None, 4, 2, 2, 1, None) == 1
    assert L.ipt_r2c_transpose(p, 16, 2, 2, 1, None) == 1
    assert L.ipt_transpose(1, None, 4, 2, 2, None) == 1
    assert L.ipt_transpose_variant(p, 4, 2, 2, 1, 99, None, None) == 1  # bad variant
    odd = ctypes.c_void_p(buf_addr := ctypes.addressof(buf) + 1)
    assert L.ipt_transpose_213(odd, 4, 2, 2, 1, None) == 1           # misaligned
    del buf_addr


def test_memory_kind_needs_no_device(ipt):
    """ipt_memory_kind enqueues nothing and answers without a GPU: a NULL or
    unknown host pointer is pageable host memory, with no owning device."""
    L = ipt.lib()
    dev = ctypes.c_int(7)
    assert L.ipt_memory_kind(None, ctypes.byref(dev)) == ipt.MEM_PAGEABLE and dev.value == -1
    buf = ctypes.create_string_buffer(64)
    dev.value = 7
    kind = L.ipt_memory_kind(ctypes.cast(buf, ctypes.c_void_p), ctypes.byref(dev))
    assert kind == ipt.MEM_PAGEABLE and dev.value == -1
    assert L.ipt_memory_kind(ctypes.cast(buf, ctypes.c_void_p), None) == ipt.MEM_PAGEABLE


def _rand_points(rng, R, C, n):
    return rng.integers(0, R, size=n), rng.integers(0, C, size=n)


MAPS = [  # (library map id, oracle map id)
    (0, 0), (1, 1), (2, 2), (3, 3), (4, 4), (5, 5)]


def test_library_maps_match_oracle_tiny_shapes(ipt, oracle):
    O = oracle.lib()
    for R in range(1, 25):
        for C in range(1, 25):
            p = oracle.plan(R, C)
            ii, jj = np.meshgrid(np.arange(R), np.arange(C), indexing="ij")
            ii, jj = ii.reshape(-1), jj.reshape(-1)
            for lib_id, or_id in MAPS:
                exp = np.array([O.ipt_oracle_map(or_id, ctypes.byref(p), int(a), int(b))
                                for a, b in zip(ii, jj)])
                for bits in (32, 64):
                    got = ipt.eval_map(lib_id, R, C, bits, ii, jj)
                    assert np.array_equal(got, exp), (lib_id, R, C, bits)


def test_library_maps_match_oracle_large_shapes(ipt, oracle):
    """Config-like and extreme shapes, including 64-bit plans (R*C >= 2^31,
    b^2 >= 2^31) and the widest allowed slices."""
    O = oracle.lib()
    rng = np.random.default_rng(3)
    shapes = [(1000, 3000), (384, 512), (256, 1000), (32, 1 << 24), (7, 1 << 22), (4093, 4099),
              (16384, 16384), (65536, 65535), (3, 1 << 30), (46341, 46340), (1 << 20, 1 << 20)]
    shapes += [(int(a), int(b)) for a, b in rng.integers(2, 1 << 17, size=(25, 2))]
    for R, C in shapes:
        p = oracle.plan(R, C)
        ii, jj = _rand_points(rng, R, C, 400)
        info_bits = [64] + ([32] if (R * C + R < (1 << 31) and p.b * p.b < (1 << 31)
                                     and p.c * R + R < (1 << 31) and (p.a + 2) * p.a < (1 << 31))
                            else [])
        for lib_id, or_id in MAPS:
            exp = np.array([O.ipt_oracle_map(or_id, ctypes.byref(p), int(a), int(b))
                            for a, b in zip(ii, jj)])
            for bits in info_bits:
                got = ipt.eval_map(lib_id, R, C, bits, ii, jj)
                assert np.array_equal(got, exp), (lib_id, R, C, bits)


def test_plan_query_matches_oracle_plan(ipt, oracle):
    for d1, d2 in [(3000, 1000), (512, 384), (256, 1000), (1 << 24, 32), (7, 1 << 22), (4099, 4093),
                   (16384, 16384), (5, 5), (1, 9)]:
        info = ipt.plan_query(4, d1, d2, 1)
        R, C = (d2, d1) if d1 >= d2 else (d1, d2)
        p = oracle.plan(R, C)
        for f in ("R", "C", "c", "a", "b", "k", "q"):
            assert info[f] == getattr(p, f), (d1, d2, f)
        assert info["index_bits"] in (32, 64)
        assert 1 <= info["heuristic_variant"] < ipt.N_VARIANTS


def test_eval_map_rejects_bad_input(ipt):
    with pytest.raises(ipt.TransposeError):
        ipt.eval_map(7, 4, 4, 64, [0], [0])
    with pytest.raises(ipt.TransposeError):
        ipt.eval_map(0, 4, 4, 64, [4], [0])  # i out of range
    with pytest.raises(ipt.TransposeError):
        ipt.eval_map(4, 32, 1 << 24, 32, [0], [0])  # 32-bit plan not allowed here


def test_views_without_exact_64bit_plan_are_refused(ipt):
    """b = C / gcd(R, C) above ~3e9 puts k * (x mod b) of the C2R shuffle past
    2^63: the plan query reports no index width, the host map evaluation and
    every decomposition refuse the view (no silent wrong permutation)."""
    import ctypes
    big = (1 << 33) + 3  # coprime to 3: b = C
    info = ipt.plan_query(4, big, 3)
    assert info["b"] == big and info["index_bits"] == 0
    i = np.zeros(1, np.int64)
    out = np.zeros(1, np.int64)
    L = ipt.lib()
    rc = L.ipt_debug_eval_map(ipt.MAP_C2R_SHUFFLE, 3, big, 64, i.ctypes.data, i.ctypes.data, 1,
                              out.ctypes.data)
    assert rc == 3  # IPT_ERR_TOO_LARGE
    ok = (1 << 31) + 11  # b^2 < 2^63: the 64-bit plan is exact
    assert ipt.plan_query(4, ok, 3)["index_bits"] == 64


def _gcs_owners(esz, d1, d2, g, K, banded):
    """Owner member of every tile of a d2 x d1 slice (nti x ntj), restated
    from the tile orders' definitions: banded, positions walked band by band
    (bands 8 tile rows tall, the last one shorter), column by column inside a
    band, member m owning positions m*K .. m*K + K - 1; strided, tile t = ti *
    ntj + tj owned by member t % g."""
    TS = 32 if esz == 4 else 16
    nti, ntj = -(-d2 // TS), -(-d1 // TS)
    owner = np.empty((nti, ntj), dtype=np.int64)
    if not banded:
        owner[:] = (np.arange(nti * ntj) % g).reshape(nti, ntj)
        return owner, TS
    bh = min(8, nti)
    p = 0
    for b0 in range(0, nti, bh):
        h = min(bh, nti - b0)
        for tj in range(ntj):
            for r in range(h):
                owner[b0 + r, tj] = p // K
                p += 1
    return owner, TS


def _gcs_dep_violations(esz, d1, d2, g, K, banded, lo, hi):
    """Elements whose loader is neither their writer nor in the writer's
    wait range lo..hi (brute force over every element of the slice)."""
    owner, TS = _gcs_owners(esz, d1, d2, g, K, banded)
    R, C = d2, d1
    e = np.arange(R * C, dtype=np.int64)
    loader = owner[(e // C) // TS, (e % C) // TS]        # holds input element e
    writer = owner[(e % R) // TS, (e // R) // TS]        # stores output element e
    lo64, hi64 = lo.astype(np.int64), hi.astype(np.int64)
    ok = (loader == writer) | ((loader >= lo64[writer]) & (loader <= hi64[writer]))
    return int((~ok).sum())


def test_gcs_barrier_waits_cover_every_overwritten_tile(ipt):
    """The cooperative kernels' group barrier lets a member store once the
    members in its range lo..hi have arrived: every member whose staged tiles
    hold an element it overwrites must be in that range (brute force over
    every element, tile owners restated from the orders' definitions), on
    random slices, both element sizes, both tile orders, with and without
    idle members.  The banded ranges must also be narrower than the group
    somewhere, and a range cut short by one member must be caught."""
    rng = np.random.default_rng(5)
    narrowed = caught = 0
    for case in range(160):
        esz = 4 if case % 2 else 8
        TS = 32 if esz == 4 else 16
        d1 = int(rng.integers(1, 900)) if case % 5 else int(rng.integers(1, 40))
        d2 = int(rng.integers(1, 900)) if case % 7 else int(rng.integers(1, 40))
        ntiles = (-(-d2 // TS)) * (-(-d1 // TS))
        K = int(rng.integers(1, max(2, ntiles // 3) + 1))
        g = -(-ntiles // K) + int(rng.integers(0, 3))
        for banded in (True, False):
            lo, hi = ipt.gcs_deps(esz, d1, d2, g, K, banded)
            assert _gcs_dep_violations(esz, d1, d2, g, K, banded, lo, hi) == 0, (esz, d1, d2, g, K, banded)
            live = lo <= hi
            if banded and live.any() and (hi[live] - lo[live] + 1).max() < g:
                narrowed += 1
            # a planted mistake: one member's range cut short at its top
            cut = np.where(live & (hi > lo))[0]
            if len(cut):
                m = int(cut[rng.integers(0, len(cut))])
                hi2 = hi.copy()
                hi2[m] -= 1
                caught += _gcs_dep_violations(esz, d1, d2, g, K, banded, lo, hi2) > 0
    assert narrowed >= 20, narrowed
    assert caught >= 40, caught
