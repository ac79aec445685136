"""Pins the CPU oracle before anything is checked against it.

Checks (each chosen so that a plausible slip — a dropped term, a wrong sign or
index, a transposed operand — fails at least one):
  * seq_tt against the reference's own gen_ans output (golden fixtures made by
    tests/golden/make_golden.py from oracle/_ref/gen_ans) and against numpy's
    axis permutation for random ranks / permutations;
  * extended_gcd and the reduced divisors against the reference's gcd.cpp and
    reduced_math.cpp compiled from source (oracle/_ref/libref_math.so), and
    against exact integer arithmetic;
  * every pass map: bijective per row / column; fused column op == rotate +
    scatter_permute; p(pi(i)) = i; the composition of the C2R and R2C passes
    == seq_tt on every shape with d1, d2 <= 40 and on random larger shapes;
  * the in-place drivers (the OpenMP CPU path) on seeded random bits.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def load_golden():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)["cases"]


def test_seq_tt_matches_reference_gen_ans_small(oracle):
    cases = [c for c in load_golden() if "file" in c]
    assert len(cases) >= 20
    for c in cases:
        d1, d2, d3, esz = c["d1"], c["d2"], c["d3"], c["esz"]
        with open(os.path.join(GOLDEN, c["file"]), "rb") as f:
            ref = np.frombuffer(f.read(), dtype=oracle.dtype_of(esz))
        got = oracle.seq_tt_213(oracle.iota(d1 * d2 * d3, esz), d1, d2, d3)
        assert np.array_equal(got, ref), c


@pytest.mark.parametrize("shape", [(3000, 1000, 1, 4), (512, 384, 64, 4), (256, 1000, 17, 8),
                                   (4099, 4093, 1, 4)])
def test_seq_tt_matches_reference_gen_ans_config_hash(oracle, shape):
    case = [c for c in load_golden() if (c["d1"], c["d2"], c["d3"], c["esz"]) == shape]
    assert case, shape
    d1, d2, d3, esz = shape
    got = oracle.seq_tt_213(oracle.iota(d1 * d2 * d3, esz), d1, d2, d3)
    assert hashlib.sha256(got.tobytes()).hexdigest() == case[0]["sha256"]


def test_seq_tt_rank_n_matches_numpy(oracle):
    rng = np.random.default_rng(5)
    for _ in range(60):
        rank = int(rng.integers(1, 5))
        dims = [int(x) for x in rng.integers(1, 7, size=rank)]
        perm = [int(x) for x in rng.permutation(rank)]
        vol = int(np.prod(dims))
        a = oracle.random_bits(vol, 4, int(rng.integers(1 << 30)))
        got = oracle.seq_tt(a, dims, perm)
        # dims are fastest-first: numpy's C order is slowest-first.
        arr = a.reshape(dims[::-1])
        axes = [rank - 1 - p for p in perm[::-1]]
        exp = np.transpose(arr, axes).reshape(-1)
        assert np.array_equal(got, exp), (dims, perm)
    # rank 3 {1,0,2} general == closed form
    a = oracle.random_bits(5 * 7 * 3, 8, 3)
    assert np.array_equal(oracle.seq_tt(a, [5, 7, 3], [1, 0, 2]), oracle.seq_tt_213(a, 5, 7, 3))


def _ref_math():
    path = os.path.join(oracle_dir(), "_ref", "libref_math.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref/libref_math.so not built (reference checkout absent)")
    L = ctypes.CDLL(path)
    L.ref_extended_gcd.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]
    L.ref_find_divisor32.argtypes = [ctypes.c_uint, ctypes.POINTER(ctypes.c_uint),
                                     ctypes.POINTER(ctypes.c_uint)]
    L.ref_find_divisor64.argtypes = [ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_ulonglong),
                                     ctypes.POINTER(ctypes.c_uint)]
    return L


def oracle_dir():
    return os.path.join(os.path.dirname(HERE), "oracle")


def test_extended_gcd_matches_reference_and_identity(oracle):
    L = _ref_math()
    rng = np.random.default_rng(1)
    pairs = [(a, b) for a in range(1, 60) for b in range(1, 60)]
    pairs += [(int(x), int(y)) for x, y in rng.integers(1, 1 << 30, size=(3000, 2))]
    for a, b in pairs:
        g, m = oracle.extended_gcd(a, b)
        rg, rm = ctypes.c_int(), ctypes.c_int()
        L.ref_extended_gcd(a, b, ctypes.byref(rg), ctypes.byref(rm))
        assert (g, m) == (rg.value, rm.value), (a, b)
        assert g == np.gcd(a, b)
        if g == 1 and b > 1:
            assert (a * m) % b == 1 and 0 <= m < b
        else:
            assert m == 0


def test_reduced_divisor_matches_reference_and_exact(oracle):
    L = _ref_math()
    O = oracle.lib()
    rng = np.random.default_rng(2)
    divisors = list(range(1, 300)) + [1 << s for s in range(31)] + \
        [int(x) for x in rng.integers(2, (1 << 31) - 1, size=400)]
    for d in divisors:
        m, s = ctypes.c_uint32(), ctypes.c_uint32()
        O.ipt_oracle_find_divisor32(d, ctypes.byref(m), ctypes.byref(s))
        rm, rs = ctypes.c_uint(), ctypes.c_uint()
        assert L.ref_find_divisor32(d, ctypes.byref(rm), ctypes.byref(rs)) == 0
        assert (m.value, s.value) == (rm.value, rs.value), d
        xs = [0, 1, d - 1, d, d + 1, (1 << 31) - 1] + [int(x) for x in rng.integers(0, 1 << 31, 50)]
        for x in xs:
            if x < 0 or x >= (1 << 31):
                continue
            assert O.ipt_oracle_rdiv32(x, d, m.value, s.value) == x // d, (x, d)
    for d in [2, 3, 7, 1 << 40, (1 << 62) + 1] + [int(x) for x in rng.integers(2, 1 << 62, 200)]:
        m, s = ctypes.c_uint64(), ctypes.c_uint32()
        O.ipt_oracle_find_divisor64(d, ctypes.byref(m), ctypes.byref(s))
        rm, rs = ctypes.c_ulonglong(), ctypes.c_uint()
        assert L.ref_find_divisor64(d, ctypes.byref(rm), ctypes.byref(rs)) == 0
        assert (m.value, s.value) == (rm.value, rs.value), d
        for x in [0, d - 1, d, (1 << 63) - 1] + [int(x) for x in rng.integers(0, 1 << 62, 30)]:
            assert O.ipt_oracle_rdiv64(x, d, m.value, s.value) == x // d, (x, d)
    # the reference throws for a zero divisor
    rm, rs = ctypes.c_uint(), ctypes.c_uint()
    assert L.ref_find_divisor32(0, ctypes.byref(rm), ctypes.byref(rs)) == -1


def _shapes_tiny():
    return [(R, C) for R in range(1, 25) for C in range(1, 25)]


def test_maps_are_bijections(oracle):
    O = oracle.lib()
    for R, C in _shapes_tiny():
        p = oracle.plan(R, C)
        for which in (oracle.MAP_C2R_PRE, oracle.MAP_C2R_POST, oracle.MAP_R2C_PRE,
                      oracle.MAP_R2C_POST):
            for j in range(C):
                col = sorted(O.ipt_oracle_map(which, ctypes.byref(p), i, j) for i in range(R))
                assert col == list(range(R)), (which, R, C, j)
        for which in (oracle.MAP_C2R_SHUFFLE, oracle.MAP_R2C_SHUFFLE):
            for i in range(R):
                row = sorted(O.ipt_oracle_map(which, ctypes.byref(p), i, j) for j in range(C))
                assert row == list(range(C)), (which, R, C, i)


def test_permuters_are_inverse(oracle):
    O = oracle.lib()
    for R, C in _shapes_tiny() + [(384, 512), (1000, 256), (96, 36)]:
        p = oracle.plan(R, C)
        for i in range(R):
            pi = O.ipt_oracle_permuter(oracle.PERM_C2R_POST, ctypes.byref(p), i)
            assert O.ipt_oracle_permuter(oracle.PERM_R2C_PRE, ctypes.byref(p), pi) == i, (R, C, i)


def test_fused_column_ops_equal_rotate_then_permute(oracle):
    """c2r post: gather (i*C - i/a + j) mod R == rotate(j mod R) then scatter by p;
    r2c pre: gather p((i-j) mod R) == scatter by pi then rotate((R - j mod R) mod R)."""
    O = oracle.lib()
    for R, C in _shapes_tiny():
        p = oracle.plan(R, C)
        P = [O.ipt_oracle_permuter(oracle.PERM_R2C_PRE, ctypes.byref(p), i) for i in range(R)]
        Pi = [O.ipt_oracle_permuter(oracle.PERM_C2R_POST, ctypes.byref(p), i) for i in range(R)]
        for j in range(C):
            # c2r: T[i] = Y[(i + rot(j)) % R]; Z[P[i]] = T[i]  =>  Z[i] = Y[src(i)]
            rot = O.ipt_oracle_rotator(oracle.ROT_C2R_POST, ctypes.byref(p), j)
            Z = [None] * R
            for i in range(R):
                Z[P[i]] = (i + rot) % R
            fused = [O.ipt_oracle_map(oracle.MAP_C2R_POST, ctypes.byref(p), i, j) for i in range(R)]
            assert Z == fused, (R, C, j)
            # r2c: T[Pi[i]] = A[i]; Y[i] = T[(i + rot2(j)) % R]
            T = [None] * R
            for i in range(R):
                T[Pi[i]] = i
            rot2 = O.ipt_oracle_rotator(oracle.ROT_R2C_PRE, ctypes.byref(p), j)
            Y = [T[(i + rot2) % R] for i in range(R)]
            fused2 = [O.ipt_oracle_map(oracle.MAP_R2C_PRE, ctypes.byref(p), i, j) for i in range(R)]
            assert Y == fused2, (R, C, j)


def test_reference_fused_c2r_argument_defect_is_not_followed(oracle):
    """src/transpose.cu:22 passes d2/c where the map needs c.  Replanting that
    slip in a copy of the oracle's C2R post map breaks the composition for most
    shapes, while the oracle's form passes everywhere (checked below)."""
    broken = 0
    total = 0
    for R in range(2, 21):
        for C in range(R, 31):
            p = oracle.plan(R, C)
            if p.c == R // p.c:
                continue
            total += 1
            # full C2R on index values with the replanted post map
            idx = np.arange(R * C, dtype=np.int64).reshape(R, C)
            X = idx.copy()
            if p.c > 1:
                X = np.array([[X[(i + j // p.b) % R][j] for j in range(C)] for i in range(R)])
            O = oracle.lib()
            Y = np.array([[X[i][O.ipt_oracle_map(oracle.MAP_C2R_SHUFFLE, ctypes.byref(p), i, j)]
                           for j in range(C)] for i in range(R)])
            # as written: m.div(i * (d2/c)) = floor(i * a / R) = i // c (instead of i // a)
            Z = np.array([[Y[(i * C - i // p.c + j) % R][j] for j in range(C)]
                          for i in range(R)])
            if not np.array_equal(Z.reshape(C, R), idx.T):
                broken += 1
    assert total > 0 and broken / total > 0.8


@pytest.mark.parametrize("esz", [4, 8])
def test_inplace_drivers_all_tiny_shapes(oracle, esz):
    """C2R / R2C, fused and unfused column ops, and the 213 dispatch, on every
    slice shape with d1, d2 <= 40 (d3 = 2), seeded random bits."""
    for d1 in range(1, 41):
        for d2 in range(1, 41):
            a = oracle.random_bits(d1 * d2 * 2, esz, 1000 * d1 + d2)
            exp = oracle.seq_tt_213(a, d1, d2, 2)
            for fused in (True, False):
                b = a.copy()
                oracle.c2r(b, d1, d2, 2, fused=fused, threads=1)
                assert np.array_equal(b, exp), ("c2r", d1, d2, fused)
                b = a.copy()
                oracle.r2c(b, d2, d1, 2, fused=fused, threads=1)
                assert np.array_equal(b, exp), ("r2c", d1, d2, fused)
            b = a.copy()
            oracle.transpose_213(b, d1, d2, 2, threads=1)
            assert np.array_equal(b, exp), ("213", d1, d2)


def test_inplace_drivers_random_larger_shapes(oracle):
    rng = np.random.default_rng(11)
    for t in range(150):
        d1, d2 = (int(x) for x in rng.integers(33, 401, size=2))
        d3 = int(rng.integers(1, 4))
        esz = 4 if t % 2 else 8
        a = oracle.random_bits(d1 * d2 * d3, esz, t)
        exp = oracle.seq_tt_213(a, d1, d2, d3)
        b = a.copy()
        oracle.transpose_213(b, d1, d2, d3)
        assert np.array_equal(b, exp), (d1, d2, d3, esz)
        b = a.copy()
        oracle.c2r(b, d1, d2, d3, fused=False)
        assert np.array_equal(b, exp), ("c2r unfused", d1, d2, d3, esz)
        b = a.copy()
        oracle.r2c(b, d2, d1, d3, fused=False)
        assert np.array_equal(b, exp), ("r2c unfused", d1, d2, d3, esz)


def test_openmp_entry_row_and_column_major(oracle):
    m, n = 37, 53
    a = oracle.random_bits(m * n, 4, 9)
    b = a.copy()
    assert oracle.openmp_transpose(True, b, m, n) == 0
    assert np.array_equal(b.reshape(n, m), a.reshape(m, n).T)
    b = a.copy()
    assert oracle.openmp_transpose(False, b, m, n) == 0
    # column-major m x n: element (i,j) at i + j*m  ->  becomes n x m column-major
    assert np.array_equal(b.reshape(m, n), a.reshape(n, m).T)


def test_invalid_arguments_rejected(oracle):
    a = oracle.random_bits(16, 4, 1)
    O = oracle.lib()
    assert O.ipt_oracle_c2r(a.ctypes.data, 3, 4, 4, 1, 1, 1) == -1
    assert O.ipt_oracle_c2r(a.ctypes.data, 4, 0, 4, 1, 1, 1) == -1
    assert O.ipt_oracle_transpose_213(None, 4, 4, 4, 1, 1) == -1
