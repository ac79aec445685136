"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bit-exact comparisons on seeded random bits (NaN payloads included) and on the
reference's iota input against the golden fixtures written by the reference's
own gen_ans.  Every variant and every pass tier is exercised: tiers that only
appear at large sizes are forced at small sizes with a small LDS cap.
"""
import hashlib
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def to_dev(a):
    return torch.from_numpy(a.view(np.int32 if a.itemsize == 4 else np.int64)).cuda()


def to_host(t, esz):
    return t.cpu().numpy().view(np.uint32 if esz == 4 else np.uint64)


def run_case(ipt, oracle, d1, d2, d3, esz, variant, seed, tune=None):
    print(f"[case] d1={d1} d2={d2} d3={d3} esz={esz} variant={variant} seed={seed}", flush=True)
    a = oracle.random_bits(d1 * d2 * d3, esz, seed)
    exp = oracle.seq_tt_213(a, d1, d2, d3)
    t = to_dev(a)
    ipt.transpose_variant(t, d1, d2, d3, variant, tune=tune)
    torch.cuda.synchronize()
    got = to_host(t, esz)
    assert np.array_equal(got, exp), (d1, d2, d3, esz, variant)


SHAPES = [
    (1, 1, 1), (1, 9, 2), (9, 1, 3), (2, 3, 1), (3, 2, 5), (17, 17, 3), (32, 31, 2), (31, 32, 2),
    (40, 24, 3), (24, 40, 3), (64, 48, 4), (48, 64, 4), (97, 41, 2), (41, 97, 2), (128, 96, 1),
    (300, 100, 1), (100, 300, 2), (512, 12, 3), (12, 512, 3), (1000, 3, 2), (3, 1000, 2),
    (255, 257, 1), (768, 512, 1), (512, 768, 1), (1000, 999, 1), (1024, 1024, 1), (333, 333, 2),
    (4096, 7, 1), (7, 4096, 1), (2000, 64, 2), (64, 2000, 2),
]


@pytest.mark.parametrize("esz", [4, 8])
def test_auto_dispatch_matches_oracle(ipt, oracle, esz):
    for k, (d1, d2, d3) in enumerate(SHAPES):
        run_case(ipt, oracle, d1, d2, d3, esz, ipt.VARIANT_AUTO, 100 + k)


@pytest.mark.parametrize("variant", ["c2r", "r2c", "slice_lds", "square_swap", "gcs", "block_cycle"])
@pytest.mark.parametrize("esz", [4, 8])
def test_each_variant_matches_oracle(ipt, oracle, variant, esz):
    v = {"c2r": ipt.VARIANT_C2R, "r2c": ipt.VARIANT_R2C, "slice_lds": ipt.VARIANT_SLICE_LDS,
         "square_swap": ipt.VARIANT_SQUARE_SWAP, "gcs": ipt.VARIANT_GCS,
         "block_cycle": ipt.VARIANT_BLOCK_CYCLE}[variant]
    ran = 0
    for k, (d1, d2, d3) in enumerate(SHAPES):
        if not ipt.feasible(v, esz, d1, d2, d3):
            with pytest.raises(ipt.TransposeError):
                ipt.transpose_variant(to_dev(oracle.random_bits(d1 * d2 * d3, esz, 1)), d1, d2, d3, v)
            continue
        run_case(ipt, oracle, d1, d2, d3, esz, v, 200 + k)
        ran += 1
    assert ran > 0


@pytest.mark.parametrize("esz", [4, 8])
def test_square_swap_any_side_and_base(ipt, oracle, esz):
    """Square swap on sides that are not a multiple of 16 bytes (partial
    micro-rows in the last tile column) and bases off 16-byte alignment by
    1-3 elements, several slices: the neighbours of the slice are untouched."""
    L = ipt.lib()
    import ctypes
    for k, (n, d3, off) in enumerate([(63, 3, 1), (65, 2, 0), (97, 1, 3), (129, 2, 2), (255, 1, 1),
                                      (1001, 1, 0), (1001, 2, 3), (2049, 1, 1)]):
        if esz == 8 and n == 2049:
            d3 = 2
        print(f"[case] square n={n} d3={d3} esz={esz} off={off}", flush=True)
        m = n * n * d3
        a = oracle.random_bits(m + 8, esz, 700 + k)
        t = to_dev(a)
        view = t[off:off + m]
        st = L.ipt_transpose_variant(ctypes.c_void_p(view.data_ptr()), esz, n, n, d3,
                                     ipt.VARIANT_SQUARE_SWAP, None,
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert st == 0, (st, L.ipt_last_error())
        torch.cuda.synchronize()
        exp = a.copy()
        exp[off:off + m] = oracle.seq_tt_213(a[off:off + m], n, n, d3)
        assert np.array_equal(to_host(t, esz), exp), (n, d3, esz, off)


def test_square_swap_diagonal_pair_order(ipt, oracle):
    """Squares whose row pitch is a power of two of 16 KB or more take their
    tile pairs in diagonal order: f32 4096^2 x 2, f64 4096^2 and 8192^2 x 2,
    f32 16384^2 off 16-byte alignment through the buffer-resource form, f32
    16384^2 with half-size tiles, all against the oracle on random bits; f64
    16384^2 (2 GiB, the 16-byte form) against the closed form on iota input."""
    import ctypes
    L = ipt.lib()
    for k, (n, d3, esz, off, tile) in enumerate([(4096, 2, 4, 0, 0), (4096, 1, 8, 2, 0),
                                                  (8192, 2, 8, 0, 0), (16384, 1, 4, 1, 0),
                                                  (16384, 1, 4, 0, 32)]):
        print(f"[case] square n={n} d3={d3} esz={esz} off={off} tile={tile}", flush=True)
        m = n * n * d3
        a = oracle.random_bits(m + 4, esz, 900 + k)
        t = to_dev(a)
        view = t[off:off + m]
        tune = ipt.Tune(tile=tile) if tile else None
        st = L.ipt_transpose_variant(ctypes.c_void_p(view.data_ptr()), esz, n, n, d3,
                                     ipt.VARIANT_SQUARE_SWAP, ctypes.byref(tune) if tune else None,
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert st == 0, (st, L.ipt_last_error())
        torch.cuda.synchronize()
        exp = a.copy()
        exp[off:off + m] = oracle.seq_tt_213(a[off:off + m], n, n, d3)
        assert np.array_equal(to_host(t, esz), exp), (n, d3, esz, off, tile)
        del t, view, a, exp
        torch.cuda.empty_cache()
    n = 16384
    print(f"[case] square n={n} esz=8 iota", flush=True)
    t = iota_dev(n * n, 8)
    ipt.transpose_variant(t, n, n, 1, ipt.VARIANT_SQUARE_SWAP)
    torch.cuda.synchronize()
    e = torch.arange(n * n, dtype=torch.int64, device="cuda").view(n, n).t().contiguous().view(-1)
    assert torch.equal(t, e)
    del t, e
    torch.cuda.empty_cache()


BLOCK_CYCLE_SHAPES = [  # (d1, d2, d3): skinny wide / tall, then large rectangles
    (65536, 32, 1), (32, 65536, 1), (7, 1 << 18, 2), (1 << 16, 5, 3), (128, 4096, 2),
    (4096, 128, 2), (3, 1 << 19, 1), (1 << 19, 3, 1), (2000, 64, 2), (64, 2000, 2),
    (1000, 3, 2), (3, 1000, 2), (6000, 100, 1), (100, 6000, 1),
    (6144, 4096, 1), (4096, 6144, 1), (5000, 3000, 2), (3000, 5000, 1), (8192, 2048, 3),
    (10007, 2000, 1), (2000, 10007, 1), (1500, 1000, 3),
    # chunk sizes that are not multiples of 16 bytes (odd T): 8- and 4-byte vectors
    (3003, 1001, 2), (1001, 3003, 1), (777, 4095, 1), (3751, 3723, 1),
]


@pytest.mark.parametrize("esz", [4, 8])
def test_block_cycle_path_matches_oracle(ipt, oracle, esz):
    """Block-cycle path (blocks transposed by a one-sweep kernel + segmented
    cycle-following chunk permutation), skinny and large rectangular slices:
    up to millions of chunks, several slices per call, every block kernel;
    back-to-back calls reuse the cached segment tables."""
    ran = 0
    for k, (d1, d2, d3) in enumerate(BLOCK_CYCLE_SHAPES):
        if not ipt.feasible(ipt.VARIANT_BLOCK_CYCLE, esz, d1, d2, d3):
            continue
        for rep in range(2):
            run_case(ipt, oracle, d1, d2, d3, esz, ipt.VARIANT_BLOCK_CYCLE, 900 + 10 * k + rep)
        ran += 1
    assert ran >= 16


def test_block_cycle_unaligned_data(ipt, oracle):
    """Data one element past a 16-byte boundary: the chunk step moves 4-byte
    vectors instead of 16-byte ones; forced and AUTO both exact."""
    d1, d2, d3, esz = 7, 1 << 22, 1, 4  # beyond GCS: AUTO resolves to the block-cycle path
    a = oracle.random_bits(d1 * d2 * d3 + 1, esz, 990)
    t = to_dev(a)
    view = t[1:]  # 4-byte offset
    ipt.transpose_variant(view, d1, d2, d3, ipt.VARIANT_BLOCK_CYCLE)
    torch.cuda.synchronize()
    got = to_host(t, esz)
    assert got[0] == a[0]
    exp = oracle.seq_tt_213(a[1:], d1, d2, d3)
    assert np.array_equal(got[1:], exp)
    ipt.transpose(view, d2, d1, d3)  # AUTO, back
    torch.cuda.synchronize()
    got = to_host(t, esz)
    assert np.array_equal(got, a)


@pytest.mark.parametrize("esz", [4, 8])
def test_gcs_groups_and_barriers(ipt, oracle, esz):
    """Cooperative kernel with large groups (small LDS caps), many slices per
    launch (several rounds of slots), rows of any length and
    back-to-back launches (counters restored to zero by each launch)."""
    cases = [((512, 384, 64), None), ((256, 1000, 17), None), ((3000, 1000, 1), None),
             ((97, 203, 40), ipt.Tune(lds_bytes=8192)), ((300, 100, 9), ipt.Tune(lds_bytes=4608)),
             ((1000, 999, 1), ipt.Tune(lds_bytes=65536)), ((64, 2000, 3), ipt.Tune(lds_bytes=20000)),
             ((33, 35, 500), ipt.Tune(lds_bytes=4608))]
    ran = 0
    for k, ((d1, d2, d3), tune) in enumerate(cases):
        if not ipt.feasible(ipt.VARIANT_GCS, esz, d1, d2, d3):
            continue
        for rep in range(2):
            try:
                run_case(ipt, oracle, d1, d2, d3, esz, ipt.VARIANT_GCS, 600 + 10 * k + rep, tune=tune)
            except ipt.TransposeError as e:
                # a small LDS cap can make the group larger than one per CU
                assert tune is not None and e.status == 2, e
                break
            ran += 1
        ipt.device_check()
    assert ran >= 8


@pytest.mark.parametrize("esz", [4, 8])
def test_gcs_any_row_length_and_base(ipt, oracle, esz):
    """The cooperative kernel's 16-byte tiles on rows of any length and on a
    base one element past a 16-byte boundary: partial micro-rows at the right
    and bottom edges, loads that run past the end of the slice, element-wise
    stores in the last row band.  f64 3751 x 3723 (112 MB) and f32 5400 x 5300
    (114 MB) need the largest register-tile class (18 / 9 tiles per wave) to
    fit one group."""
    cases = [(1001, 999, 3), (333, 77, 40), (97, 203, 5), (4093, 4099, 1), (35, 33, 300),
             (1023, 1, 1), (1, 1023, 2), (17, 4099, 2), (4099, 17, 2), (3751, 3723, 1),
             (5400, 5300, 1)]
    ran = 0
    for k, (d1, d2, d3) in enumerate(cases):
        if not ipt.feasible(ipt.VARIANT_GCS, esz, d1, d2, d3):
            continue
        for off in (0, 1):
            print(f"[case] gcs d1={d1} d2={d2} d3={d3} esz={esz} offset={off}", flush=True)
            n = d1 * d2 * d3
            a = oracle.random_bits(n + 1, esz, 700 + 10 * k + off)
            t = to_dev(a)
            ipt.transpose_variant(t[off:off + n], d1, d2, d3, ipt.VARIANT_GCS)
            torch.cuda.synchronize()
            got = to_host(t, esz)
            exp = a.copy()
            exp[off:off + n] = oracle.seq_tt_213(a[off:off + n], d1, d2, d3)
            assert np.array_equal(got, exp), (d1, d2, d3, esz, off)
            ran += 1
    ipt.device_check()
    assert ran >= 12


def test_gcs_watchdog_reports_and_recovers(ipt, oracle):
    """A barrier that cannot complete (one arrival more than the group) gives
    up after a few polls: the kernel still ends, ipt_device_check reports
    the watchdog and clears it, and the next transposes are exact."""
    d1, d2, d3, esz = 512, 384, 8, 4
    a = oracle.random_bits(d1 * d2 * d3, esz, 3100)
    t = to_dev(a)
    ipt.set_watchdog(2000, 1)
    try:
        ipt.transpose_variant(t, d1, d2, d3, ipt.VARIANT_GCS)
        torch.cuda.synchronize()
        with pytest.raises(ipt.TransposeError):
            ipt.device_check()
    finally:
        ipt.set_watchdog(0, 0)
    ipt.device_check()
    for k, (e1, e2, e3, ez) in enumerate([(512, 384, 8, 4), (256, 1000, 3, 8)]):
        run_case(ipt, oracle, e1, e2, e3, ez, ipt.VARIANT_GCS, 3110 + k)
    ipt.device_check()


@pytest.mark.parametrize("esz", [4, 8])
def test_scratch_tiers_forced_by_small_lds_cap(ipt, oracle, esz):
    """lds_bytes=256 pushes every column pass with R*esz > 256 and every row
    pass with C*esz > 240 to the scratch gather + copy-back tiers."""
    tune = ipt.Tune(lds_bytes=256)
    for k, (d1, d2, d3) in enumerate([(300, 100, 2), (100, 300, 1), (97, 41, 3), (768, 512, 1),
                                      (1000, 3, 2), (3, 1000, 2)]):
        for v in (ipt.VARIANT_C2R, ipt.VARIANT_R2C):
            run_case(ipt, oracle, d1, d2, d3, esz, v, 300 + k, tune=tune)


@pytest.mark.parametrize("esz", [4, 8])
def test_tune_knobs_do_not_change_results(ipt, oracle, esz):
    for k, tune in enumerate([ipt.Tune(block=64), ipt.Tune(block=1024, strip=256),
                              ipt.Tune(strip=1), ipt.Tune(grid_mult=1, lds_bytes=4096),
                              ipt.Tune(lds_bytes=32768, block=512, strip=16)]):
        for d1, d2, d3 in [(300, 100, 2), (100, 300, 1), (512, 384, 2), (64, 2000, 1)]:
            for v in (ipt.VARIANT_C2R, ipt.VARIANT_R2C):
                run_case(ipt, oracle, d1, d2, d3, esz, v, 400 + k, tune=tune)


@pytest.mark.parametrize("esz", [4, 8])
def test_each_pass_matches_oracle_pass(ipt, oracle, esz):
    """Per-pass parity: one GPU gather pass vs the oracle's same pass."""
    # (4096, 96) and (2500, 300): tall columns, so the rotation maps take the
    # wide-strip scratch tier by choice
    # (1100, 1024, 16): many strips per slice, several slices
    cases = [(384, 512, 2), (256, 1000, 1), (1000, 256, 1), (96, 36, 3), (7, 4096, 1), (45, 45, 2),
             (4096, 96, 2), (2500, 300, 1), (1100, 1024, 16)]
    for k, (R, C, d3) in enumerate(cases):
        for map_id in range(6):
            for tune in (None, ipt.Tune(lds_bytes=256)):
                print(f"[pass] R={R} C={C} d3={d3} esz={esz} map={map_id} tune={tune is not None}",
                      flush=True)
                a = oracle.random_bits(R * C * d3, esz, 500 + k)
                exp = a.copy()
                if map_id < 4:
                    oracle.apply_col_gather(exp, R, C, d3, map_id)
                else:
                    oracle.apply_row_gather(exp, R, C, d3, map_id)
                t = to_dev(a)
                ipt.run_pass(t, R, C, d3, map_id, tune=tune)
                torch.cuda.synchronize()
                assert np.array_equal(to_host(t, esz), exp), (R, C, d3, map_id)


@pytest.mark.parametrize("esz", [4, 8])
def test_decomposition_tall_columns(ipt, oracle, esz):
    """C2R and R2C on slices with tall columns and gcd > 1: the rotation
    passes go through the wide-strip scratch tier, the others through narrow
    LDS strips."""
    for k, (d1, d2, d3) in enumerate([(96, 4096, 2), (4096, 96, 2), (300, 2500, 1), (2500, 300, 1),
                                      (1100, 1024, 16)]):
        for v in (ipt.VARIANT_C2R, ipt.VARIANT_R2C):
            run_case(ipt, oracle, d1, d2, d3, esz, v, 1400 + 10 * k + v)


PERMS3 = [(0, 1, 2), (1, 0, 2), (0, 2, 1), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


@pytest.mark.parametrize("esz", [4, 8])
def test_permute3_matches_oracle(ipt, oracle, esz):
    """Every rank-3 permutation against the oracle's rank-N seq_tt: inner dims
    of one element (super-element 213), two (8-byte super-elements for f32),
    odd sizes (4- / 8-byte chunk vectors) and 16-byte multiples, several outer
    shapes; and a base one element past a 16-byte boundary."""
    shapes = [(5, 7, 9), (1, 50, 40), (2, 33, 31), (16, 33, 20), (3, 100, 64), (64, 48, 30),
              (8, 2, 300), (300, 5, 7), (33, 17, 65)]
    for k, dims in enumerate(shapes):
        n = dims[0] * dims[1] * dims[2]
        for perm in PERMS3:
            for off in (0, 1) if k % 3 == 0 else (0,):
                print(f"[case] permute3 dims={dims} perm={perm} esz={esz} offset={off}", flush=True)
                a = oracle.random_bits(n + 1, esz, 3300 + 10 * k)
                t = to_dev(a)
                ipt.permute3(t[off:off + n], dims, perm)
                torch.cuda.synchronize()
                got = to_host(t, esz)
                exp = a.copy()
                exp[off:off + n] = oracle.seq_tt(a[off:off + n], list(dims), list(perm))
                assert np.array_equal(got, exp), (dims, perm, esz, off)
    ipt.device_check()


@pytest.mark.parametrize("esz", [4, 8])
def test_permute_rank_n_matches_oracle(ipt, oracle, esz):
    """Random permutations of ranks 1-6 (dims of 1-9, some of size 1) against
    the oracle's rank-N seq_tt."""
    rng = np.random.default_rng(77 + esz)
    for k in range(40):
        rank = int(rng.integers(1, 7))
        dims = [int(x) for x in rng.integers(1, 10, size=rank)]
        if rank >= 5:
            dims = [max(1, d // 2) for d in dims]
        perm = [int(x) for x in rng.permutation(rank)]
        n = int(np.prod(dims))
        off = k % 2
        print(f"[case] permute dims={dims} perm={perm} esz={esz} offset={off}", flush=True)
        a = oracle.random_bits(n + 1, esz, 3500 + k)
        t = to_dev(a)
        ipt.permute(t[off:off + n], dims, perm)
        torch.cuda.synchronize()
        got = to_host(t, esz)
        exp = a.copy()
        exp[off:off + n] = oracle.seq_tt(a[off:off + n], dims, perm)
        assert np.array_equal(got, exp), (dims, perm, esz, off)
    ipt.device_check()


@pytest.mark.parametrize("esz", [4, 8])
def test_permute_large_inner_chunks(ipt, oracle, esz):
    """Exchanges whose inner chunks are large (tens of KB, beyond one
    segment's staging) move them in parts; odd chunk lengths narrow the
    vectors.  Against the oracle's rank-N seq_tt."""
    cases = [([20000, 3, 4], [0, 2, 1]), ([6001, 5, 3], [0, 2, 1]), ([3000, 7, 2, 3], [0, 3, 1, 2]),
             ([9000, 2, 5], [0, 2, 1])]
    for k, (dims, perm) in enumerate(cases):
        n = int(np.prod(dims))
        print(f"[case] permute large chunks dims={dims} perm={perm} esz={esz}", flush=True)
        a = oracle.random_bits(n, esz, 3700 + k)
        t = to_dev(a)
        ipt.permute(t, dims, perm)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(t, esz), oracle.seq_tt(a, dims, perm)), (dims, perm, esz)
    ipt.device_check()


def test_permute3_rejects_bad_perm(ipt):
    t = torch.zeros(24, dtype=torch.int32, device="cuda")
    for perm in [(0, 0, 1), (0, 1, 3), (-1, 1, 2)]:
        with pytest.raises(ipt.TransposeError):
            ipt.permute3(t, (2, 3, 4), perm)
    with pytest.raises(ipt.TransposeError):
        ipt.permute(t, (2, 3, 2, 2, 1, 1, 1), (0, 1, 2, 3, 4, 5, 6))  # rank 7


def test_reference_entry_points(ipt, oracle):
    d1, d2, d3 = 300, 100, 2
    a = oracle.random_bits(d1 * d2 * d3, 4, 7)
    exp = oracle.seq_tt_213(a, d1, d2, d3)
    for fn in (ipt.transpose, ipt.c2r.transpose):
        t = to_dev(a)
        fn(t, d1, d2, d3)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(t, 4), exp)
    # r2c::transpose(data, d1, d2, d3): slices of d1 rows x d2 cols -> d2 x d1
    t = to_dev(a)
    ipt.r2c.transpose(t, d2, d1, d3)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(t, 4), exp)
    # openmp-style row-/column-major m x n
    m, n = 123, 77
    b = oracle.random_bits(m * n, 8, 8)
    t = to_dev(b)
    ipt.openmp_style(True, t, m, n)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(t, 8).reshape(n, m), b.reshape(m, n).T)
    t = to_dev(b)
    ipt.openmp_style(False, t, m, n)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(t, 8).reshape(m, n), b.reshape(n, m).T)


def test_transpose_twice_is_identity_and_stream_ordered(ipt, oracle):
    s = torch.cuda.Stream()
    a = oracle.random_bits(1000 * 999 * 2, 4, 9)
    t = to_dev(a)
    with torch.cuda.stream(s):
        ipt.transpose(t, 1000, 999, 2)
        ipt.transpose(t, 999, 1000, 2)
    s.synchronize()
    assert np.array_equal(to_host(t, 4), a)


def test_wrong_sizes_raise(ipt, oracle):
    t = to_dev(oracle.random_bits(64, 4, 1))
    with pytest.raises(ipt.TransposeError):
        ipt.transpose(t, 8, 9, 1)
    with pytest.raises(ipt.TransposeError):
        ipt.transpose_variant(t, 16, 4, 1, ipt.VARIANT_SQUARE_SWAP)


def golden_case(shape):
    with open(os.path.join(HERE, "golden", "golden.json")) as f:
        for c in json.load(f)["cases"]:
            if (c["d1"], c["d2"], c["d3"], c["esz"]) == shape:
                return c
    raise KeyError(shape)


def iota_dev(n, esz):
    dt = torch.int32 if esz == 4 else torch.int64
    return torch.arange(n, dtype=dt, device="cuda")


def sha256_dev(t):
    h = hashlib.sha256()
    flat = t.view(-1)
    step = 1 << 27
    for s in range(0, flat.numel(), step):
        h.update(flat[s:s + step].cpu().numpy().tobytes())
    return h.hexdigest()


CONFIGS = [(3000, 1000, 1, 4), (512, 384, 64, 4), (256, 1000, 17, 8), (7, 1 << 22, 1, 4),
           (4099, 4093, 1, 4), (16384, 16384, 1, 4), (1 << 24, 32, 1, 8)]


@pytest.mark.parametrize("shape", CONFIGS, ids=lambda s: "x".join(map(str, s)))
def test_config_shapes_match_reference_gen_ans(ipt, shape):
    """Full-size configurations on the reference's iota input: the SHA-256 of
    the GPU result equals that of the reference's own gen_ans output."""
    d1, d2, d3, esz = shape
    ref = golden_case(shape)["sha256"]
    variants = [ipt.VARIANT_AUTO]
    if d1 * d2 * d3 * esz <= (1 << 28):
        variants += [ipt.VARIANT_C2R, ipt.VARIANT_R2C]
    for v in variants:
        print(f"[config] {shape} variant={v}", flush=True)
        t = iota_dev(d1 * d2 * d3, esz)
        ipt.transpose_variant(t, d1, d2, d3, v)
        torch.cuda.synchronize()
        assert sha256_dev(t) == ref, (shape, v)
        del t
        torch.cuda.empty_cache()


@pytest.mark.parametrize("esz", [4, 8])
def test_out_of_place_comparator_matches_oracle(ipt, oracle, esz):
    """ipt_transpose_oop (the out-of-place comparator): dst is the oracle's
    seq_tt of src, src untouched; 16-byte and element-wise forms."""
    for k, (d1, d2, d3) in enumerate([(512, 384, 3), (3000, 1000, 1), (4099, 4093, 1), (7, 300, 5),
                                      (64, 64, 2), (1, 9, 2)]):
        print(f"[case] oop d1={d1} d2={d2} d3={d3} esz={esz}", flush=True)
        a = oracle.random_bits(d1 * d2 * d3, esz, 7000 + k)
        src = to_dev(a)
        dst = torch.empty_like(src)
        ipt.transpose_oop(dst, src, d1, d2, d3)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(dst, esz), oracle.seq_tt_213(a, d1, d2, d3)), (d1, d2, d3)
        assert np.array_equal(to_host(src, esz), a)
    with pytest.raises(ipt.TransposeError):
        ipt.transpose_oop(src, src, 1, 9, 2)  # overlapping buffers are refused
