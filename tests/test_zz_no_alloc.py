"""ipt_transpose_213_noalloc (the entry point that allocates nothing) against
ipt_transpose_213 through the C ABI: bit-exact on the same inputs, on two
calls in a row, the second with a smaller input, for every route the entry
point takes (cooperative groups with their barrier state in the static array,
whole slices, square swaps, the decomposition staging through the array in
place of the block-cycle path, cooperative groups whose barrier state does
not fit the array's share); the library's own arenas are left as they were
(nothing allocated); a shape whose staging does not fit is refused before
anything is enqueued."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(a.view(np.int32 if a.itemsize == 4 else np.int64)).cuda()


def _both(ipt, oracle, shape, seed, check_oracle=True):
    d1, d2, d3, esz = shape
    print(f"[case] noalloc d1={d1} d2={d2} d3={d3} esz={esz}", flush=True)
    a = oracle.random_bits(d1 * d2 * d3, esz, seed)
    ref = _dev(a)
    got = ref.clone()
    L = ipt.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.ipt_transpose_213(ctypes.c_void_p(ref.data_ptr()), esz, d1, d2, d3, stream) == 0
    assert L.ipt_transpose_213_noalloc(ctypes.c_void_p(got.data_ptr()), esz, d1, d2, d3, stream) == 0
    torch.cuda.synchronize()
    assert torch.equal(got, ref), shape
    if check_oracle:
        exp = oracle.seq_tt_213(a, d1, d2, d3)
        assert np.array_equal(got.cpu().numpy().view(a.dtype), exp), shape


PAIRS = [  # (first call, second call with a smaller input)
    ((512, 384, 8, 4), (256, 1000, 3, 8)),        # cooperative groups (barrier state in the array)
    ((100, 60, 3, 4), (33, 17, 2, 8)),            # whole slices in LDS
    ((1024, 1024, 1, 4), (512, 512, 1, 8)),       # square swap
    ((7, 1 << 18, 1, 4), (5, 1 << 17, 1, 8)),     # decomposition through the staging
    ((1 << 18, 9, 1, 8), (1 << 17, 3, 1, 4)),     # the same, tall
    ((3000, 1000, 1, 4), (1000, 999, 1, 4)),      # single slices
]


@pytest.mark.parametrize("pair", PAIRS)
def test_noalloc_matches_transpose_213(ipt, oracle, pair):
    before = (ipt.arena_info(ipt.ARENA_SYNC), ipt.arena_info(ipt.ARENA_SCRATCH))
    first, second = pair
    _both(ipt, oracle, first, 800)
    _both(ipt, oracle, second, 801)
    torch.cuda.synchronize()
    # the ipt_transpose_213 calls may have grown the arenas; run the no-alloc
    # entry alone again and check it leaves them as they are
    after_ref = (ipt.arena_info(ipt.ARENA_SYNC), ipt.arena_info(ipt.ARENA_SCRATCH))
    for k, (d1, d2, d3, esz) in enumerate((first, second)):
        a = oracle.random_bits(d1 * d2 * d3, esz, 810 + k)
        t = _dev(a)
        ipt.transpose_noalloc(t, d1, d2, d3)
        torch.cuda.synchronize()
        assert np.array_equal(t.cpu().numpy().view(a.dtype), oracle.seq_tt_213(a, d1, d2, d3))
    assert (ipt.arena_info(ipt.ARENA_SYNC), ipt.arena_info(ipt.ARENA_SCRATCH)) == after_ref
    del before
    ipt.device_check()


def test_noalloc_barrier_state_beyond_its_share(ipt, oracle):
    """3000 slices of two-member groups need more barrier state than the
    array's 1 MiB: the entry point takes the decomposition instead."""
    _both(ipt, oracle, (300, 200, 3000, 4), 820, check_oracle=False)
    _both(ipt, oracle, (300, 200, 2900, 4), 821, check_oracle=False)
    ipt.device_check()


def test_noalloc_refuses_staging_beyond_the_array(ipt, oracle):
    """Rows of 33.5 MB (3 x 8388617 f32 with no one-sweep or block-cycle route
    that fits) exceed the 31 MiB staging: IPT_ERR_TOO_LARGE before anything
    is enqueued, the data untouched; a smaller call after it is exact."""
    d1, d2 = 8388617, 3
    a = oracle.random_bits(d1 * d2, 4, 830)
    t = _dev(a)
    with pytest.raises(ipt.TransposeError) as err:
        ipt.transpose_noalloc(t, d1, d2, 1)
    assert err.value.status == 3
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy().view(a.dtype), a)
    _both(ipt, oracle, (512, 384, 2, 4), 831)
    ipt.device_check()
