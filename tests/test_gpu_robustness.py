"""Failure reporting and the lifetime of library memory under graph capture.

- A cooperative barrier that gives up (tripped on purpose through the
  watchdog test hook: one arrival more than the group, a few thousand polls,
  so every kernel still ends) is reported by the next library call on the
  device, without any explicit device check, and the calls after it are exact.
- Library buffers that a caller's captured graph references (the barrier
  arena, the scratch arena) stay allocated when the arena later grows: the
  old buffer is retired, not freed.  That is checked on the host first (the
  retired count and the old pointer's allocation); only then are the graphs
  replayed and checked against the oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def to_dev(a):
    return torch.from_numpy(a.view(np.int32 if a.itemsize == 4 else np.int64)).cuda()


def to_host(t, a):
    return t.cpu().numpy().view(a.dtype)


def test_tripped_barrier_is_reported_by_the_next_call(ipt, oracle):
    d1, d2, d3, esz = 512, 384, 8, 4  # cooperative groups of several members
    a = oracle.random_bits(d1 * d2 * d3, esz, 5100)
    t = to_dev(a)
    ipt.device_check()
    ipt.set_watchdog(2000, 1)
    try:
        print("[case] tripping the watchdog on purpose (GCS 512x384x8 f32)", flush=True)
        ipt.transpose_variant(t, d1, d2, d3, ipt.VARIANT_GCS)
        torch.cuda.synchronize()
    finally:
        ipt.set_watchdog(0, 0)
    b = oracle.random_bits(300 * 200, 4, 5101)
    u = to_dev(b)
    with pytest.raises(ipt.TransposeError) as err:
        ipt.transpose(u, 300, 200, 1)  # reports the earlier trip, does nothing
    assert "watchdog" in str(err.value)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(u, b), b)  # untouched
    ipt.device_check()  # the state was restored by the reporting call
    for k, (e1, e2, e3, ez) in enumerate([(512, 384, 8, 4), (256, 1000, 3, 8), (300, 200, 1, 4)]):
        c = oracle.random_bits(e1 * e2 * e3, ez, 5110 + k)
        w = to_dev(c)
        ipt.transpose_variant(w, e1, e2, e3, ipt.VARIANT_GCS)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(w, c), oracle.seq_tt_213(c, e1, e2, e3))
    ipt.device_check()


@pytest.mark.parametrize("which,case", [
    ("sync", (512, 384, 8, 4, "gcs", 0)),         # group barriers in the barrier arena
    ("scratch", (300, 100, 2, 4, "c2r", 256)),    # scratch tiers forced by a 256-byte LDS cap
])
def test_captured_arena_outlives_growth(ipt, oracle, which, case):
    d1, d2, d3, esz, name, lds = case
    arena = ipt.ARENA_SYNC if which == "sync" else ipt.ARENA_SCRATCH
    v = ipt.VARIANT_GCS if name == "gcs" else ipt.VARIANT_C2R
    tune = ipt.Tune(lds_bytes=lds) if lds else None
    print(f"[case] arena {which}: {name} {d1}x{d2}x{d3} esz={esz}", flush=True)
    a = oracle.random_bits(d1 * d2 * d3, esz, 5200 + lds)
    t = to_dev(a)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ipt.transpose_variant(t, d1, d2, d3, v, tune=tune)  # warm-up outside capture
    s.synchronize()
    exp = oracle.seq_tt_213(a, d1, d2, d3)
    assert np.array_equal(to_host(t, a), exp)
    ptr0, bytes0, ret0 = ipt.arena_info(arena)
    assert ptr0 and bytes0 > 0
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        ipt.transpose_variant(t, d1, d2, d3, v, tune=tune)
    assert ipt.arena_info(arena)[0] == ptr0  # the graph holds this buffer
    ipt.arena_grow(arena, bytes0 + 1)
    ptr1, bytes1, ret1 = ipt.arena_info(arena)
    # host checks before any replay: a new buffer, the captured one retired
    # and still allocated
    assert ptr1 != ptr0 and bytes1 > bytes0 and ret1 == ret0 + 1
    assert ipt.memory_kind(ptr0)[0] == ipt.MEM_DEVICE
    side = torch.zeros(1 << 20, device="cuda")
    for r in range(3):
        g.replay()
        torch.cuda.synchronize()
        exp = oracle.seq_tt_213(exp, d1, d2, d3)
        assert np.array_equal(to_host(t, a), exp), r
        side.add_(1.0)  # eager work between replays
        ipt.transpose_variant(t, d1, d2, d3, v, tune=tune)  # eager use of the new buffer
        torch.cuda.synchronize()
        exp = oracle.seq_tt_213(exp, d1, d2, d3)
        assert np.array_equal(to_host(t, a), exp), r
    del g
    ipt.device_check()
