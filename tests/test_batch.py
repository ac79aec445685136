"""Batched uneven transposes: one persistent launch (+ graph capture) against
the oracle, per instance, bit-exact."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def carve(total_elems_bytes):
    return torch.empty(total_elems_bytes // 8 + 64, dtype=torch.int64, device="cuda")


def make_instances(oracle, shapes, seed):
    """One device buffer carved into 256-byte aligned instances."""
    offs, nbytes = [], 0
    for d1, d2, d3, esz in shapes:
        offs.append(nbytes)
        nbytes += (d1 * d2 * d3 * esz + 255) // 256 * 256
    buf = torch.empty(nbytes // 8 + 1, dtype=torch.int64, device="cuda")
    raw = buf.view(torch.uint8)
    items, host = [], []
    for k, ((d1, d2, d3, esz), off) in enumerate(zip(shapes, offs)):
        n = d1 * d2 * d3
        a = oracle.random_bits(n, esz, seed + k)
        t = raw[off:off + n * esz].view(torch.int32 if esz == 4 else torch.int64)
        t.copy_(torch.from_numpy(a.view(np.int32 if esz == 4 else np.int64)))
        items.append((t, d1, d2, d3))
        host.append(a)
    return buf, items, host


def check(oracle, items, host, times):
    for (t, d1, d2, d3), a in zip(items, host):
        exp = a
        for _ in range(times):
            exp = oracle.seq_tt_213(exp, d1, d2, d3)
        got = t.cpu().numpy().view(a.dtype)
        assert np.array_equal(got, exp), (d1, d2, d3, a.itemsize)


SHAPES = [  # (d1, d2, d3, esz): slice-in-LDS, GCS, square-swap and big routes, plus no-ops
    (1, 1, 1, 4), (7, 1, 3, 8), (3, 5, 7, 4), (33, 17, 4, 8), (100, 30, 2, 4), (64, 64, 3, 8),
    (512, 384, 2, 4), (256, 1000, 1, 8), (1000, 999, 1, 4), (3000, 1000, 1, 4), (97, 203, 5, 8),
    (1024, 1024, 1, 4), (2048, 2048, 1, 8), (4096, 40, 1, 4), (40, 4096, 1, 8), (129, 131, 1, 4),
    (2, 3000, 2, 8), (3000, 2, 1, 4), (700, 700, 2, 4), (255, 257, 1, 8), (4096, 4096, 1, 4),
    (129, 129, 2, 4), (333, 333, 1, 8), (1001, 1001, 1, 4),
]


@pytest.mark.gpu
def test_batch_eager_and_graph_replays(ipt, oracle):
    buf, items, host = make_instances(oracle, SHAPES, 900)
    # AUTO everywhere except the 2048^2 f64 instance, forced onto the tile-pair
    # route (which AUTO takes depends on the shipped selector's ranking)
    v = [ipt.VARIANT_SQUARE_SWAP if s == (2048, 2048, 1, 8) else ipt.VARIANT_AUTO for s in SHAPES]
    b = ipt.Batch([it + (vv,) for it, vv in zip(items, v)])
    info = b.info()
    print("[batch]", info, flush=True)
    assert info["n_slice_items"] > 0 and info["n_gcs_items"] > 0 and info["n_swap_items"] > 0
    assert info["skipped"] == 2
    b.run(graph=False)
    torch.cuda.synchronize()
    b.check()
    check(oracle, items, host, 1)
    # graph: captured on first run, replayed; eager work and syncs in between
    b.run(graph=True)
    torch.cuda.synchronize()
    check(oracle, items, host, 2)
    scratch = torch.zeros(1 << 20, device="cuda")
    scratch += 1
    b.run(graph=True)
    torch.cuda.synchronize()
    check(oracle, items, host, 3)
    b.run(graph=True)
    b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    check(oracle, items, host, 5)
    b.close()


@pytest.mark.gpu
def test_batch_forced_variants_and_small_lds(ipt, oracle):
    shapes = [(300, 100, 2, 4), (100, 300, 1, 8), (512, 512, 1, 4), (768, 512, 1, 4)]
    variants = [ipt.VARIANT_GCS, ipt.VARIANT_R2C, ipt.VARIANT_SQUARE_SWAP, ipt.VARIANT_C2R]
    buf, items, host = make_instances(oracle, shapes, 950)
    items = [it + (v,) for it, v in zip(items, variants)]
    b = ipt.Batch(items, tune=ipt.Tune(lds_bytes=65536))
    info = b.info()
    assert info["n_big"] == 2 and info["n_gcs_items"] > 0 and info["n_swap_items"] > 0
    b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    check(oracle, [it[:4] for it in items], host, 1)
    b.close()


@pytest.mark.gpu
def test_batch_block_cycle_instances(ipt, oracle):
    """Instances on the block-cycle path inside a batch: chunk steps before
    (chunked columns) and after (chunked rows) the persistent kernel, blocks as
    its items; a shape the path cannot take falls back to C2R / R2C."""
    BC, AUTO = ipt.VARIANT_BLOCK_CYCLE, ipt.VARIANT_AUTO
    cases = [((65536, 8, 2, 4), BC), ((8, 65536, 1, 8), BC), ((6144, 2048, 1, 4), BC),
             ((2048, 6144, 1, 4), BC), ((4096, 96, 3, 8), BC), ((96, 4096, 2, 4), AUTO),
             ((1009, 3, 1, 4), BC), ((300, 200, 1, 4), AUTO)]
    buf, items, host = make_instances(oracle, [c for c, _ in cases], 1200)
    b = ipt.Batch([it + (v,) for it, (_, v) in zip(items, cases)])
    info = b.info()
    print("[batch block-cycle]", info, flush=True)
    assert info["n_block_cycle"] >= 5 and info["n_big"] >= 1
    b.run(graph=False)
    torch.cuda.synchronize()
    check(oracle, items, host, 1)
    b.run(graph=True)
    torch.cuda.synchronize()
    check(oracle, items, host, 2)
    scratch = torch.zeros(1 << 20, device="cuda")
    scratch += 1
    b.run(graph=True)
    b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    check(oracle, items, host, 4)
    b.close()


@pytest.mark.gpu
def test_batch_heuristic_routes_oversized_instances(ipt, oracle):
    """With the selector off, instances too large for one cooperative group
    of the persistent kernel: f64 3970 x 2975 (94 MB) and 3751 x 3723 (112
    MB) fit one group of the per-call kernel and run as separate cooperative
    launches; f64 3900 x 4096 (128 MB) fits neither and takes the block-cycle
    route with its blocks in the batch; none falls to a separate C2R / R2C
    launch.  Beside ordinary instances, every instance exact after an eager
    run and a graph replay."""
    shapes = [(3900, 4096, 1, 8), (3751, 3723, 1, 8), (3970, 2975, 1, 8), (512, 384, 2, 4),
              (33, 17, 4, 8)]
    buf, items, host = make_instances(oracle, shapes, 1300)
    ipt.selector_load(None)
    try:
        b = ipt.Batch(items)
    finally:
        ipt.load_default_selector()
    routes = list(b.routes())
    print("[batch heuristic oversized]", b.info(), routes, flush=True)
    assert routes[:3] == [ipt.ROUTE_BLOCK_CYCLE, ipt.ROUTE_GCS_LAUNCH, ipt.ROUTE_GCS_LAUNCH]
    assert ipt.ROUTE_SEPARATE not in routes
    b.run(graph=False)
    torch.cuda.synchronize()
    check(oracle, items, host, 1)
    b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    check(oracle, items, host, 2)
    b.close()


@pytest.mark.gpu
def test_tiny_batches(ipt, oracle):
    """Batches with fewer work items than twice the XCD count run one shared
    queue (and a grid of a few workgroups)."""
    for k, shapes in enumerate([[(300, 200, 1, 4)], [(300, 200, 1, 8), (5, 7, 1, 4)],
                                [(64, 64, 1, 4)], [(200, 300, 2, 4)]]):
        buf, items, host = make_instances(oracle, shapes, 1300 + 10 * k)
        b = ipt.Batch(items)
        info = b.info()
        print("[tiny batch]", shapes, info, flush=True)
        b.run(graph=False)
        torch.cuda.synchronize()
        check(oracle, items, host, 1)
        b.run(graph=True)
        b.run(graph=True)
        torch.cuda.synchronize()
        b.check()
        check(oracle, items, host, 3)
        b.close()


@pytest.mark.gpu
def test_uneven_workload_sample(ipt, oracle):
    """A 300-instance sample of the uneven-batch workload (m, n log-uniform in
    [2, 1024]), one graph launch, every instance checked."""
    shapes = [(n, m, 1, e) for m, n, e in ipt.uneven_shapes(300, 2, 1024, seed=3)]
    buf, items, host = make_instances(oracle, shapes, 1000)
    b = ipt.Batch(items)
    b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    check(oracle, items, host, 1)
    b.close()


def test_batch_rejects_bad_descriptors(ipt):
    import ctypes
    L = ipt.lib()
    descs = (ipt.MatrixDesc * 1)()
    descs[0] = ipt.MatrixDesc(None, 4, 0, 4, 4, 1)
    h = ctypes.c_void_p()
    assert L.ipt_batch_create(descs, 1, None, ctypes.byref(h)) == 1
    descs[0] = ipt.MatrixDesc(256, 3, 0, 4, 4, 1)
    assert L.ipt_batch_create(descs, 1, None, ctypes.byref(h)) == 1
    descs[0] = ipt.MatrixDesc(256, 4, 9, 4, 4, 1)
    assert L.ipt_batch_create(descs, 1, None, ctypes.byref(h)) == 1
    assert L.ipt_batch_create(descs, -1, None, ctypes.byref(h)) == 1
    assert L.ipt_batch_run(None, None) == 1
    assert L.ipt_batch_destroy(None) == 1


@pytest.mark.gpu
def test_uneven_workload_full_4096(ipt, oracle):
    """The uneven-batch workload at its stated size: 4096 matrices, m and n
    log-uniform in [2, 4096] (seed 4, as the benchmark), 4- and 8-byte
    elements, 7.57 GB of seeded random bits.  Every instance is checked
    against the oracle after one eager run and again after two graph
    replays (the oracle applied three times)."""
    shapes = [(n, m, 1, e) for m, n, e in ipt.uneven_shapes(4096, 2, 4096, seed=4)]
    offs, nbytes = [], 0
    for d1, d2, d3, esz in shapes:
        offs.append(nbytes)
        nbytes += (d1 * d2 * d3 * esz + 255) // 256 * 256
    print(f"[case] uneven batch 4096 instances, {nbytes / 1e9:.2f} GB", flush=True)
    host = np.empty(nbytes, dtype=np.uint8)
    views = []
    for k, ((d1, d2, d3, esz), off) in enumerate(zip(shapes, offs)):
        n = d1 * d2 * d3
        a = oracle.random_bits(n, esz, 40000 + k)
        host[off:off + n * esz] = a.view(np.uint8)
        views.append((off, n, esz))
    buf = torch.from_numpy(host).cuda()
    items = []
    for (d1, d2, d3, esz), (off, n, _) in zip(shapes, views):
        items.append((buf[off:off + n * esz].view(torch.int32 if esz == 4 else torch.int64),
                      d1, d2, d3))
    b = ipt.Batch(items)
    info = b.info()
    print("[batch 4096]", info, flush=True)
    assert info["count"] == 4096
    assert ipt.ROUTE_SEPARATE not in set(b.routes())  # no C2R / R2C fallbacks

    def check(times):
        got_all = buf.cpu().numpy()
        bad = 0
        for (d1, d2, d3, esz), (off, n, _) in zip(shapes, views):
            dt = np.uint32 if esz == 4 else np.uint64
            exp = host[off:off + n * esz].view(dt)
            for _ in range(times):
                exp = oracle.seq_tt_213(exp, d1, d2, d3)
            bad += not np.array_equal(got_all[off:off + n * esz].view(dt), exp)
        return bad

    b.run(graph=False)
    torch.cuda.synchronize()
    b.check()
    assert check(1) == 0
    print("[batch 4096] eager run exact", flush=True)
    b.run(graph=True)
    filler = torch.zeros(1 << 20, device="cuda")
    filler += 1  # eager work between the replays
    b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    assert check(3) == 0
    print("[batch 4096] two graph replays exact", flush=True)
    b.close()


@pytest.mark.gpu
def test_batch_rows_off_the_line_grid(ipt, oracle):
    """Cooperative instances whose rows (or bases) are off the 128-byte line
    grid, where the persistent kernel loads whole lines and realigns them in
    registers: row pitches at every 4- / 8-byte offset class, bases off by
    1..31 elements, several slices (each slice's base has its own offset),
    members that straddle bands, narrow slices (fewer than 8 tile rows), and
    line-aligned neighbours; every instance exact after an eager run and two
    graph replays."""
    rng = np.random.default_rng(20261204)
    shapes, offs = [], []
    for k in range(40):
        esz = 4 if k % 2 == 0 else 8
        d1 = int(rng.integers(300, 3000))
        d2 = int(rng.integers(40, 1600)) if k % 5 else int(rng.integers(33, 200))
        d3 = int(rng.integers(1, 3))
        if k % 7 == 0:  # line-aligned pitch, base off the grid
            d1 = (d1 // (128 // esz) + 1) * (128 // esz)
        shapes.append((d1, d2, d3, esz))
        offs.append(int(rng.integers(0, 32)) if k % 3 else 0)
    nbytes = sum((d1 * d2 * d3 * e + o * e + 255) // 256 * 256 for (d1, d2, d3, e), o in zip(shapes, offs))
    buf = torch.empty(nbytes // 8 + 1, dtype=torch.int64, device="cuda")
    raw = buf.view(torch.uint8)
    items, host, pos = [], [], 0
    for k, ((d1, d2, d3, esz), o) in enumerate(zip(shapes, offs)):
        n = d1 * d2 * d3
        a = oracle.random_bits(n, esz, 1500 + k)
        start = pos + o * esz
        t = raw[start:start + n * esz].view(torch.int32 if esz == 4 else torch.int64)
        t.copy_(torch.from_numpy(a.view(np.int32 if esz == 4 else np.int64)))
        items.append((t, d1, d2, d3, ipt.VARIANT_GCS))
        host.append(a)
        pos += (n * esz + o * esz + 255) // 256 * 256
    b = ipt.Batch(items)
    print("[batch off-grid]", b.info(), flush=True)
    plain = [it[:4] for it in items]
    b.run(graph=False)
    torch.cuda.synchronize()
    b.check()
    check(oracle, plain, host, 1)
    b.run(graph=True)
    b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    check(oracle, plain, host, 3)
    b.close()


@pytest.mark.gpu
def test_batch_barrier_watchdog_reports_and_recovers(ipt, oracle):
    """The batch kernel's group barriers (per-member flags and dependency
    ranges) give up through the watchdog hook (every member waits for an
    arrival that never comes for 2000 polls, then proceeds, so every kernel
    exits normally): ipt_batch_check reports it and restores the barrier
    state, and the next run is exact again."""
    shapes = [(1500, 1000, 1, 4), (777, 1023, 1, 8), (300, 2000, 2, 4), (40, 50, 3, 4)]
    buf, items, host = make_instances(oracle, shapes, 1300)
    b = ipt.Batch(items)
    info = b.info()
    print("[batch]", info, flush=True)
    assert info["n_gcs_items"] > 0
    b.set_watchdog(2000, 1)
    print("[case] tripping the batch barriers on purpose", flush=True)
    b.run(graph=False)
    with pytest.raises(ipt.TransposeError) as err:
        b.check()
    assert "watchdog" in str(err.value)
    b.check()  # state restored, settings cleared
    buf2, items2, host2 = make_instances(oracle, shapes, 1400)
    b2 = ipt.Batch(items2)
    b2.run(graph=False)
    torch.cuda.synchronize()
    b2.check()
    check(oracle, items2, host2, 1)
    # the tripped batch itself runs normally again on fresh data
    for (t, d1, d2, d3), a in zip(items, host):
        t.copy_(torch.from_numpy(a.view(np.int32 if a.itemsize == 4 else np.int64)).cuda())
    b.run(graph=False)
    torch.cuda.synchronize()
    b.check()
    check(oracle, items, host, 1)
    b.close()
    b2.close()


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [11, 12])
def test_uneven_batches_other_seeds(ipt, oracle, seed):
    """Uneven batches drawn with other seeds (800 matrices, m and n
    log-uniform in [2, 3000]): cooperative groups of every size from one
    member to the whole grid, each member waiting only for the members its
    stores depend on; every instance exact after an eager run and a graph
    replay."""
    shapes = [(n, m, 1, e) for m, n, e in ipt.uneven_shapes(800, 2, 3000, seed=seed)]
    print(f"[case] uneven batch seed {seed}: "
          f"{sum(a * b * c * e for a, b, c, e in shapes) / 1e9:.2f} GB", flush=True)
    buf, items, host = make_instances(oracle, shapes, 20000 + 1000 * seed)
    b = ipt.Batch(items)
    print("[batch]", b.info(), flush=True)
    b.run(graph=False)
    torch.cuda.synchronize()
    b.check()
    check(oracle, items, host, 1)
    b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    check(oracle, items, host, 2)
    b.close()
