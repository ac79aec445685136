"""Benchmark of the in-place transpose engine on one MI355X.

Headline workload (one step = one launch of the planned batch's captured
hipGraph): the uneven batch -- 4096 row-major matrices, m and n log-uniform in
[2, 4096] (seed 4), 4- or 8-byte elements (p = 1/2), 7.57 GB of payload,
resident in HBM (well above the 256 MB Infinity Cache).  Every matrix is
transposed in place; value = 2 * payload bytes (each element read once and
written once) / step time, in GB/s.

The JSON line also carries:
  speedup       t(plain per-call heuristic loop: one ipt_transpose_213 per
                matrix) / t(learned selector + batched graph), the metric's
                second part, split into the batching gain (heuristic per-call
                loop vs heuristic batch) and the learned gain (heuristic batch
                vs learned batch).
  verified      every matrix checked after an untimed run: iota input, output
                compared on the device with the closed form (the iota matrix
                transposed by torch).
  roofline      the batch's persistent kernel k_batch: algorithmic bytes per
                launch (2 * the bytes of the instances whose work it does: it
                reads and writes each of their elements once; the few
                instances run as separate launches are excluded) / its
                duration, from HIP events its own launch records on the run's
                stream (eager runs, median of 5), against the 8 TB/s HBM
                peak.  `traffic` is FETCH_SIZE + WRITE_SIZE of k_batch per
                launch from two child `rocprofv3 --pmc` runs of
                `bench.py --profile-batch`, each scaled by what the counter
                reads per byte for the plain copy kernel copying the whole
                payload in the same run.
  copy_GBps     a plain device copy of the whole payload (one launch).
  rows          the reference's single-tensor configs, each transposed per
                call with the default dispatch (median of event-timed blocks),
                each with its residency (MALL when it fits the 256 MB Infinity
                Cache: such rates are cache-bound and are not reported as HBM
                results), its own roofline, and the out-of-place comparator's
                rate on the same shape (ipt_transpose_oop).
  cpu_baseline  the oracle's OpenMP C2R/R2C path (a port: the reference's
                OpenMP transpose is declared but never implemented) over the
                same 4096 matrices on this host's cores; beside it (cpu_product)
                the library's own host path (ipt_cpu_transpose), matrices
                spread over the same number of threads.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --configs [--only 1,3c]     every config shape, checked
                                                    against its gen_ans digest
       python bench.py --profile-batch | --profile-config NAME | --profile
                                                    (sequences for rocprofv3)
       python bench.py --batch-routes | --move-lists   (breakdowns)
"""
import argparse
import concurrent.futures
import glob
import hashlib
import importlib
import json
import math
import os
import sqlite3
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
MALL_BYTES = 256 << 20
METRIC = ("in-place GB/s (2×bytes/t, % of 8TB/s HBM); learned+batched vs heuristic "
          "speedup")
WORKLOAD = [  # (d1, d2, d3, esz): the 213 permutes of --profile (2a and 2b)
    (512, 384, 64, 4),
    (256, 1000, 17, 8),
]
BATCH_COUNT, BATCH_SEED, BATCH_HI = 4096, 4, 4096


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE child runs")
    ap.add_argument("--no-rows", action="store_true", help="skip the single-config rows")
    ap.add_argument("--cands", default=None,
                    help="--profile: selector candidate per workload tensor, comma-separated")
    ap.add_argument("--profile-batch", action="store_true",
                    help="calibration copy of the payload, then the uneven batch's graph "
                         "warmup + steps times (for a profiler)")
    ap.add_argument("--profile", action="store_true",
                    help="2a and 2b: calibration copies, then each tensor transposed back to "
                         "back (for profilers)")
    ap.add_argument("--profile-config", default=None, metavar="NAME",
                    help="one reference config shape (1, 2a, ..., 3d): calibration copies of "
                         "its bytes, then warmup + steps transposes with the default dispatch "
                         "(for profilers)")
    ap.add_argument("--profile-uniform", default=None, metavar="D1xD2xESZ",
                    help="a batch of copies of one row-major shape (d1 columns, d2 rows, esz "
                         "bytes; about 4 GB in all), run eagerly warmup + steps times (for "
                         "profilers: row alignment against traffic counters)")
    ap.add_argument("--batch-routes", action="store_true",
                    help="the uneven batch's time by route and by slice size: the workload's "
                         "instances of each route run as a batch of their own")
    ap.add_argument("--move-lists", action="store_true",
                    help="the block-cycle chunk step's move lists at the plan limit (2^23 chunk "
                         "positions): the device builder against the host walk")
    ap.add_argument("--configs", action="store_true",
                    help="time every reference config shape with the default dispatch (one JSON line)")
    ap.add_argument("--only", default=None,
                    help="--configs: comma-separated config names to run (default: all)")
    return ap.parse_args()


CONFIGS = [  # (name, (d1, d2, d3, esz)): the reference's config shapes
    ("1", (3000, 1000, 1, 4)), ("2a", (512, 384, 64, 4)), ("2b", (256, 1000, 17, 8)),
    ("3a", (16777216, 32, 1, 8)), ("3b", (7, 4194304, 1, 4)), ("3c", (4099, 4093, 1, 4)),
    ("3d", (16384, 16384, 1, 4)),
]


def configs_main(args, torch, ipt):
    """Every config shape, iota input: one transpose checked against the
    gen_ans digest and undone; then, for the transpose and for its inverse
    apart, 3 warm-up calls and the median over 5 event-timed blocks of 10
    back-to-back calls; the same for the out-of-place comparator
    (ipt_transpose_oop into a second buffer, its output checked against the
    same digest).  The work of an in-place permutation does not depend
    on the values, so repeated calls time the same work."""
    import numpy as np
    out = []
    only = set(args.only.split(",")) if args.only else None
    for name, (d1, d2, d3, esz) in CONFIGS:
        if only and name not in only:
            continue
        dt = torch.int32 if esz == 4 else torch.int64
        t = torch.arange(d1 * d2 * d3, dtype=dt, device="cuda")
        ipt.transpose(t, d1, d2, d3)
        torch.cuda.synchronize()
        sha = hashlib.sha256(t.cpu().numpy().tobytes()).hexdigest()
        ok = sha == golden_sha((d1, d2, d3, esz))
        ipt.transpose(t, d2, d1, d3)

        def per_call(a, b, call=lambda a, b: ipt.transpose(t, a, b, d3)):
            for _ in range(3):
                call(a, b)
            blocks = []
            for _ in range(5):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(10):
                    call(a, b)
                e1.record()
                e1.synchronize()
                blocks.append(e0.elapsed_time(e1) / 10)
            return float(np.median(blocks))
        ms = per_call(d1, d2)
        ms_inv = per_call(d2, d1)
        # the out-of-place comparator on the same shape, from the iota input
        t.copy_(torch.arange(d1 * d2 * d3, dtype=dt, device="cuda"))
        dst = torch.empty_like(t)
        ipt.transpose_oop(dst, t, d1, d2, d3)
        torch.cuda.synchronize()
        oop_ok = hashlib.sha256(dst.cpu().numpy().tobytes()).hexdigest() == golden_sha(
            (d1, d2, d3, esz))
        ms_oop = per_call(d1, d2, lambda a, b: ipt.transpose_oop(dst, t, a, b, d3))
        del dst
        ipt.device_check()
        nb = 2 * d1 * d2 * d3 * esz
        k = ipt.select_candidate(esz, d1, d2, d3)
        out.append({"config": name, "shape": [d1, d2, d3, esz], "verified_vs_gen_ans": ok,
                    "variant": ipt.VARIANT_NAMES[ipt.candidate(k)[0]], "ms": round(ms, 4),
                    "GBps": round(nb / (ms * 1e-3) / 1e9, 1),
                    "frac_of_hbm_peak": round(nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "inverse_variant": ipt.VARIANT_NAMES[ipt.candidate(
                        ipt.select_candidate(esz, d2, d1, d3))[0]],
                    "inverse_GBps": round(nb / (ms_inv * 1e-3) / 1e9, 1),
                    "oop_GBps": round(nb / (ms_oop * 1e-3) / 1e9, 1), "oop_verified": oop_ok})
        print(f"[config {name}] {out[-1]}", file=sys.stderr, flush=True)
        del t
        torch.cuda.empty_cache()
    print(json.dumps({"configs": out, "timing": "median of 5 HIP-event-timed blocks of 10 "
                                                "back-to-back calls, default dispatch"}))


def profile_config_main(args, torch, ipt):
    """Calibration copies of the config's bytes, then warmup + steps
    transposes of it with the default dispatch, all with the same dims (the
    work of an in-place permutation does not depend on the values, so every
    call does the same work; runs under rocprofv3)."""
    d1, d2, d3, esz = dict(CONFIGS)[args.profile_config]
    dt = torch.int32 if esz == 4 else torch.int64
    t = torch.arange(d1 * d2 * d3, dtype=dt, device="cuda")
    raw = t.view(torch.uint8)[:t.numel() * esz // 16 * 16]  # the copy moves whole 16-byte words
    sink = torch.empty_like(raw)
    for _ in range(args.warmup + args.steps):
        ipt.copy(sink, raw)
    torch.cuda.synchronize()
    del sink, raw
    torch.cuda.empty_cache()
    for _ in range(args.warmup + args.steps):
        ipt.transpose(t, d1, d2, d3)
    torch.cuda.synchronize()
    ipt.device_check()
    print(f"profile run done: config {args.profile_config} {d1}x{d2}x{d3}x{esz}", flush=True)


def golden_sha(shape):
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        for c in json.load(f)["cases"]:
            if (c["d1"], c["d2"], c["d3"], c["esz"]) == shape:
                return c["sha256"]
    return None


def time_calls(torch, fn, tensor, reps, blocks=5):
    """Mean time per call (ms) of `reps` back-to-back calls, timed with
    events on the launch stream; the median over `blocks` such blocks, so a
    single hiccup of the box does not set the figure."""
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    for _ in range(2):
        fn(tensor)
    torch.cuda.synchronize()
    per_block = []
    for _ in range(blocks):
        start.record(stream)
        for _ in range(reps):
            fn(tensor)
        end.record(stream)
        end.synchronize()
        per_block.append(start.elapsed_time(end) / reps)
    return sorted(per_block)[len(per_block) // 2]


def make_tensors(torch):
    tensors = []
    for d1, d2, d3, esz in WORKLOAD:
        dt = torch.int32 if esz == 4 else torch.int64
        tensors.append(torch.arange(d1 * d2 * d3, dtype=dt, device="cuda"))
    return tensors


def profile_main(args, torch, ipt):
    """The 2a + 2b kernels' measurement sequence and nothing else (runs
    under rocprofv3): calibration copies of both tensors first, then each
    tensor transposed back to back on its own.  --cands picks the selector
    candidate per tensor; without it the library's dispatch decides."""
    tensors = make_tensors(torch)
    sinks = [torch.empty_like(t) for t in tensors]
    for _ in range(args.warmup + args.steps):
        for t, s in zip(tensors, sinks):
            ipt.copy(s, t)
    torch.cuda.synchronize()
    del sinks
    torch.cuda.empty_cache()
    cands = [int(x) for x in args.cands.split(",")] if args.cands else [None] * len(tensors)
    for t, (d1, d2, d3, esz), k in zip(tensors, WORKLOAD, cands):
        if k is None:
            fn = lambda: ipt.transpose(t, d1, d2, d3)
        else:
            v, tune = ipt.candidate(k)
            fn = lambda: ipt.transpose_variant(t, d1, d2, d3, v, tune=tune)
        for _ in range(args.warmup + args.steps):
            fn()
        torch.cuda.synchronize()
    print("profile run done", flush=True)


def pmc_per_dispatch(db, counter):
    """{kernel name: [counter value of each dispatch, in dispatch order]}
    from a rocpd database."""
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, dispatch_id, sum(value) from counters_collection "
                       "where counter_name = ? group by kernel_name, dispatch_id "
                       "order by dispatch_id", (counter,))
    acc = {}
    for name, _, val in rows:
        acc.setdefault(name, []).append(val)
    return acc


def measure_traffic(match, copy_bytes, child_args):
    """Bytes per launch of the kernel whose name contains every string in
    `match`, from FETCH_SIZE and WRITE_SIZE in two child rocprofv3 runs of
    this script with `child_args`; each counter is scaled per byte by the
    copy kernel's dispatches of the same run, which copy `copy_bytes` each."""
    totals = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="ipt_pmc_")
        cmd = ["timeout", "-s", "KILL", "240", "rocprofv3", "--pmc", counter, "-d", out, "-o",
               "pmc", "--", sys.executable, os.path.abspath(__file__)] + child_args
        r = subprocess.run(cmd, cwd=tempfile.gettempdir(), capture_output=True, text=True)
        dbs = glob.glob(os.path.join(out, "**", "*.db"), recursive=True)
        if r.returncode not in (0, 139) or not dbs:
            return None, f"rocprofv3 --pmc {counter} failed (rc={r.returncode})"
        per = pmc_per_dispatch(dbs[0], counter)
        kern = [v for k, v in per.items() if all(part in k for part in match)]
        copies = [v for k, v in per.items() if "k_copy16" in k]
        if not copies or not kern:
            return None, f"{counter}: kernel not found in profile"
        kern_mean = sum(kern[0]) / len(kern[0])
        copy_mean = sum(copies[0]) / len(copies[0])
        totals[counter] = kern_mean * (copy_bytes / copy_mean)
    return totals, None


def uneven_items(torch, ipt, count=BATCH_COUNT, seed=BATCH_SEED):
    """The uneven-batch workload: (tensor, d1, d2, d3) views into one buffer,
    every instance 256-byte aligned."""
    shapes = [(n, m, 1, e) for m, n, e in ipt.uneven_shapes(count, 2, BATCH_HI, seed=seed)]
    offs, total = [], 0
    for d1, d2, d3, esz in shapes:
        offs.append(total)
        total += (d1 * d2 * d3 * esz + 255) // 256 * 256
    buf = torch.empty(total // 8 + 1, dtype=torch.int64, device="cuda")
    raw = buf.view(torch.uint8)
    items = []
    for (d1, d2, d3, esz), off in zip(shapes, offs):
        n = d1 * d2 * d3
        items.append((raw[off:off + n * esz].view(torch.int32 if esz == 4 else torch.int64),
                      d1, d2, d3))
    payload = sum(d1 * d2 * d3 * esz for d1, d2, d3, esz in shapes)
    return buf, items, payload


def profile_uniform_main(args, torch, ipt):
    """A batch of identical instances (one shape, 256-byte aligned, about
    4 GB in all), run eagerly warmup + steps times: each run is one k_batch
    dispatch over the same work (runs under rocprofv3)."""
    d1, d2, esz = (int(x) for x in args.profile_uniform.lower().split("x"))
    n = max(1, int(4e9 // (d1 * d2 * esz)))
    pitch = (d1 * d2 * esz + 255) // 256 * 256
    buf = torch.zeros(n * pitch // 8 + 1, dtype=torch.int64, device="cuda")
    raw = buf.view(torch.uint8)
    dt = torch.int32 if esz == 4 else torch.int64
    items = [(raw[i * pitch:i * pitch + d1 * d2 * esz].view(dt), d1, d2, 1) for i in range(n)]
    b = ipt.Batch(items)
    for _ in range(args.warmup + args.steps):
        b.run(graph=False)
    torch.cuda.synchronize()
    b.check()
    b.close()
    print(f"profile run done: {n} x {d1}x{d2}x{esz}", flush=True)


def profile_batch_main(args, torch, ipt):
    """Calibration copies of the whole payload (one k_copy16 launch each),
    then the uneven batch's graph, run warmup + steps times (runs under
    rocprofv3)."""
    buf, items, payload = uneven_items(torch, ipt)
    raw = buf.view(torch.uint8)[:payload // 16 * 16]
    sink = torch.empty_like(raw)
    for _ in range(2):
        ipt.copy(sink, raw)
    torch.cuda.synchronize()
    del sink
    torch.cuda.empty_cache()
    b = ipt.Batch(items)
    for _ in range(args.warmup + args.steps):
        b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    b.close()
    print("profile run done", flush=True)


MOVE_LISTS = [(32, 262143, 16), (7, 1198371, 64), (2896, 2896, 128), (32, 8192, 16)]


def move_lists_main(args, torch, ipt):
    """Build the chunk step's move lists of each (rows, cols, segment) chunk
    matrix with the device builder and with the host walk (ipt_debug_chunk_moves:
    build, copy the lists to the host, free), median of 3 after one warm-up
    build of each; one JSON line."""
    import numpy as np
    out = []
    ipt.chunk_moves(32, 512, 16)
    ipt.chunk_moves(32, 512, 16, device=False)
    for rc, cc, ell in MOVE_LISTS:
        row = {"rows": rc, "cols": cc, "positions": rc * cc, "segment": ell}
        for name, dev in (("device_ms", True), ("host_ms", False)):
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                lists = ipt.chunk_moves(rc, cc, ell, device=dev)
                ts.append((time.perf_counter() - t0) * 1e3)
            row[name] = round(float(np.median(ts)), 2)
            row["moves"], row["heads"] = len(lists[1]), len(lists[3])
        print(f"[move lists] {row}", file=sys.stderr, flush=True)
        out.append(row)
    print(json.dumps({"move_lists": out, "timing": "wall clock per build including the "
                      "copy of the lists to the host, median of 3"}))


def batch_routes_main(args, torch, ipt):
    """The uneven batch broken down: the instances of each route
    (ipt_batch_get_routes of the whole batch) as a batch of their own, the
    cooperative ones also by slice size; graph runs, median of 5 blocks of
    3.  One line per part, then a plain copy of the payload."""
    buf, items, payload = uneven_items(torch, ipt)
    shapes = [(d1, d2, d3, t.element_size()) for t, d1, d2, d3 in items]

    def graph_ms(b):
        b.run(graph=True)
        torch.cuda.synchronize()
        return time_calls(torch, lambda _: b.run(graph=True), None, 3)

    def part(name, idx):
        nb = sum(shapes[i][0] * shapes[i][1] * shapes[i][2] * shapes[i][3] for i in idx)
        b = ipt.Batch([items[i] for i in idx])
        ms = graph_ms(b)
        b.check()
        b.close()
        print(json.dumps({"part": name, "instances": len(idx), "bytes": nb, "ms": round(ms, 4),
                          "GBps": round(2 * nb / (ms * 1e-3) / 1e9, 1)}), flush=True)

    b = ipt.Batch(items)
    routes = b.routes().tolist()
    b.close()
    part("whole batch", list(range(len(items))))
    for r in sorted(set(routes)):
        idx = [i for i, x in enumerate(routes) if x == r]
        if r == ipt.ROUTE_SKIP:
            continue
        part(ipt.ROUTE_NAMES[r], idx)
        if r == ipt.ROUTE_GCS:
            edges = [0, 1 << 20, 4 << 20, 16 << 20, 48 << 20, 1 << 62]
            for lo, hi in zip(edges[:-1], edges[1:]):
                sub = [i for i in idx if lo <= shapes[i][0] * shapes[i][1] * shapes[i][3] < hi]
                if sub:
                    part(f"gcs slices [{lo >> 20}, {hi >> 20}) MB", sub)
    raw = buf.view(torch.uint8)[:payload // 16 * 16]
    sink = torch.empty_like(raw)
    ms = time_calls(torch, lambda _: ipt.copy(sink, raw), None, 3)
    print(json.dumps({"part": "plain copy of the payload", "bytes": raw.numel(), "ms": round(ms, 4),
                      "GBps": round(2 * raw.numel() / (ms * 1e-3) / 1e9, 1)}), flush=True)


def fill_iota(torch, items):
    for t, d1, d2, d3 in items:
        t.copy_(torch.arange(t.numel(), dtype=t.dtype, device="cuda"))


def verify_items(torch, items, times):
    """Every instance after `times` applications of its transpose to iota
    input equals the closed form (the iota matrix transposed by torch, on
    the device).  Returns the number of mismatching instances."""
    bad = 0
    for t, d1, d2, d3 in items:
        e = torch.arange(t.numel(), dtype=t.dtype, device="cuda")
        a, b = d1, d2
        for _ in range(times):
            e = e.view(d3, b, a).transpose(1, 2).contiguous().view(-1)
            a, b = b, a
        bad += not torch.equal(t, e)
    return bad


def batch_leg(torch, ipt, args):
    """The headline: per-call heuristic loop, heuristic batch, learned batch
    (graph); verification, k_batch roofline, plain copy of the payload."""
    buf, items, payload = uneven_items(torch, ipt)

    def per_call():
        for t, d1, d2, d3 in items:
            ipt.transpose(t, d1, d2, d3)

    def graph_ms(b, reps=3):
        b.run(graph=True)  # capture + first replay
        torch.cuda.synchronize()
        return time_calls(torch, lambda _: b.run(graph=True), None, reps)

    ipt.selector_load(None)  # plain heuristic
    ms_loop_h = time_calls(torch, lambda _: per_call(), None, 2)
    bh = ipt.Batch(items)
    ms_batch_h = graph_ms(bh)
    bh.check()
    bh.close()
    learned = ipt.load_default_selector()
    t0 = time.perf_counter()
    b = ipt.Batch(items)
    plan_ms = (time.perf_counter() - t0) * 1e3
    info = b.info()
    routes = b.routes()
    # bytes of the instances whose work k_batch does: the persistent-kernel
    # routes and the block steps of block-cycle instances (their chunk steps,
    # and the separate launches, are other kernels)
    inside = {ipt.ROUTE_SLICE, ipt.ROUTE_GCS, ipt.ROUTE_SWAP, ipt.ROUTE_BLOCK_CYCLE}
    kernel_payload = sum(t.numel() * t.element_size()
                         for (t, _, _, _), r in zip(items, routes.tolist()) if r in inside)

    # verified: one untimed graph run on iota input, every instance checked
    fill_iota(torch, items)
    b.run(graph=True)
    torch.cuda.synchronize()
    b.check()
    bad = verify_items(torch, items, 1)

    # timed steps: warmup, then exactly `steps` graph launches between syncs
    for _ in range(args.warmup):
        b.run(graph=True)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        b.run(graph=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms_step = ev0.elapsed_time(ev1) / args.steps
    b.check()

    # the persistent kernel alone: events its launch records (eager runs)
    b.set_timing(True)
    kms = []
    for _ in range(6):
        b.run(graph=False)
        kms.append(b.kernel_ms())
    kms = sorted(kms[1:])
    k_ms = kms[len(kms) // 2]
    b.check()
    b.close()

    raw = buf.view(torch.uint8)[:payload // 16 * 16]
    sink = torch.empty_like(raw)
    copy_ms = time_calls(torch, lambda _: ipt.copy(sink, raw), None, 3)
    copy_gbps = 2 * raw.numel() / (copy_ms * 1e-3) / 1e9
    del sink, raw, items, buf
    torch.cuda.empty_cache()
    by_route = {}
    for r in routes.tolist():
        name = ipt.ROUTE_NAMES[r]
        by_route[name] = by_route.get(name, 0) + 1
    return {
        "kernel_payload": kernel_payload,
        "ms_step": ms_step, "wall": wall, "payload": payload, "verified_bad": bad,
        "k_batch_ms": k_ms, "copy_GBps": copy_gbps, "plan_ms": plan_ms, "learned": learned,
        "ms_per_call_heuristic": ms_loop_h, "ms_batch_heuristic": ms_batch_h, "info": info,
        "routes": by_route,
    }


def row_leg(torch, ipt, name, shape):
    """One single-tensor config: default dispatch per call (median of 5
    event-timed blocks), checked once on the device against the closed form,
    plus the out-of-place comparator on the same shape."""
    d1, d2, d3, esz = shape
    dt = torch.int32 if esz == 4 else torch.int64
    n = d1 * d2 * d3
    t = torch.arange(n, dtype=dt, device="cuda")
    ipt.transpose(t, d1, d2, d3)
    e = torch.arange(n, dtype=dt, device="cuda").view(d3, d2, d1).transpose(1, 2).contiguous().view(-1)
    ok = bool(torch.equal(t, e))
    del e
    reps = 10 if n * esz < (64 << 20) else 3
    ms = time_calls(torch, lambda x: ipt.transpose(x, d1, d2, d3), t, reps)
    dst = torch.empty_like(t)
    oop_ms = time_calls(torch, lambda x: ipt.transpose_oop(dst, x, d1, d2, d3), t, reps)
    ipt.device_check()
    del dst, t
    torch.cuda.empty_cache()
    nb = 2 * n * esz
    gbps = nb / (ms * 1e-3) / 1e9
    v = ipt.candidate(ipt.select_candidate(esz, d1, d2, d3))[0]
    mall = n * esz <= MALL_BYTES
    sweeps = {ipt.VARIANT_BLOCK_CYCLE: 2, ipt.VARIANT_C2R: 3, ipt.VARIANT_R2C: 3}.get(v, 1)
    return {"config": name, "shape": [d1, d2, d3, esz], "verified": ok,
            "variant": ipt.VARIANT_NAMES[v], "ms": round(ms, 5), "GBps": round(gbps, 1),
            "residency": "MALL (fits the 256 MB Infinity Cache)" if mall else "HBM",
            "roofline": {"bound": "mall" if mall else "hbm", "achieved": round(gbps, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": None if mall else round(gbps / HBM_PEAK_GBS, 4),
                         "frac_of_hbm_peak_for_reference": round(gbps / HBM_PEAK_GBS, 4),
                         "sweeps": sweeps, "timing": "whole call, HIP events"},
            "oop_GBps": round(nb / (oop_ms * 1e-3) / 1e9, 1)}


def cpu_threads():
    """Threads for the CPU baselines: OMP_NUM_THREADS when set (the CPU share
    the host gives this job), else the cores this process may run on."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return (int(env) if env.isdigit() and int(env) > 0 else aff), aff


def cpu_legs(ipt):
    """The oracle's OpenMP path, and the library's host path, over the same
    4096 matrices (host arrays, iota bit patterns)."""
    import numpy as np
    import oracle_ctypes as oracle
    threads, aff = cpu_threads()
    shapes = [(n, m, 1, e) for m, n, e in ipt.uneven_shapes(BATCH_COUNT, 2, BATCH_HI,
                                                             seed=BATCH_SEED)]
    arrays = [np.arange(d1 * d2, dtype=np.uint32 if esz == 4 else np.uint64)
              for d1, d2, d3, esz in shapes]
    payload = sum(a.nbytes for a in arrays)

    def oracle_pass():
        for a, (d1, d2, d3, esz) in zip(arrays, shapes):
            oracle.transpose_213(a, d1, d2, d3, threads=threads)

    def product_pass():
        with concurrent.futures.ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda k: ipt.cpu_transpose_213(arrays[k], shapes[k][0], shapes[k][1], 1,
                                                        threads=1), range(len(arrays))))

    out = {}
    for key, fn in (("oracle", oracle_pass), ("product", product_pass)):
        times = []
        t_begin = time.perf_counter()
        while len(times) < 3 and (not times or time.perf_counter() - t_begin < 20.0):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        out[key] = (sorted(times)[len(times) // 2], len(times))
    nproc = os.cpu_count()
    med, n = out["oracle"]
    pmed, pn = out["product"]
    return {"value": round(2 * payload / med / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port",
            "sample": f"the whole uneven batch ({len(arrays)} matrices, {payload / 1e9:.2f} GB) transposed "
                      f"in place by the oracle's OpenMP C2R/R2C path, {threads} OpenMP threads per "
                      f"matrix, median of {n} passes ({med:.2f} s)",
            "nproc": nproc, "affinity_cores": aff,
            "cpu_product": {"value": round(2 * payload / pmed / 1e9, 3), "unit": "GB/s",
                            "cores": threads,
                            "sample": f"the same batch through the library's host path "
                                      f"(ipt_cpu_transpose_213), matrices spread over {threads} "
                                      f"threads, median of {pn} passes ({pmed:.2f} s)"}}


def selector_eval(ipt):
    """Held-out evaluation of the shipped selector on its own sweep data (CPU)."""
    path = os.path.join(ROOT, "213transpose_amd", "selector", "sweep.npz")
    if not os.path.exists(path) or not os.path.exists(ipt.SELECTOR_WEIGHTS):
        return None
    import numpy as np
    at = importlib.import_module("213transpose_amd.tuning")
    d = np.load(path)
    p = np.fromfile(ipt.SELECTOR_WEIGHTS, dtype=np.float32)
    res = at.evaluate_arrays(d["shapes"][-2000:], d["feats"][-2000:], d["times"][-2000:], p)
    return {k: round(v, 4) for k, v in res.items()}


ROWS = ["2a+2b", "1", "3c", "3d", "3a"]


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(slices and batch instances shard without exchange; "
                                   "multi-GPU runs are not part of this build)"}))
        sys.exit(2)
    import torch

    ipt = importlib.import_module("213transpose_amd")
    torch.cuda.set_device(0)
    if args.profile:
        profile_main(args, torch, ipt)
        return
    if args.profile_batch:
        profile_batch_main(args, torch, ipt)
        return
    if args.profile_config:
        profile_config_main(args, torch, ipt)
        return
    if args.profile_uniform:
        profile_uniform_main(args, torch, ipt)
        return
    if args.batch_routes:
        batch_routes_main(args, torch, ipt)
        return
    if args.move_lists:
        move_lists_main(args, torch, ipt)
        return
    if args.configs:
        configs_main(args, torch, ipt)
        return

    bl = batch_leg(torch, ipt, args)
    payload = bl["payload"]
    value = 2 * payload / (bl["ms_step"] * 1e-3) / 1e9
    k_gbps = 2 * bl["kernel_payload"] / (bl["k_batch_ms"] * 1e-3) / 1e9
    print(f"[bench] batch {bl['ms_step']:.3f} ms/step {value:.0f} GB/s, k_batch "
          f"{bl['k_batch_ms']:.3f} ms", file=sys.stderr, flush=True)

    traffic, traffic_note = None, "not measured (--no-traffic)"
    if not args.no_traffic:
        totals, err = measure_traffic(["k_batch"], payload // 16 * 16,
                                      ["--profile-batch", "--steps", "2", "--warmup", "1"])
        if totals:
            traffic = int(totals["FETCH_SIZE"] + totals["WRITE_SIZE"])
            traffic_note = (f"FETCH {totals['FETCH_SIZE'] / 1e9:.3f} GB + WRITE "
                            f"{totals['WRITE_SIZE'] / 1e9:.3f} GB per launch, copy-calibrated "
                            f"({traffic / (2 * bl['kernel_payload']):.3f}x the algorithmic bytes)")
        else:
            traffic_note = err

    rows = []
    if not args.no_rows:
        cfg = dict(CONFIGS)
        for name in ROWS:
            if name == "2a+2b":
                a, b = row_leg(torch, ipt, "2a", cfg["2a"]), row_leg(torch, ipt, "2b", cfg["2b"])
                nb = sum(2 * r["shape"][0] * r["shape"][1] * r["shape"][2] * r["shape"][3]
                         for r in (a, b))
                ms = a["ms"] + b["ms"]
                rows.append({"config": "2a+2b", "GBps": round(nb / (ms * 1e-3) / 1e9, 1),
                             "ms": round(ms, 5), "residency": a["residency"],
                             "verified": a["verified"] and b["verified"], "parts": [a, b]})
            else:
                rows.append(row_leg(torch, ipt, name, cfg[name]))
            print(f"[bench] row {rows[-1]['config']}: {rows[-1]['GBps']} GB/s", file=sys.stderr,
                  flush=True)

    cpu = None if args.no_cpu else cpu_legs(ipt)
    speedup = bl["ms_per_call_heuristic"] / bl["ms_step"]
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bl["ms_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32+u64 (4- and 8-byte payloads moved as raw bits)",
        "data": "synthetic (iota bit patterns per matrix), device-resident",
        "config": {"workload": "uneven batch: 4096 row-major matrices, m and n log-uniform in "
                               "[2, 4096] (seed 4), 4- or 8-byte elements, one captured hipGraph "
                               "launch per step (learned selector + batched persistent kernel)",
                   "matrices": BATCH_COUNT, "payload_bytes": payload,
                   "residency": f"HBM ({payload / 1e9:.2f} GB > 256 MB Infinity Cache)",
                   "parallelism": "single GPU"},
        "frac_of_hbm_peak": round(value / HBM_PEAK_GBS, 4),
        "speedup": round(speedup, 2),
        "speedup_note": "t(per-call heuristic loop) / t(learned + batched graph)",
        "batching_gain": round(bl["ms_per_call_heuristic"] / bl["ms_batch_heuristic"], 3),
        "learned_gain_in_batch": round(bl["ms_batch_heuristic"] / bl["ms_step"], 3),
        "ms_per_call_heuristic": round(bl["ms_per_call_heuristic"], 3),
        "ms_batch_heuristic": round(bl["ms_batch_heuristic"], 4),
        "selector_loaded": bl["learned"],
        "verified": bl["verified_bad"] == 0,
        "verified_note": f"all {BATCH_COUNT} matrices vs the closed form after one untimed run "
                         f"({bl['verified_bad']} mismatching)",
        "copy_GBps": round(bl["copy_GBps"], 1),
        "wall_s": round(bl["wall"], 3),
        "plan_ms": round(bl["plan_ms"], 1),
        "items": {k: bl["info"][k] for k in ("n_slice_items", "n_gcs_items", "n_swap_items",
                                             "n_block_cycle", "n_big", "grid")},
        "routes": bl["routes"],
        "roofline": {"bound": "hbm", "achieved": round(k_gbps, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(k_gbps / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_note": traffic_note, "alg_bytes": 2 * bl["kernel_payload"],
                     "alg_bytes_note": "2 x the bytes of the instances whose work k_batch does "
                                       "(persistent-kernel routes and block-cycle block steps; "
                                       "separate launches and chunk steps are other kernels)",
                     "kernel": "k_batch", "kernel_ms": round(bl["k_batch_ms"], 4),
                     "timing": "HIP events recorded by the kernel's own launch on the run's "
                               "stream, eager runs, median of 5"},
        "rows": rows,
        "selector_heldout": selector_eval(ipt),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
