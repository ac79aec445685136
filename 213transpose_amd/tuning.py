"""Autotuner sweeps, selector training and evaluation.

    python 213transpose_amd/tuning.py sweep --count 22000 --seed 5 \
        --start 0 --stop 5000 --out tune_part0.npz
    python 213transpose_amd/tuning.py merge --out data.npz part0.npz part1.npz ...
    python 213transpose_amd/tuning.py train --data data.npz --heldout 2000 \
        --out 213transpose_amd/selector/weights.f32
    python 213transpose_amd/tuning.py evaluate --data data.npz --heldout 2000 \
        --weights 213transpose_amd/selector/weights.f32

sweep: seeded shapes with m, n log-uniform in [2, 65536] and 4- or 8-byte
elements (a quarter of them with d3 > 1, log-uniform up to 512, total bytes
capped); every candidate configuration timed by ipt_autotune.  The dataset
holds the shapes, the features and the per-candidate times (negative =
infeasible).  train: the MLP trained by the library's HIP kernels on log2
times of the first count - heldout shapes.  evaluate: on the last `heldout`
shapes, regret of the selector's choice against exhaustive search and total
time against the plain heuristic's choice.
"""
import argparse
import importlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ipt = importlib.import_module("213transpose_amd")

# The sweep's domain is truncated at MAX_BYTES per shape: timing every
# candidate on the multi-GB tail of the log-uniform distribution would cost
# GPU-hours.  The selector decides only inside this domain (see
# kSelectorMaxBytes in the library); larger shapes, and shapes with a side
# beyond the sweep's 65536 (kSelectorMaxSide), keep the heuristic.
MAX_BYTES = 1 << 30


def sample_shapes(count, seed, lo=2, hi=65536, d3_frac=0.25, d3_max=512, max_bytes=MAX_BYTES):
    """[(d1, d2, d3, esz)] for the sweep (row-major m x n -> d1 = n, d2 = m).
    m, n log-uniform in [lo, hi]; draws whose m*n*esz exceeds max_bytes are
    drawn again; with probability d3_frac a slice count d3 log-uniform in
    [1, d3_max] is added, capped so the total stays within max_bytes."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        m, n = (int(x) for x in np.minimum(np.floor(np.exp(rng.uniform(
            np.log(lo), np.log(hi + 1), size=2))), hi))
        esz = 4 if rng.random() < 0.5 else 8
        use_d3 = rng.random() < d3_frac
        d3r = int(np.floor(np.exp(rng.uniform(0, np.log(d3_max + 1)))))
        if m * n * esz > max_bytes:
            continue
        d3 = max(1, min(d3r, max_bytes // (m * n * esz))) if use_d3 else 1
        out.append((n, m, d3, esz))
    return out


def reps_for(nbytes):
    """Timed runs per candidate (ipt_autotune adds one warm-up run)."""
    if nbytes < (1 << 20):
        return 10
    if nbytes < (16 << 20):
        return 4
    return 2


def sweep(args):
    import torch
    shapes = sample_shapes(args.count, args.seed)[args.start:args.stop]
    nbytes = max(d1 * d2 * d3 * esz for d1, d2, d3, esz in shapes)
    torch.cuda.set_device(0)
    buf = torch.empty(nbytes // 8 + 2, dtype=torch.int64, device="cuda")
    raw = buf.view(torch.uint8)
    feats = np.zeros((len(shapes), ipt.MLP_IN), np.float32)
    times = np.full((len(shapes), ipt.N_CANDIDATES), -1.0, np.float32)
    done = 0
    t0 = last = time.time()
    for k, (d1, d2, d3, esz) in enumerate(shapes):
        n = d1 * d2 * d3
        t = raw[:n * esz].view(torch.int32 if esz == 4 else torch.int64)
        feats[k] = ipt.shape_features(esz, d1, d2, d3)
        times[k], _ = ipt.autotune(t, d1, d2, d3, reps=reps_for(n * esz))
        done = k + 1
        if time.time() - last > 20 or done == len(shapes):
            last = time.time()
            np.savez(args.out, shapes=np.array(shapes[:done], np.int64), feats=feats[:done],
                     times=times[:done], first=args.start)
            print(f"[sweep] {args.start + done}/{args.stop} shapes, {last - t0:.0f} s", flush=True)
        if args.budget and time.time() - t0 > args.budget:
            print(f"[sweep] budget reached at {args.start + done}", flush=True)
            break
    np.savez(args.out, shapes=np.array(shapes[:done], np.int64), feats=feats[:done],
             times=times[:done], first=args.start)
    ipt.device_check()


def merge(args):
    parts = sorted((np.load(p) for p in args.parts), key=lambda d: int(d["first"]))
    np.savez_compressed(args.out, shapes=np.concatenate([d["shapes"] for d in parts]),
                        seed=args.seed,
                        feats=np.concatenate([d["feats"] for d in parts]),
                        times=np.concatenate([d["times"] for d in parts]))


def targets(times):
    mask = (times > 0).astype(np.float32)
    y = np.where(times > 0, np.log2(np.maximum(times, 1e-3)), 0.0).astype(np.float32)
    return y, mask


def init_params(seed):
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((ipt.MLP_IN, ipt.MLP_HIDDEN), (ipt.MLP_HIDDEN, ipt.MLP_HIDDEN),
                            (ipt.MLP_HIDDEN, ipt.N_CANDIDATES)):
        parts.append(rng.normal(0, np.sqrt(2.0 / fan_in), size=fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def train(args):
    d = np.load(args.data)
    n = len(d["feats"]) - args.heldout
    y, mask = targets(d["times"][:n])
    params = init_params(args.seed)
    t0 = time.time()
    params, loss = ipt.selector_train(d["feats"][:n], y, mask, params, epochs=args.epochs,
                                      lr=args.lr)
    print(f"[train] {n} shapes, {args.epochs} epochs, final masked MSE {loss:.4f} "
          f"(log2 us), {time.time() - t0:.1f} s", flush=True)
    params.astype(np.float32).tofile(args.out)


def choose(pred, times):
    """argmin of predictions over the feasible candidates."""
    p = np.where(times > 0, pred, np.inf)
    return np.argmin(p, axis=1)


def evaluate_arrays(shapes, feats, times, params):
    pred = ipt.selector_infer_host(params, feats)
    sel = choose(pred, times)
    heur = np.array([ipt.heuristic_candidate(int(e), int(a), int(b), int(c))
                     for a, b, c, e in shapes])
    rows = np.arange(len(times))
    t_sel = times[rows, sel]
    t_heur = times[rows, heur]
    t_best = np.where(times > 0, times, np.inf).min(axis=1)
    ok = (t_heur > 0) & (t_sel > 0)
    regret = t_sel[ok] / t_best[ok] - 1.0
    return {
        "shapes": int(ok.sum()),
        "regret_mean": float(regret.mean()), "regret_median": float(np.median(regret)),
        "regret_p90": float(np.quantile(regret, 0.9)),
        "heuristic_regret_mean": float((t_heur[ok] / t_best[ok] - 1.0).mean()),
        "picked_best_frac": float((t_sel[ok] == t_best[ok]).mean()),
        "total_time_speedup_vs_heuristic": float(t_heur[ok].sum() / t_sel[ok].sum()),
        "geomean_speedup_vs_heuristic": float(np.exp(np.log(t_heur[ok] / t_sel[ok]).mean())),
        "total_time_vs_oracle_best": float(t_sel[ok].sum() / t_best[ok].sum()),
    }


def evaluate(args):
    d = np.load(args.data)
    params = np.fromfile(args.weights, dtype=np.float32)
    h = args.heldout
    res = evaluate_arrays(d["shapes"][-h:], d["feats"][-h:], d["times"][-h:], params)
    print(res)
    return res


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("sweep")
    s.add_argument("--count", type=int, default=22000)
    s.add_argument("--seed", type=int, default=5)
    s.add_argument("--start", type=int, default=0)
    s.add_argument("--stop", type=int, default=22000)
    s.add_argument("--budget", type=float, default=0, help="seconds; 0 = no limit")
    s.add_argument("--out", required=True)
    m = sub.add_parser("merge")
    m.add_argument("--out", required=True)
    m.add_argument("--seed", type=int, default=5)
    m.add_argument("parts", nargs="+")
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--heldout", type=int, default=2000)
    t.add_argument("--epochs", type=int, default=3000)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("--data", required=True)
    e.add_argument("--heldout", type=int, default=2000)
    e.add_argument("--weights", required=True)
    a = ap.parse_args()
    {"sweep": sweep, "merge": merge, "train": train, "evaluate": evaluate}[a.cmd](a)


if __name__ == "__main__":
    main()
