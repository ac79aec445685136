"""Randomized parity: seeded random shapes, element sizes, base offsets and
variants (every forced variant where feasible, AUTO, and the decomposition
with small LDS budgets that force its scratch tiers), each compared bit for
bit with the oracle's seq_tt over the whole buffer (the element before the
view must stay untouched).  Deterministic: fixed seed and case count."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

CASES = 2500


def draw(rng):
    kind = rng.integers(0, 4)
    if kind == 0:    # small and mid-size slices
        d1, d2 = (int(x) for x in np.exp(rng.uniform(0, np.log(3000), 2)))
        d3 = int(rng.integers(1, 5))
    elif kind == 1:  # many slices: several rounds of cooperative groups
        d1, d2 = (int(x) for x in np.exp(rng.uniform(np.log(200), np.log(2500), 2)))
        d3 = int(rng.integers(2, 12))
    elif kind == 2:  # skinny slices: the block-cycle path
        a = int(np.exp(rng.uniform(np.log(2), np.log(64))))
        b = int(np.exp(rng.uniform(np.log(1000), np.log(400000))))
        d1, d2 = (a, b) if rng.random() < 0.5 else (b, a)
        d3 = int(rng.integers(1, 3))
    else:            # tall columns: the decomposition's scratch tiers
        d1, d2, d3 = int(rng.integers(40, 400)), int(rng.integers(3000, 12000)), 1
        if rng.random() < 0.5:
            d1, d2 = d2, d1
    return max(d1, 1), max(d2, 1), d3


def test_random_shapes_and_variants_match_oracle(ipt, oracle):
    L = ipt.lib()
    rng = np.random.default_rng(20261203)
    variants = {"auto": ipt.VARIANT_AUTO, "gcs": ipt.VARIANT_GCS, "block_cycle": ipt.VARIANT_BLOCK_CYCLE,
                "c2r": ipt.VARIANT_C2R, "r2c": ipt.VARIANT_R2C, "slice_lds": ipt.VARIANT_SLICE_LDS,
                "square_swap": ipt.VARIANT_SQUARE_SWAP}
    names = list(variants)
    ran = 0
    while ran < CASES:
        d1, d2, d3 = draw(rng)
        esz = int(rng.choice([4, 8]))
        off = int(rng.integers(0, 2))
        name = names[int(rng.integers(0, len(names)))]
        tune_bytes = int(rng.choice([256, 4096, 20000, 65536])) if rng.random() < 0.5 else 0
        seed = int(rng.integers(1 << 30))
        n = d1 * d2 * d3
        v = variants[name]
        if n > 6_000_000 or (v and not ipt.feasible(v, esz, d1, d2, d3)):
            continue
        tune = ipt.Tune(lds_bytes=tune_bytes) if tune_bytes and name in ("c2r", "r2c") else None
        print(f"[case] {name} d1={d1} d2={d2} d3={d3} esz={esz} off={off} "
              f"lds_bytes={tune.lds_bytes if tune else 0} seed={seed}", flush=True)
        a = oracle.random_bits(n + 1, esz, seed)
        t = torch.from_numpy(a.view(np.int32 if esz == 4 else np.int64)).cuda()
        view = t[off:off + n]
        st = L.ipt_transpose_variant(ctypes.c_void_p(view.data_ptr()), esz, d1, d2, d3, v,
                                     ctypes.byref(tune) if tune else None,
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        if st == 2:  # declined at run time (IPT_ERR_UNSUPPORTED), nothing written
            continue
        assert st == 0, (st, L.ipt_last_error())
        torch.cuda.synchronize()
        got = t.cpu().numpy().view(np.uint32 if esz == 4 else np.uint64)
        exp = a.copy()
        exp[off:off + n] = oracle.seq_tt_213(a[off:off + n], d1, d2, d3)
        assert np.array_equal(got, exp), (name, d1, d2, d3, esz, off, tune_bytes)
        ran += 1
    ipt.device_check()
