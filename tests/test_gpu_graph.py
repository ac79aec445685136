"""Per-call transposes captured into a caller's HIP graph (torch.cuda.graph):
after one warm-up call outside capture (the block-cycle path builds its move
lists on a shape's first use), the cooperative kernel, the whole-slice
kernel, the square swap, the block-cycle path and the decomposition are
captured, and every replay, with eager work and synchronisations between
replays, matches the oracle applied as many times."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

CASES = [  # (d1, d2, d3, esz, variant name)
    (512, 384, 64, 4, "gcs"), (256, 1000, 17, 8, "gcs"), (100, 60, 3, 4, "slice_lds"),
    (1024, 1024, 1, 4, "square_swap"), (7, 1 << 18, 1, 4, "block_cycle"), (1000, 3000, 1, 8, "c2r"),
    (777, 333, 2, 4, "r2c"),
]


def test_captured_transposes_replay_exactly(ipt, oracle):
    variants = {"gcs": ipt.VARIANT_GCS, "slice_lds": ipt.VARIANT_SLICE_LDS,
                "square_swap": ipt.VARIANT_SQUARE_SWAP, "block_cycle": ipt.VARIANT_BLOCK_CYCLE,
                "c2r": ipt.VARIANT_C2R, "r2c": ipt.VARIANT_R2C}
    side = torch.zeros(1 << 20, device="cuda")
    for k, (d1, d2, d3, esz, name) in enumerate(CASES):
        v = variants[name]
        if not ipt.feasible(v, esz, d1, d2, d3):
            continue
        print(f"[case] graph {name} d1={d1} d2={d2} d3={d3} esz={esz}", flush=True)
        n = d1 * d2 * d3
        a = oracle.random_bits(n, esz, 6100 + k)
        t = torch.from_numpy(a.view(np.int32 if esz == 4 else np.int64)).cuda()
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            ipt.transpose_variant(t, d1, d2, d3, v)  # warm-up outside capture
        s.synchronize()
        exp = oracle.seq_tt_213(a, d1, d2, d3)
        assert np.array_equal(t.cpu().numpy().view(a.dtype), exp)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            ipt.transpose_variant(t, d1, d2, d3, v)
        for r in range(3):
            g.replay()
            torch.cuda.synchronize()
            exp = oracle.seq_tt_213(exp, d1, d2, d3)
            assert np.array_equal(t.cpu().numpy().view(a.dtype), exp), (name, r)
            side.mul_(1.5).add_(1.0)  # eager work between replays
            torch.cuda.synchronize()
        del g
    ipt.device_check()
