"""Several host threads, each on its own stream, running the path's variants
at once (cooperative kernel, block-cycle, decomposition with its scratch
tier): the per-device arenas and their cross-stream ordering keep every
result exact."""
import threading

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def to_dev(a):
    return torch.from_numpy(a.view(np.int32 if a.itemsize == 4 else np.int64)).cuda()


def to_host(t, esz):
    return t.cpu().numpy().view(np.uint32 if esz == 4 else np.uint64)


def test_threads_on_separate_streams(ipt, oracle):
    cases = [((512, 384, 8), 4, ipt.VARIANT_GCS), ((7, 65536, 2), 4, ipt.VARIANT_BLOCK_CYCLE),
             ((96, 4096, 2), 8, ipt.VARIANT_C2R), ((300, 200, 3), 8, ipt.VARIANT_AUTO),
             ((4096, 96, 2), 4, ipt.VARIANT_R2C)]
    nthreads, reps = 4, 3
    inputs = [[oracle.random_bits(d1 * d2 * d3, esz, 2000 + 10 * tid + k)
               for k, ((d1, d2, d3), esz, _) in enumerate(cases)] for tid in range(nthreads)]
    results, errors = {}, []

    def work(tid):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                ts = [to_dev(a) for a in inputs[tid]]
                for _ in range(reps):
                    for t, ((d1, d2, d3), esz, v) in zip(ts, cases):
                        ipt.transpose_variant(t, d1, d2, d3, v, stream=s)
                        ipt.transpose(t, d2, d1, d3, stream=s)  # back
                for t, ((d1, d2, d3), esz, v) in zip(ts, cases):
                    ipt.transpose_variant(t, d1, d2, d3, v, stream=s)
                s.synchronize()
                results[tid] = [to_host(t, esz) for t, (_, esz, _) in zip(ts, cases)]
        except Exception as e:  # reported below, on the main thread
            errors.append((tid, repr(e)))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(nthreads)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    ipt.device_check()
    for tid in range(nthreads):
        for k, ((d1, d2, d3), esz, v) in enumerate(cases):
            exp = oracle.seq_tt_213(inputs[tid][k], d1, d2, d3)
            assert np.array_equal(results[tid][k], exp), (tid, d1, d2, d3, esz, v)
