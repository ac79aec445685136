"""The product's host CPU path (ipt_cpu_transpose / ipt_cpu_transpose_213,
the replacement of the reference's declared inplace::openmp::transpose,
include/openmp.h:7-8), checked bit-exact against the oracle's seq_tt on
seeded random bits.  CPU only: these run without a GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpu_213_random_shapes(ipt, oracle):
    rng = np.random.default_rng(31)
    for k in range(250):
        d1, d2 = int(rng.integers(1, 260)), int(rng.integers(1, 260))
        if k % 6 == 0:
            d2 = d1  # square slices take the tile-pair swap
        d3 = int(rng.integers(1, 5))
        esz = int(rng.choice([4, 8]))
        a = oracle.random_bits(d1 * d2 * d3, esz, 4000 + k)
        exp = oracle.seq_tt_213(a, d1, d2, d3)
        ipt.cpu_transpose_213(a, d1, d2, d3, threads=int(rng.integers(1, 9)))
        assert np.array_equal(a, exp), (d1, d2, d3, esz)


@pytest.mark.parametrize("shape", [
    (3000, 1000, 1, 4),     # config 1 (gcd 1000: C2R with a pre-rotation)
    (512, 384, 16, 4),      # a quarter of config 2a (workers take whole slices)
    (256, 1000, 3, 8),      # config 2b's slices (R2C view)
    (4099, 4093, 1, 4),     # config 3c (coprime)
    (7, 1 << 20, 1, 4),     # skinny, fewer rows than workers (shared rows)
    (1 << 21, 3, 1, 8),     # skinny with rows above the shared-row threshold
    (777, 777, 2, 8),       # squares, ragged last tile
])
def test_cpu_config_shapes(ipt, oracle, shape):
    d1, d2, d3, esz = shape
    a = oracle.random_bits(d1 * d2 * d3, esz, 77)
    exp = oracle.seq_tt_213(a, d1, d2, d3)
    ipt.cpu_transpose_213(a, d1, d2, d3, threads=8)
    assert np.array_equal(a, exp)


def test_cpu_row_and_column_major(ipt, oracle):
    m, n = 123, 457
    for esz in (4, 8):
        a = oracle.random_bits(m * n, esz, 5)
        b = a.copy()
        ipt.cpu_transpose(True, b, m, n)  # row-major m x n -> (d1 = n, d2 = m)
        assert np.array_equal(b, oracle.seq_tt_213(a, n, m, 1))
        b = a.copy()
        ipt.cpu_transpose(False, b, m, n)  # column-major m x n -> (d1 = m, d2 = n)
        assert np.array_equal(b, oracle.seq_tt_213(a, m, n, 1))


def test_cpu_caller_scratch(ipt, oracle):
    d1, d2, d3, esz, threads = 1000, 300, 2, 8, 4
    need = ipt.cpu_scratch_bytes(esz, d1, d2, d3, threads)
    assert need > 0
    tmp = np.zeros(need, dtype=np.uint8)
    a = oracle.random_bits(d1 * d2 * d3, esz, 9)
    exp = oracle.seq_tt_213(a, d1, d2, d3)
    ipt.cpu_transpose_213(a, d1, d2, d3, threads=threads, tmp=tmp)
    assert np.array_equal(a, exp)
    assert ipt.cpu_scratch_bytes(4, 1, 500, 1, 4) == 0  # one column: nothing moves
    assert ipt.cpu_scratch_bytes(3, 10, 10, 1, 4) == -1


def test_cpu_transpose_twice_is_identity(ipt, oracle):
    d1, d2, esz = 641, 1279, 4
    a = oracle.random_bits(d1 * d2, esz, 10)
    b = a.copy()
    ipt.cpu_transpose_213(b, d1, d2, 1)
    ipt.cpu_transpose_213(b, d2, d1, 1)
    assert np.array_equal(a, b)


def test_cpu_rejects_bad_arguments(ipt):
    import ctypes
    L = ipt.lib()
    buf = np.zeros(64, dtype=np.uint32)
    p = ctypes.c_void_p(buf.ctypes.data)
    assert L.ipt_cpu_transpose_213(None, 4, 4, 4, 1, None, 1) == 1
    assert L.ipt_cpu_transpose_213(p, 2, 4, 4, 1, None, 1) == 1
    assert L.ipt_cpu_transpose_213(p, 4, 0, 4, 1, None, 1) == 1
    assert L.ipt_cpu_transpose_213(ctypes.c_void_p(buf.ctypes.data + 2), 4, 4, 4, 1, None, 1) == 1
    assert L.ipt_cpu_transpose(1, p, 4, -3, 4, None, 1) == 1
    # a view whose C2R shuffle would need k * (x mod b) >= 2^63 has no exact
    # plan (b = C / gcd above ~3e9; 2^33 + 3 is coprime to 3, so b = C):
    # refused before touching the data
    big = (1 << 33) + 3
    assert ipt.plan_query(4, big, 3)["b"] == big
    assert L.ipt_cpu_transpose_213(p, 4, big, 3, 1, None, 1) == 3


def test_cpp_mirror_openmp_form_runs_on_host(tmp_path):
    """inplace::openmp::transpose compiled against include/inplace_transpose.hpp
    and linked with the library: a host malloc'd matrix transposed in place."""
    src = tmp_path / "mirror.cpp"
    src.write_text(r'''
#include "inplace_transpose.hpp"
#include <cstdio>
#include <cstdlib>
#include <vector>
int main() {
    const int m = 37, n = 91;
    std::vector<float> a(m * n), orig(m * n);
    for (int i = 0; i < m * n; i++) { unsigned u = 2654435761u * (unsigned)i; float f; __builtin_memcpy(&f, &u, 4); a[i] = orig[i] = f; }
    inplace::openmp::transpose(true, a.data(), m, n, (float*)nullptr, 3);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++)
            if (__builtin_memcmp(&a[j * m + i], &orig[i * n + j], 4)) { std::printf("mismatch\n"); return 1; }
    try {
        inplace::openmp::transpose(true, (float*)nullptr, m, n);
        return 2;
    } catch (const inplace::error& e) {
        if (e.status != IPT_ERR_INVALID_ARG) return 3;
    }
    std::printf("ok\n");
    return 0;
}
''')
    exe = tmp_path / "mirror"
    libdir = os.path.join(ROOT, "213transpose_amd")
    r = subprocess.run(["g++", "-std=c++14", "-O1", "-I", os.path.join(ROOT, "include"),
                        "-I", "/opt/rocm/include", str(src), "-o", str(exe), "-L", libdir,
                        "-linplace_hip", f"-Wl,-rpath,{libdir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, (r.returncode, r.stdout, r.stderr)
