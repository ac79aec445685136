"""Writes the golden fixtures under tests/golden/ from the reference's own
answer generator, gen_ans (src/gen_ans.cpp + src/tensor_util.cpp), built from
source into oracle/_ref/gen_ans by oracle/Makefile.

gen_ans d1 d2 d3 esz [out] initialises data[i] = i (int for esz 4, long long
for esz 8; src/tensor_util.cpp:42-47) and writes seq_tt's permutation {1,0,2}
of it as raw little-endian bytes (src/tensor_util.cpp:69-92).

Outputs:
  small/<d1>_<d2>_<d3>_<esz>.bin   the raw answer bytes for small shapes
  golden.json                      shape -> sha256 + byte count, for the small
                                   shapes and for every full-size config shape

Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GEN_ANS = os.path.join(ROOT, "oracle", "_ref", "gen_ans")

# (d1, d2, d3, esz): small shapes stored byte for byte.
SMALL = [
    (1, 1, 1, 4), (1, 7, 2, 4), (7, 1, 2, 8), (2, 3, 1, 4), (3, 2, 1, 8),
    (5, 5, 3, 4), (30, 7, 3, 4), (7, 30, 3, 8), (33, 33, 2, 4), (32, 48, 2, 8),
    (40, 24, 1, 4), (24, 40, 1, 4), (64, 33, 3, 8), (97, 41, 1, 4), (41, 97, 2, 8),
    (128, 96, 2, 4), (96, 128, 1, 8), (300, 100, 1, 4), (100, 300, 1, 8),
    (512, 12, 2, 4), (12, 512, 2, 4), (1000, 3, 1, 8), (3, 1000, 1, 4),
]
# Full-size config shapes (hash only): row-major f32 1000x3000; 213
# [512][384][64] f32 and [256][1000][17] f64; int64 32 x 2^24; f32 2^22 x 7;
# f32 4093 x 4099; f32 16384^2.
LARGE = [
    (3000, 1000, 1, 4), (512, 384, 64, 4), (256, 1000, 17, 8),
    (1 << 24, 32, 1, 8), (7, 1 << 22, 1, 4), (4099, 4093, 1, 4), (16384, 16384, 1, 4),
]


def run(d1, d2, d3, esz):
    h = hashlib.sha256()
    n = 0
    with subprocess.Popen([GEN_ANS, str(d1), str(d2), str(d3), str(esz)],
                          stdout=subprocess.PIPE) as p:
        chunks = []
        while True:
            b = p.stdout.read(1 << 24)
            if not b:
                break
            h.update(b)
            n += len(b)
            if n <= (1 << 20):
                chunks.append(b)
        if p.wait() != 0:
            raise RuntimeError(f"gen_ans failed for {(d1, d2, d3, esz)}")
    assert n == d1 * d2 * d3 * esz, (n, d1, d2, d3, esz)
    return h.hexdigest(), n, b"".join(chunks)


def main():
    if not os.path.exists(GEN_ANS):
        sys.exit("oracle/_ref/gen_ans missing: run make -C oracle first")
    os.makedirs(os.path.join(HERE, "small"), exist_ok=True)
    table = []
    for shape in SMALL + LARGE:
        digest, n, data = run(*shape)
        entry = {"d1": shape[0], "d2": shape[1], "d3": shape[2], "esz": shape[3],
                 "bytes": n, "sha256": digest}
        if shape in SMALL:
            name = "%d_%d_%d_%d.bin" % shape
            with open(os.path.join(HERE, "small", name), "wb") as f:
                f.write(data)
            entry["file"] = "small/" + name
        table.append(entry)
        print(entry, flush=True)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump({"generator": "oracle/_ref/gen_ans (reference src/gen_ans.cpp, src/tensor_util.cpp)",
                   "input": "data[i] = i (int32 for esz 4, int64 for esz 8)",
                   "cases": table}, f, indent=1)


if __name__ == "__main__":
    main()
