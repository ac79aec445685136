"""The test_213inplace CLI (reference driver interface, src/test_213inplace.cu:66-90)
writes answer files byte-identical to the reference's gen_ans."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "213transpose_amd", "test_213inplace")


def test_cli_output_matches_gen_ans_files(tmp_path):
    assert os.path.exists(CLI), "test_213inplace not built"
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        cases = [c for c in json.load(f)["cases"] if "file" in c]
    for c in cases[::3]:
        out = tmp_path / "out.bin"
        args = [CLI, str(c["d1"]), str(c["d2"]), str(c["d3"]), str(c["esz"]), str(out)]
        print("[cli]", " ".join(args[1:5]), flush=True)
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert "Effective Bandwidth" in r.stdout
        with open(os.path.join(ROOT, "tests", "golden", c["file"]), "rb") as g:
            assert out.read_bytes() == g.read(), c


def test_cli_usage_without_arguments():
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage" in r.stdout
