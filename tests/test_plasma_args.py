"""aph_plasma's option checks that run before any device work: image sizes must be whole
numbers in [1, 16384] (a negative or huge value was cast to uint32_t unchecked)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "aurora_amd", "apps", "_build", "aph_plasma")


@pytest.mark.parametrize("dims", ["0,5", "-3,5", "1e12,5", "2.5,4", "8,16385"])
def test_bad_dims_are_reported(tmp_path, dims):
    assert os.path.exists(BIN), "aurora_amd/apps/_build/aph_plasma is not built (run the build)"
    obj = tmp_path / "t.obj"
    obj.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    p = subprocess.run([BIN, str(obj), "--dims", dims], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    assert "--dims takes a whole number in [1, 16384]" in p.stderr
