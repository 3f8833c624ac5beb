"""The reference's own bootstrapping tests, written against the reference's
function names through the C++ drop-in header, run on the GPU (one program per
torus width, as the reference is compiled)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("width", [32, 56])
def test_dropin_program(width):
    exe = os.path.join(HERE, "yatfhe_amd", f"dropin_test_t{width}")
    assert os.path.exists(exe), "build the drop-in test programs with yatfhe_amd.build"
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "PASSED" in r.stdout
