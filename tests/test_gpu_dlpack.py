"""Zero-copy torch observations (the analogue of the reference's
gpu2gpu_transfer, sensor_wrapper.py:97-124) in a fresh process: torch's GPU
must be initialised before libgbr loads, which this test process (where
libgbr is already loaded) cannot do."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_torch_observations_alias_the_planes():
    script = os.path.join(os.path.dirname(__file__), "torch_dlpack_check.py")
    r = subprocess.run([sys.executable, "-u", script], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "DLPACK OK" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
