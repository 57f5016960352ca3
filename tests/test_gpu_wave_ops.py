"""GPU: the wavefront scan, reduction, lane-shift and segmented-sum helpers
the kernels are built on (ugu_amd/csrc/device_common.h), lane by lane against
sequential results on the host (tests/wave_ops_check.hip, built with the
library)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_bin", "wave_ops_check")


def test_wave_ops(gpu_available):
    assert os.path.exists(BIN), "build() compiles tests/_bin/wave_ops_check"
    print("[wave-ops] running the check", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "2048 lanes checked, 0 mismatches" in r.stdout
