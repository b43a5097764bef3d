"""The row pass finds each pupil's columns of a grid row as one interval
(pa::circle_cols) instead of testing every element.  Built for the host from
the same header the kernels use, it is checked here against the per-element
predicate pa::in_circle (float64, no fused multiply-add, as the numpy mask of
pupils.py:8-10) on 288,000 rows: random pupils, radii that are whole numbers
of cells (circle edge through grid points), centres half-way between grid
points, sub-cell pupils and pupils partly off the grid."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_circle_cols_equals_per_element_predicate(tmp_path):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "pupil_interval_check")
    # -mfma -ffp-contract=fast: contract wherever the helpers allow it, as the device build does
    subprocess.run([hipcc, "--offload-host-only", "-O2", "-mfma", "-ffp-contract=fast", "-std=c++17", "-x", "hip",
                    os.path.join(HERE, "pupil_interval_check.cpp"), "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], check=False, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
