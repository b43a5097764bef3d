"""The polynomial block's Bessel values (pa::bessel_j<32>, Miller's backward
recurrence; host build of the header the kernels use) against scipy.special.jv
over the z range the block uses, |z| <= 14, including tiny and negative z."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_bessel_j_matches_scipy(tmp_path):
    jv = pytest.importorskip("scipy.special").jv
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "bessel_check")
    subprocess.run([hipcc, "--offload-host-only", "-O2", "-mfma", "-ffp-contract=fast", "-std=c++17", "-x", "hip",
                    os.path.join(HERE, "bessel_check.cpp"), "-o", exe], check=True, capture_output=True)
    z = np.concatenate([np.linspace(-14, 14, 2801), [0.0, 1e-300, 1e-12, -1e-7, 3e-3, 13.999999, -14.0]])
    out = subprocess.run([exe], input="\n".join(f"{v:.17g}" for v in z), capture_output=True, text=True,
                         timeout=60, check=True).stdout
    got = np.array([[float(t) for t in line.split()] for line in out.strip().splitlines()])
    assert got.shape == (len(z), 33) and np.array_equal(got[:, 0], z)
    ref = jv(np.arange(32)[None, :], z[:, None])
    # absolute error relative to the largest |J_m(z)| (<= 1): the coefficients
    # enter the block as 2 J_m(z) next to terms of size ~1
    assert np.abs(got[:, 1:] - ref).max() < 1e-14
