"""The reference's PortableRuntime over the MI355X kernel.

tests/runtime/runtime_driver.cpp is built twice from the same source: over
the reference kernel (oracle/_ref/runtime_driver_ref, CPU; its output is
tests/golden/runtime_driver.npz) and over the drop-in adapter
(oracle/_ref/hip_runtime/runtime_driver_hip: the contract's
class name and header over libwvamd).  The runtime sources -- the forcing
engine, the constant-stratification integration system, WVFixedStepRK4
(advanceToTime with a clipped last step), WVAdaptiveRK23, the built-in
forcings of the frozen schedule besides nonlinear advection (adaptive
damping, beta-plane PV advection, linear and quadratic bottom friction,
pseudo-topographic wave generation, fixed amplitudes) and the observers
coupled to the right-hand side (Lagrangian floats with spline and drifters
with linear velocity sampling, an antialiased F-grid dye, all fed by the
kernel's advection fields) -- are compiled unchanged in both.  Results must
agree: states and right-hand sides at 1e-12 relative (max|diff| /
max|ref|), times and step counts exactly.
"""
import os
import subprocess
import sys

import numpy as np
import pytest
from conftest import GOLDEN, ROOT

sys.path.insert(0, os.path.join(ROOT, "tests", "runtime"))
from driver_io import read_records  # noqa: E402

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "oracle", "_ref", "hip_runtime", "runtime_driver_hip")
# scalars compared exactly: end times, step counts and sizes, RHS evaluations
EXACT = {"rk4_nh_t", "rk4_nh_steps", "rk4_nh_last_step", "rk4_h_t", "rk4_h_steps", "rk4_h_last_step", "rk23_t",
         "rk23_accepted", "rk23_rejected", "rk23_rhs", "obs_rhs_count"}


def test_runtime_over_hip_kernel_matches_reference_kernel(torch_gpu, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("runtime driver over the adapter was not built (needs the contract headers at build time)")
    out = tmp_path / "hip.bin"
    print("[gpu] runtime driver over the HIP kernel ...", flush=True)
    r = subprocess.run([DRIVER, str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    mine = read_records(str(out))
    ref = dict(np.load(os.path.join(GOLDEN, "runtime_driver.npz"), allow_pickle=False))
    assert sorted(mine) == sorted(ref)
    for name in sorted(ref):
        if name in EXACT:
            assert np.array_equal(mine[name], ref[name]), (name, mine[name], ref[name])
        else:
            err = np.max(np.abs(mine[name] - ref[name])) / np.max(np.abs(ref[name]))
            print(f"[gpu] {name}: rel err {err:.2e}", flush=True)
            assert err <= 1e-12, (name, err)
