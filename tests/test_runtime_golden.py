"""The runtime golden vectors (tests/golden/runtime_driver.npz, written by the
reference PortableRuntime over the reference kernel) exercise what they are
meant to: the scheduled forcings change the right-hand side and the fixed
amplitudes hold, so the GPU comparison of the same driver
(tests/test_gpu_runtime.py) covers them."""
import os

import numpy as np
from conftest import GOLDEN


def _golden():
    return dict(np.load(os.path.join(GOLDEN, "runtime_driver.npz"), allow_pickle=False))


def test_topographic_forcing_contributes_to_the_right_hand_side():
    g = _golden()
    for p in ("topo_nh_", "topo_h_"):
        rhs, plain = g[p + "rhs"], g[p + "rhs_nonlinear"]
        assert np.all(np.isfinite(rhs)) and np.abs(rhs).max() > 0
        # pseudo-topographic generation is most of the tendency of this flow
        assert np.abs(rhs - plain).max() > 0.5 * np.abs(rhs).max()


def test_fixed_amplitudes_hold_through_the_rk4_steps():
    # WVFixedAmplitudeForcing restores Ap[6], Ap[13] and A0[11] after every step
    g = _golden()
    for p in ("topo_nh_", "topo_h_"):
        s = g[p + "state"].view(np.complex128)
        n = s.size // 3
        assert s[6] == 2e-3 + 1e-3j and s[13] == -1e-3 + 5e-4j
        assert s[2 * n + 11] == 4e-3 - 2e-3j
        assert np.abs(s).max() > 1e-2  # the rest of the state evolved


def test_forced_schedule_differs_from_plain_advection():
    g = _golden()
    for p in ("forced_nh_", "forced_h_"):
        assert np.all(np.isfinite(g[p + "rhs"])) and np.abs(g[p + "rhs"]).max() > 0
