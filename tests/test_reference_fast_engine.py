"""The build-authored fast WVFFTEngine (oracle/fast_fft_engine.cpp) that times
the reference kernel on the host cores (the benchmark's CPU baseline),
checked against the reference's own direct-DFT engine
(CompiledKernel/adapters/reference/WVReferenceFFTEngine.cpp) plan by plan and
through the reference kernel's nonlinear flux.

The direct engine sums N (N^2 in 2-D) rounded terms, so the two differ by its
accumulated rounding: a few 1e-15 at 16-point lengths, about 1e-14 at 64 x 64
and at 129 points.  A wrong sign convention, a missing Hermitian half or a
shifted sine index gives O(1).
"""
import numpy as np
import pytest

from oracle import Reference, ReferenceFast, make_config

pytestmark = pytest.mark.skipif(not (ReferenceFast.available() and Reference.available()),
                                reason="reference kernel not built here")

PLANS = [("r2c", 16, 16), ("r2c", 12, 8), ("r2c", 16, 24), ("r2c", 64, 64),
         ("c2r", 16, 16), ("c2r", 12, 24), ("c2r", 64, 64),
         ("dct1", 2, 1), ("dct1", 9, 1), ("dct1", 10, 1), ("dct1", 33, 1), ("dct1", 129, 1),
         ("dst1", 1, 1), ("dst1", 7, 1), ("dst1", 8, 1), ("dst1", 31, 1), ("dst1", 127, 1)]


@pytest.mark.parametrize("kind,n0,n1", PLANS, ids=[f"{k}-{a}x{b}" for k, a, b in PLANS])
def test_fast_engine_matches_direct_engine(kind, n0, n1):
    tol = 3e-14 if max(n0, n1) >= 64 else 1e-14
    for seed in (1, 2):
        err = ReferenceFast.engine_compare(kind, n0, n1, batch=3, seed=seed)
        assert 0.0 <= err <= tol, (kind, n0, n1, err)


@pytest.mark.parametrize("shape", [(16, 16, 9, 5, False), (32, 32, 17, 11, False), (32, 24, 17, 10, True)])
def test_reference_kernel_flux_is_engine_independent(shape):
    Nx, Ny, Nz, Nj, hydro = shape
    cfg = make_config(Nx, Ny, Nz, Nj, Ly=40e3 if Ny != Nx else None, isHydrostatic=hydro)
    fast, direct = ReferenceFast(cfg), Reference(cfg)
    rng = np.random.default_rng(3)
    A = [1e-3 * (rng.standard_normal((fast.Nkl, Nj)) + 1j * rng.standard_normal((fast.Nkl, Nj))) for _ in range(3)]
    a = fast.nonlinear_flux(*A, t=10.0, t0=1.0)
    b = direct.nonlinear_flux(*A, t=10.0, t0=1.0)
    den = max(np.abs(x).max() for x in b)
    assert max(np.abs(x - y).max() for x, y in zip(a, b)) / den < 1e-12
