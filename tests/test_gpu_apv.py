"""GPU tests of wv_apv_device, the available potential vorticity of a state
(Operations/APVOperation.m:15-26 with eta_true = eta for the constant-N
background), through the C-ABI.

Parity bar: FP64 against the oracle's APV (composed from its state-field
derivatives, the composition pinned in tests/test_apv.py) at 1e-12 relative
(max|diff| / max|ref|); the mixed-precision kernel at 1e-4.
"""
import numpy as np
import pytest
from conftest import rel_err

from oracle import Oracle

pytestmark = pytest.mark.gpu

CASES = [
    dict(Nx=16, Ny=16, Nz=9, Nj=5),
    dict(Nx=32, Ny=24, Nz=17, Nj=11, Ly=40e3),
    dict(Nx=16, Ny=16, Nz=17, Nj=10, isHydrostatic=True),
    dict(Nx=64, Ny=64, Nz=33, Nj=21),
]


def _state(o, seed, scale=1e-3):
    rng = np.random.default_rng(seed)
    return [scale * (rng.standard_normal((o.Nkl, o.Nj)) + 1j * rng.standard_normal((o.Nkl, o.Nj)))
            for _ in range(3)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['Nx']}x{c['Ny']}x{c['Nz']}{'h' if c.get('isHydrostatic') else ''}")
def test_apv_matches_oracle(torch_gpu, case):
    import wvamd
    cfg = wvamd.make_config(**case)
    o = Oracle(cfg)
    for seed in (3, 4):
        A = _state(o, seed)
        ref = o.apv(*A, t=77.0, t0=5.0)
        for prec, tol in (("fp64", 1e-12), ("mixed", 1e-4), ("mixed-split16", 1e-4)):
            print(f"[gpu] apv {case} seed {seed} {prec}", flush=True)
            k = wvamd.WVTransformConstantStratificationKernel(cfg, prec)
            dev = [k.to_device(a) for a in A]
            mine = k.apv(*dev, t=77.0, t0=5.0).cpu().numpy()
            err = rel_err(mine, ref)
            print(f"[gpu]   rel err {err:.2e}", flush=True)
            assert err <= tol, (prec, err)


def test_apv_rejects_arbitrary_stratification_and_bad_arguments(torch_gpu):
    import wvamd
    from wvamd import stratification as S
    cfg = wvamd.make_config(16, 16, 9, 5, Lz=4000.0, isHydrostatic=True)
    N2, dN2 = S.exponential()
    ops = S.vertical_operators(S.Stratification(N2, Lz=4000.0, g=cfg["g"], dN2=dN2), 9, 5)
    k = wvamd.WVTransformHydrostaticKernel(cfg, ops, "fp64")
    A = [k.to_device(np.zeros(k.spectral_shape, complex)) for _ in range(3)]
    with pytest.raises(wvamd.WVKernelError) as e:
        k.apv(*A)
    assert e.value.status == "unsupportedOperation"
    kc = wvamd.WVTransformConstantStratificationKernel(wvamd.make_config(16, 16, 9, 5), "fp64")
    A = [kc.to_device(np.zeros(kc.spectral_shape, complex)) for _ in range(3)]
    with pytest.raises(wvamd.WVKernelError) as e:
        kc.apv(*A, t=float("nan"))
    assert e.value.status == "invalidConfiguration"
