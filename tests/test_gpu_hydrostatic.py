"""GPU parity of the arbitrary-stratification hydrostatic path
(wv_create_hydrostatic) against the numpy oracle of the MATLAB algorithm
(oracle/hydrostatic.py), through the C-ABI.

Tolerances: FP64 flux, transforms and RK4 steps 1e-12 relative (max|diff| /
max|ref| jointly over the outputs); mixed flux 1e-4 on white-noise
coefficients; with constant-N operators the stratified path reproduces the
constant-stratification kernel to 1e-9 (the accuracy of the host eigen-solve).
"""
import numpy as np
import pytest
from conftest import rel_err, rel_err_joint

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-4


def _setup(Nx, Ny, Nz, Nj, kind="exponential", Lx=100e3, Lz=4000.0, seed=1, amp=1e-2):
    import wvamd
    from hydrostatic import HydrostaticOracle
    from oracle import Oracle
    from wvamd import stratification as S
    cfg = wvamd.make_config(Nx, Ny, Nz, Nj, Lx=Lx, Lz=Lz, isHydrostatic=True)
    N2, dN2 = S.exponential() if kind == "exponential" else S.constant(cfg["N0"])
    st = S.Stratification(N2, Lz=Lz, g=cfg["g"], dN2=dN2)
    ops = S.vertical_operators(st, Nz, Nj)
    ho = HydrostaticOracle(cfg, ops, Oracle(cfg).modes())
    rng = np.random.default_rng(seed)
    sh = (ho.Nkl, Nj)
    A = [amp * (rng.standard_normal(sh) + 1j * rng.standard_normal(sh)) for _ in range(3)]
    print(f"[gpu] hydrostatic {kind} {Nx}x{Ny}x{Nz} Nj={Nj}", flush=True)
    return cfg, ops, ho, A


def _kernel(cfg, ops, precision="fp64"):
    import wvamd
    return wvamd.WVTransformHydrostaticKernel(cfg, ops, precision)


def _host(ts):
    return [t.cpu().numpy() for t in ts]


@pytest.mark.parametrize("shape", [(16, 16, 9, 8), (32, 32, 17, 16), (24, 20, 33, 21)])
def test_stratified_flux_fp64_matches_oracle(torch_gpu, shape):
    cfg, ops, ho, A = _setup(*shape)
    k = _kernel(cfg, ops)
    assert k.Nkl == ho.Nkl
    F = _host(k.nonlinear_flux(*[k.to_device(a) for a in A], t=1800.0, t0=600.0))
    R = ho.nonlinear_flux(*A, t=1800.0, t0=600.0)
    err = rel_err_joint(F, R)
    print(f"[gpu] stratified fp64 flux rel err {err:.2e}", flush=True)
    assert err <= FP64_TOL


def test_stratified_flux_mixed_matches_oracle(torch_gpu):
    cfg, ops, ho, A = _setup(32, 32, 17, 16)
    k = _kernel(cfg, ops, "mixed")
    F = _host(k.nonlinear_flux(*[k.to_device(a) for a in A], t=1800.0))
    err = rel_err_joint(F, ho.nonlinear_flux(*A, t=1800.0))
    print(f"[gpu] stratified mixed flux rel err {err:.2e}", flush=True)
    assert err <= MIXED_TOL


def test_stratified_constant_n_reproduces_constant_kernel(torch_gpu):
    import wvamd
    cfg, ops, ho, A = _setup(32, 32, 17, 11, kind="constant", Lz=1300.0)
    ks = _kernel(cfg, ops)
    kc = wvamd.WVTransformConstantStratificationKernel(cfg, "fp64")
    Fs = _host(ks.nonlinear_flux(*[ks.to_device(a) for a in A], t=900.0))
    Fc = _host(kc.nonlinear_flux(*[kc.to_device(a) for a in A], t=900.0))
    assert rel_err_joint(Fs, Fc) <= 1e-9


def test_stratified_transforms_match_oracle(torch_gpu):
    from wvamd import flows
    cfg, ops, ho, _ = _setup(32, 32, 17, 16)
    k = _kernel(cfg, ops)
    A = flows.random_flow_shape(ho.md["Kh"], cfg["Nj"], ho.ApmTE.ravel(), ho.A0TE.ravel(), seed=3, jmax=16)
    dev = [k.to_device(a) for a in A]
    f = k.transform_wave_vortex_to_uvweta(*dev, t=400.0).cpu().numpy()
    assert rel_err(f, ho.fields(*A, t=400.0)) <= FP64_TOL
    back = _host(k.transform_to_wave_vortex(torch_gpu.as_tensor(f[[0, 1, 3]].copy()).cuda(), t=400.0))
    assert rel_err_joint(back, A) <= FP64_TOL
    E = k.total_energy(*dev)
    assert abs(E - ho.energy(*A)) <= 1e-13 * abs(E)


def test_stratified_rk4_steps_match_oracle(torch_gpu):
    cfg, ops, ho, A = _setup(16, 16, 17, 12, seed=4, amp=1e-4)
    k = _kernel(cfg, ops)
    dev = [k.to_device(a) for a in A]
    y = [a.copy() for a in A]
    t = 0.0
    for step in range(3):
        t_gpu = k.rk4_advance(*dev, t=t, h=100.0, nsteps=1)
        y0_, y1_, y2_, t = ho.rk4(*y, t=t, dt=100.0, nsteps=1)
        y = [y0_, y1_, y2_]
        err = rel_err_joint(_host(dev), y)
        print(f"[gpu] stratified rk4 step {step} rel err {err:.2e}", flush=True)
        assert err <= FP64_TOL
        assert t_gpu == t


def test_c3_512x512x65_flux_and_100_steps(torch_gpu):
    """The arbitrary-stratification workload (512 x 512, 65 quadrature points,
    64 modes, N^2 = N0^2 exp(2z/b), |k|^-2 waves + |k|^-3 geostrophic flow at
    rms 0.1 m/s, dt = 0.5 CFL): one FP64 flux against the numpy oracle at
    1e-12, then 100 FP64 and 100 mixed RK4 steps with the mixed total energy
    within 1e-6 of FP64 at every 10th step."""
    import time
    import wvamd
    from hydrostatic import HydrostaticOracle
    from oracle import Oracle
    from wvamd.workloads import arbitrary_stratification_512
    cfg, ops, k64, A, dt = arbitrary_stratification_512(wvamd)
    k32 = _kernel(cfg, ops, "mixed")
    print(f"[gpu] C3 512x512x65 Nkl={k64.Nkl} dt={dt:.1f}s", flush=True)
    dev = [k64.to_device(a) for a in A]
    F = _host(k64.nonlinear_flux(*dev, t=0.0))
    t0 = time.perf_counter()
    ho = HydrostaticOracle(cfg, ops, Oracle(cfg).modes())
    R = ho.nonlinear_flux(*A, t=0.0)
    err = rel_err_joint(F, R)
    print(f"[gpu] C3 fp64 flux vs oracle rel err {err:.2e} (oracle {time.perf_counter() - t0:.1f}s)", flush=True)
    assert err <= FP64_TOL
    del ho, R
    d64 = [k64.to_device(a) for a in A]
    d32 = [k32.to_device(a) for a in A]
    e0 = k64.total_energy(*d64)
    t = 0.0
    worst = 0.0
    for block in range(10):
        t64 = k64.rk4_advance(*d64, t=t, h=dt, nsteps=10)
        t32 = k32.rk4_advance(*d32, t=t, h=dt, nsteps=10)
        assert t64 == t32
        t = t64
        e64, e32 = k64.total_energy(*d64), k64.total_energy(*d32)
        rel = abs(e32 - e64) / abs(e64)
        worst = max(worst, rel)
        assert np.isfinite(e64) and rel <= 1e-6, (block, rel)
    print(f"[gpu] C3 100 steps: E0={e0:.6e} E={e64:.6e} mixed-vs-fp64 worst energy rel err {worst:.2e}", flush=True)
    print(f"[gpu] C3 mixed-vs-fp64 coefficient rel err "
          f"{rel_err_joint([d.cpu().numpy() for d in d32], [d.cpu().numpy() for d in d64]):.2e}", flush=True)
