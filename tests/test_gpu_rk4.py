"""GPU fixed-step RK4 (WVFixedStepRK4::step, PortableRuntime/src/WVRungeKutta.cpp:335-462)
against the CPU oracle, step by step, plus size-independent properties at the
full 256x256x129 configuration.

Tolerances: FP64 state after every step within 1e-12 relative
(max|diff|/max|ref| over Ap, Am, A0 jointly); the flux conserves energy to
round-off; the mixed path's total energy stays within 1e-6 of the FP64 path.
"""
import numpy as np
import pytest
from conftest import rel_err_joint

from oracle import Oracle

pytestmark = pytest.mark.gpu


def _cfg(N, Nz, Lx, hydro=False, Nj=None):
    import wvamd
    return wvamd.make_config(N, N, Nz, Nj, Lx=Lx, isHydrostatic=hydro)


def _flow(k, seed, jmax, urms=0.05):
    """Seeded synthetic flow scaled to rms horizontal speed urms (GPU inverse transform)."""
    from wvamd import flows
    md = k.modes()
    Ap, Am, A0 = flows.random_flow_shape(md["Kh"], k.Nj, k.coefficient_table("ApmTE"),
                                         k.coefficient_table("A0TE"), seed=seed, jmax=jmax)
    f = k.transform_wave_vortex_to_uvw(*[k.to_device(a) for a in (Ap, Am, A0)]).cpu().numpy()
    s = flows.scale_factor_for_rms(f, k.config["Lz"], urms)
    return Ap * s, Am * s, A0 * s


def _dt(k, A):
    from wvamd import flows
    f = k.transform_wave_vortex_to_uvw(*[k.to_device(a) for a in A]).cpu().numpy()
    md = k.modes()
    return flows.cfl_time_step(f, md["k"], md["l"], k.coefficient_table("omega"), k.config["Lz"], 0.5)


def _kernel(cfg, precision="fp64"):
    import wvamd
    print(f"[gpu] rk4 kernel {cfg['Nx']}x{cfg['Ny']}x{cfg['Nz']} hydro={cfg['isHydrostatic']} {precision}", flush=True)
    return wvamd.WVTransformConstantStratificationKernel(cfg, precision)


@pytest.mark.parametrize("N,Nz,Lx,hydro,nsteps", [(16, 9, 50e3, False, 6), (16, 9, 50e3, True, 6),
                                                  (32, 17, 50e3, False, 4),
                                                  # runtime lengths: composite passes (60 = 15 x 4,
                                                  # 96 = 6 x 16) and the direct DFT (19)
                                                  (60, 9, 50e3, False, 3), (96, 17, 60e3, True, 2),
                                                  (19, 9, 50e3, False, 3),
                                                  # Nx = 256: the x kernel's 256-point forward transforms
                                                  (256, 9, 200e3, False, 2), (256, 9, 200e3, True, 2)])
def test_rk4_fp64_per_step_matches_oracle_small(torch_gpu, N, Nz, Lx, hydro, nsteps):
    cfg = _cfg(N, Nz, Lx, hydro)
    k = _kernel(cfg)
    o = Oracle(cfg)
    Oracle.set_threads(1)
    A = _flow(k, seed=17, jmax=k.Nj)
    h = _dt(k, A)
    dev = [k.to_device(a) for a in A]
    host = [a.copy() for a in A]
    t = 0.0
    for step in range(nsteps):
        t_gpu = k.rk4_advance(*dev, t=t, h=h, nsteps=1)
        *host, t = o.rk4(*host, t=t, dt=h, nsteps=1)
        assert t_gpu == t
        err = rel_err_joint([d.cpu().numpy() for d in dev], host)
        print(f"[gpu] step {step} rel err {err:.2e}", flush=True)
        assert err <= 1e-12


def test_rk4_fp64_256x256x9_no_antialias_per_step(torch_gpu):
    """Nx = 256 without the 2/3 rule (nkx = 129): the x kernel's first inverse
    pass then reads the whole spectrum (no zero band)."""
    import wvamd
    cfg = wvamd.make_config(256, 256, 9, Lx=200e3, shouldAntialias=False)
    k = _kernel(cfg)
    o = Oracle(cfg)
    Oracle.set_threads(1)
    A = _flow(k, seed=29, jmax=k.Nj)
    h = _dt(k, A)
    dev = [k.to_device(a) for a in A]
    host = [a.copy() for a in A]
    t = 0.0
    for step in range(2):
        k.rk4_advance(*dev, t=t, h=h, nsteps=1)
        *host, t = o.rk4(*host, t=t, dt=h, nsteps=1)
        err = rel_err_joint([d.cpu().numpy() for d in dev], host)
        print(f"[gpu] 256x256x9 no antialias step {step} rel err {err:.2e}", flush=True)
        assert err <= 1e-12


def test_rk4_fp64_64x64x33_50_steps_per_step(torch_gpu):
    """The 64x64x33 configuration (Lx=Ly=50 km, Lz=1300 m, N0=5.2e-3, lat 33,
    seeded flow j<=16, rms u 5 cm/s, dt = 0.5 CFL), 50 steps, every step."""
    cfg = _cfg(64, 33, 50e3)
    k = _kernel(cfg)
    o = Oracle(cfg)
    Oracle.set_threads(1)
    A = _flow(k, seed=4001, jmax=16)
    h = _dt(k, A)
    dev = [k.to_device(a) for a in A]
    host = [a.copy() for a in A]
    t = 0.0
    worst = 0.0
    for step in range(50):
        k.rk4_advance(*dev, t=t, h=h, nsteps=1)
        *host, t = o.rk4(*host, t=t, dt=h, nsteps=1)
        err = rel_err_joint([d.cpu().numpy() for d in dev], host)
        worst = max(worst, err)
        assert err <= 1e-12, (step, err)
    print(f"[gpu] 64x64x33 50 steps worst per-step rel err {worst:.2e} (dt={h:.3f}s)", flush=True)


def test_rk4_fp64_256x256x129_flux_and_increments_match_oracle(torch_gpu):
    """The 256x256x129 workload (seed 4001, j <= 64, rms 5 cm/s, dt = 0.5
    CFL): one FP64 flux against the oracle at 1e-12, then the per-step
    increments y_{n+1} - y_n of steps 1-5 and 196-200, each from the GPU's own
    state at step n, against the oracle's increment from that state at 1e-12
    of max|increment| (an increment is ~100x more sensitive than the state)."""
    cfg = _cfg(256, 129, 200e3)
    k = _kernel(cfg)
    A = _flow(k, seed=4001, jmax=64)
    h = _dt(k, A)
    o = Oracle(cfg)
    Oracle.set_threads(16)
    F = [f.cpu().numpy() for f in k.nonlinear_flux(*[k.to_device(a) for a in A], t=0.0)]
    print("[gpu] oracle 256x256x129 flux (CPU) ...", flush=True)
    err = rel_err_joint(F, o.nonlinear_flux(*A, t=0.0))
    print(f"[gpu] 256x256x129 flux rel err {err:.2e}", flush=True)
    assert err <= 1e-12
    dev = [k.to_device(a) for a in A]
    t = 0.0

    def check(step):
        nonlocal t
        y0 = [d.cpu().numpy() for d in dev]
        t_new = k.rk4_advance(*dev, t=t, h=h, nsteps=1)
        y1 = [d.cpu().numpy() for d in dev]
        ref = o.rk4(*y0, t=t, dt=h, nsteps=1)
        assert ref[3] == t_new
        e = rel_err_joint([a - b for a, b in zip(y1, y0)], [a - b for a, b in zip(ref[:3], y0)])
        print(f"[gpu] 256x256x129 step {step} increment rel err {e:.2e}", flush=True)
        assert e <= 1e-12, (step, e)
        t = t_new

    for step in range(1, 6):
        check(step)
    t = k.rk4_advance(*dev, t=t, h=h, nsteps=190)
    for step in range(196, 201):
        check(step)


def test_rk4_fp64_256x256x129_one_step_matches_oracle(torch_gpu):
    cfg = _cfg(256, 129, 200e3)
    k = _kernel(cfg)
    A = _flow(k, seed=4002, jmax=64)
    h = _dt(k, A)
    dev = [k.to_device(a) for a in A]
    k.rk4_advance(*dev, t=0.0, h=h, nsteps=1)
    gpu = [d.cpu().numpy() for d in dev]
    o = Oracle(cfg)
    Oracle.set_threads(16)
    print("[gpu] oracle 256x256x129 step (CPU) ...", flush=True)
    ref = o.rk4(*A, t=0.0, dt=h, nsteps=1)
    err = rel_err_joint(gpu, ref[:3])
    print(f"[gpu] 256x256x129 one-step rel err {err:.2e}", flush=True)
    assert err <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed-split16"])
def test_flux_conserves_energy_256x256x129(torch_gpu, precision):
    """sum TE*Re(conj(A) F) = 0 for the antialiased flux (UnitTests/TestNonlinearFlux.m:7-18);
    FP64 to 100 eps of the term magnitudes, mixed to its FP32 level."""
    cfg = _cfg(256, 129, 200e3)
    k = _kernel(cfg, precision)
    A = _flow(k, seed=4003, jmax=64)
    F = [f.cpu().numpy() for f in k.nonlinear_flux(*[k.to_device(a) for a in A])]
    Apm, A0t = k.coefficient_table("ApmTE"), k.coefficient_table("A0TE")
    terms = np.concatenate([(Apm * np.real(np.conj(A[0]) * F[0])).ravel(),
                            (Apm * np.real(np.conj(A[1]) * F[1])).ravel(),
                            (A0t * np.real(np.conj(A[2]) * F[2])).ravel()])
    ratio = abs(terms.sum()) / np.abs(terms).sum()
    print(f"[gpu] {precision} energy-flux imbalance {ratio:.2e}", flush=True)
    assert ratio <= (100 * np.finfo(float).eps if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision", ["mixed", "mixed-split16"])
def test_mixed_energy_tracks_fp64_256x256x129(torch_gpu, precision):
    cfg = _cfg(256, 129, 200e3)
    k64, k32 = _kernel(cfg, "fp64"), _kernel(cfg, precision)
    A = _flow(k64, seed=4004, jmax=64)
    h = _dt(k64, A)
    d64 = [k64.to_device(a) for a in A]
    d32 = [k32.to_device(a) for a in A]
    e0 = k64.total_energy(*d64)
    t = 0.0
    worst = 0.0
    for _ in range(5):
        t = k64.rk4_advance(*d64, t=t, h=h, nsteps=4)
        k32.rk4_advance(*d32, t=t - 4 * h, h=h, nsteps=4)
        e64, e32 = k64.total_energy(*d64), k32.total_energy(*d32)
        worst = max(worst, abs(e32 - e64) / e64)
    coef = rel_err_joint([d.cpu().numpy() for d in d32], [d.cpu().numpy() for d in d64])
    print(f"[gpu] E0={e0:.6e} E={e64:.6e} mixed-vs-fp64 worst rel energy err {worst:.2e}, "
          f"coefficients after 20 steps {coef:.2e}", flush=True)
    assert worst <= 1e-6
    assert coef <= 1e-5


@pytest.mark.parametrize("precision", ["mixed", "mixed-split16"])
def test_mixed_flux_256x256x129_matches_fp64_and_repeats_bitwise(torch_gpu, precision):
    """The mixed flux of a seeded flow at the benchmark size against the FP64
    flux (FP32 transforms: about 1e-5 of the largest component), and two
    evaluations of it bit for bit equal (no order- or timing-dependent
    arithmetic in any kernel at this size)."""
    cfg = _cfg(256, 129, 200e3)
    k64, k32 = _kernel(cfg, "fp64"), _kernel(cfg, precision)
    A = _flow(k64, seed=4005, jmax=64)
    d = [k32.to_device(a) for a in A]
    F1 = [f.cpu().numpy() for f in k32.nonlinear_flux(*d, t=3000.0, t0=0.0)]
    F2 = [f.cpu().numpy() for f in k32.nonlinear_flux(*d, t=3000.0, t0=0.0)]
    R = [f.cpu().numpy() for f in k64.nonlinear_flux(*[k64.to_device(a) for a in A], t=3000.0, t0=0.0)]
    errs = [float(np.abs(f - r).max() / np.abs(r).max()) for f, r in zip(F1, R)]
    print(f"[gpu] {precision} flux vs fp64 at 256x256x129: Fp {errs[0]:.2e} Fm {errs[1]:.2e} F0 {errs[2]:.2e}",
          flush=True)
    assert all(np.array_equal(a, b) for a, b in zip(F1, F2))
    assert max(errs) <= 5e-5
