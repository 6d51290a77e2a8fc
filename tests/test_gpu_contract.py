"""GPU tests of the kernel-contract members beyond the flux: state-field
derivatives, the flux with caller-provided advecting fields, metrics and
identifiers, WVFixedStepRK4::advanceToTime semantics (clipped final step,
restart determinism, a failed step leaves the state alone), and argument
validation -- all through the C-ABI.

Parity bar: FP64 results against the reference kernel's own outputs
(tests/golden/contract_*.npz, written by oracle/_ref) or the pinned oracle at
1e-12 relative (max|diff| / max|ref|); RK4 restart and failure semantics
bitwise.
"""
import ctypes
import os

import numpy as np
import pytest
from conftest import config_from_golden, golden_files, load_golden, rel_err, rel_err_joint

from oracle import Oracle

pytestmark = pytest.mark.gpu

CONTRACT = golden_files("contract_")
TOL = 1e-12


def _kernel(cfg, precision="fp64"):
    import wvamd
    print(f"[gpu] create {cfg['Nx']}x{cfg['Ny']}x{cfg['Nz']} Nj={cfg['Nj']} hydro={cfg['isHydrostatic']} "
          f"{precision}", flush=True)
    return wvamd.WVTransformConstantStratificationKernel(cfg, precision)


def _host(ts):
    return [t.cpu().numpy() for t in ts]


def _rand_state(k, seed, scale=1e-3):
    rng = np.random.default_rng(seed)
    return [scale * (rng.standard_normal(k.spectral_shape) + 1j * rng.standard_normal(k.spectral_shape))
            for _ in range(3)]


@pytest.mark.parametrize("path", CONTRACT, ids=[os.path.basename(p)[:-4] for p in CONTRACT])
def test_state_derivatives_and_flux_using_fields_match_reference(torch_gpu, path):
    g = load_golden(path)
    t, t0 = float(g["t"]), float(g["t0"])
    for prec, tol in (("fp64", TOL), ("mixed", 1e-4), ("mixed-split16", 1e-4)):
        k = _kernel(config_from_golden(g), prec)
        A = [k.to_device(a) for a in (g["Ap"], g["Am"], g["A0"])]
        for f, name in enumerate(("u", "v", "w", "eta")):
            d = k.transform_state_field_derivatives(*A, name, t=t, t0=t0).cpu().numpy()
            err = rel_err(d, g["derivatives"][f])
            print(f"[gpu] {prec} d/dx,y,z {name} rel err {err:.2e}", flush=True)
            assert err <= tol, name
        uvw = torch_gpu.as_tensor(np.ascontiguousarray(g["uvw"])).to("cuda")
        F = _host(k.nonlinear_flux_using_advection_fields(*A, uvw, t=t, t0=t0))
        err = rel_err_joint(F, (g["Fp"], g["Fm"], g["F0"]))
        print(f"[gpu] {prec} flux using fields rel err {err:.2e}", flush=True)
        assert err <= tol


def test_flux_with_fields_then_using_them_reproduces_the_flux(torch_gpu):
    """nonlinearFluxWithAdvectionFields returns the state's (u, v, w); feeding
    them back through nonlinearFluxUsingAdvectionFields gives the flux again
    (the runtime's shared-fields path, ...Kernel.cpp:1148-1171)."""
    import wvamd
    cfg = wvamd.make_config(32, 32, 17, 11)
    k = _kernel(cfg)
    o = Oracle(cfg)
    A = _rand_state(k, 3)
    dev = [k.to_device(a) for a in A]
    (F, uvw) = k.nonlinear_flux_with_advection_fields(*dev, t=40.0, t0=5.0)
    ref = o.nonlinear_flux(*A, t=40.0, t0=5.0)
    assert rel_err_joint(_host(F), ref) <= TOL
    assert rel_err(uvw.cpu().numpy(), o.inverse_uvw(*A, t=40.0, t0=5.0)) <= 1e-13
    F2 = _host(k.nonlinear_flux_using_advection_fields(*dev, uvw, t=40.0, t0=5.0))
    assert rel_err_joint(F2, ref) <= TOL


def test_midpoint_row_derivatives_at_nz_33(torch_gpu):
    """Nz = 33 (= 32 m + 1): the folded synthesis computes the midpoint row
    z = L/2 beside its MFMA tiles.  The i k and i l outputs of that row
    (F/G all-derivatives, state derivatives) must be written: two calls with
    different states must both match the oracle (a stale row would carry the
    first state into the second)."""
    import wvamd
    cfg = wvamd.make_config(32, 32, 33, 21)
    o = Oracle(cfg)
    for prec, tol in (("fp64", TOL), ("mixed", 1e-4), ("mixed-split16", 1e-4)):
        k = _kernel(cfg, prec)
        for seed in (11, 12):
            A = _rand_state(k, seed)
            dev = [k.to_device(a) for a in A]
            for fam in (True, False):
                mine = k.transform_to_spatial_domain_with_all_derivatives(fam, dev[0], dev[2]).cpu().numpy()
                assert rel_err(mine, o.fg_all(fam, A[0], A[2])) <= tol, (prec, seed, fam)
            for f in range(4):
                mine = k.transform_state_field_derivatives(*dev, f, t=9.0).cpu().numpy()
                assert rel_err(mine, o.state_field_derivatives(*A, f, t=9.0)) <= tol, (prec, seed, f)


def test_metrics_identifiers_and_scalar_preparation(torch_gpu):
    import wvamd
    k = _kernel(wvamd.make_config(16, 16, 9, 5))
    m = k.metrics()
    assert m["planCount"] == 17 and m["nonlinearFluxCallCount"] == 0
    assert m["descriptorBytes"] > 0 and m["planBytes"] > 0 and m["scratchCapacityBytes"] == k.scratch_bytes() > 0
    assert k.persistent_bytes() == m["descriptorBytes"] + m["planBytes"] + m["scratchCapacityBytes"]
    ids = k.identifiers()
    assert all(isinstance(v, str) and v for n, v in ids.items() if n.endswith("identifier") or n.endswith("identity"))
    assert ids["phase_reservation_bytes"] == (16 // 2 + 1) * 16 * 9 * 16
    A = [k.to_device(a) for a in _rand_state(k, 1)]
    k.nonlinear_flux(*A, t=1.0)
    k.nonlinear_flux(*A, t=2.0)
    m = k.metrics()
    assert m["nonlinearFluxCallCount"] == 2 and m["nonlinearFluxPhaseEvaluationCount"] == 2 * k.M
    assert m["horizontalExecutionCount"] == 2 * 9 and m["verticalExecutionCount"] == 2 * 19
    k.prepare_scalar_advection()
    k.prepare_scalar_advection()  # idempotent
    assert k.metrics()["planCount"] == 18


def _rk4_oracle_steps(o, A, t, steps, t0=0.0):
    Ap, Am, A0 = A
    for h in steps:
        Ap, Am, A0, t = o.rk4(Ap, Am, A0, t, h, 1, t0=t0)
    return [Ap, Am, A0], t


def test_advance_to_time_clips_the_final_step(torch_gpu):
    """WVFixedStepRK4::advanceToTime (WVRungeKutta.cpp:464-480; the reference
    test TestWVForcingAndRK4.cpp:253-274 advances 0 -> 0.725 with h = 0.1 in
    ... 3 steps there, here 3 full steps and a clipped fourth)."""
    import wvamd
    cfg = wvamd.make_config(16, 16, 9, 5, Lx=50e3)
    k = _kernel(cfg)
    o = Oracle(cfg)
    A = _rand_state(k, 21)
    dev = [k.to_device(a) for a in A]
    h, final = 60.0, 205.0
    t = k.rk4_advance_to_time(*dev, 0.0, final, h)
    assert t == final
    c = k.rk4_counters()
    assert c["stepCount"] == 4 and c["lastStepSize"] == 25.0
    ref, tr = _rk4_oracle_steps(o, A, 0.0, [60.0, 60.0, 60.0, 25.0])
    assert rel_err_joint(_host(dev), ref) <= TOL
    with pytest.raises(wvamd.WVKernelError) as e:
        k.rk4_advance_to_time(*dev, final, final - 1.0, h)
    assert e.value.status == "invalidConfiguration"
    with pytest.raises(wvamd.WVKernelError) as e:
        k.rk4_advance_to_time(*dev, final, float("inf"), h)
    assert e.value.status == "invalidConfiguration"
    before = _host(dev)
    assert k.rk4_advance_to_time(*dev, final, final, h) == final  # no step
    assert all(np.array_equal(a, b) for a, b in zip(_host(dev), before))


def test_restart_continuation_is_bitwise(torch_gpu):
    """TestWVForcingAndRK4.cpp:276-311: advancing 0 -> 0.6 -> 0.7 on one
    integrator equals advancing 0 -> 0.6 and continuing on a new one, bitwise;
    here also equal to one call 0 -> 0.7."""
    import wvamd
    cfg = wvamd.make_config(16, 16, 9, 5, Lx=50e3)
    k1, k2, k3, k4 = (_kernel(cfg) for _ in range(4))
    A = _rand_state(k1, 31)
    s1 = [k1.to_device(a) for a in A]
    s2 = [k2.to_device(a) for a in A]
    s4 = [k4.to_device(a) for a in A]
    h = 60.0
    t1 = k1.rk4_advance_to_time(*s1, 0.0, 360.0, h)
    t1 = k1.rk4_advance_to_time(*s1, t1, 420.0, h)
    t2 = k2.rk4_advance_to_time(*s2, 0.0, 360.0, h)
    t2 = k3.rk4_advance_to_time(*s2, t2, 420.0, h)  # a new handle continues
    t4 = k4.rk4_advance_to_time(*s4, 0.0, 420.0, h)
    assert t1 == t2 == t4 == 420.0
    for a, b, c in zip(_host(s1), _host(s2), _host(s4)):
        assert np.array_equal(a, b) and np.array_equal(a, c)


def test_failed_step_leaves_the_state_alone(torch_gpu):
    """TestWVForcingAndRK4.cpp:313-320: an RHS failure inside a step returns
    fftExecutionFailure and leaves the caller's state and time as they were
    before that step (steps already taken stay committed)."""
    import wvamd
    cfg = wvamd.make_config(16, 16, 9, 5, Lx=50e3)
    k = _kernel(cfg)
    lib = wvamd.library()
    A = _rand_state(k, 41)
    dev = [k.to_device(a) for a in A]
    before = _host(dev)
    # 7 launches for the first evaluation of a call (the evolved spectra,
    # synthesis, y transform, x kernel, y transform, analysis, stage), 6 for
    # each later one: launch 3 is inside step 1
    k.inject_launch_failure(3)
    t = ctypes.c_double(0.0)
    rc = lib.wv_rk4_advance_to_time_device(k._h, *[k._coef(a) for a in dev], ctypes.byref(t), 0.0, 240.0, 60.0,
                                           k._stream(None))
    torch_gpu.cuda.synchronize()
    assert rc == 8 and t.value == 0.0  # fftExecutionFailure
    assert all(np.array_equal(a, b) for a, b in zip(_host(dev), before))
    # step 1 takes launches 1-25; launch 30 is in step 2 (its first evaluation)
    k.inject_launch_failure(30)
    rc = lib.wv_rk4_advance_to_time_device(k._h, *[k._coef(a) for a in dev], ctypes.byref(t), 0.0, 240.0, 60.0,
                                           k._stream(None))
    torch_gpu.cuda.synchronize()
    assert rc == 8 and t.value == 60.0
    k2 = _kernel(cfg)
    ref = [k2.to_device(a) for a in A]
    k2.rk4_advance(*ref, 0.0, 60.0, 1)
    assert all(np.array_equal(a, b) for a, b in zip(_host(dev), _host(ref)))
    # the hook is spent: the next call runs normally
    assert k.rk4_advance_to_time(*dev, 60.0, 120.0, 60.0) == 120.0


def test_calls_from_two_streams_stay_ordered(torch_gpu):
    """Back-to-back calls on one handle from different streams share its
    workspaces; each call waits for the previous one."""
    import wvamd
    cfg = wvamd.make_config(64, 64, 33, 21, Lx=50e3)
    k = _kernel(cfg)
    o = Oracle(cfg)
    A, B = _rand_state(k, 51), _rand_state(k, 52)
    dA, dB = [k.to_device(a) for a in A], [k.to_device(b) for b in B]
    torch_gpu.cuda.synchronize()
    s1, s2 = torch_gpu.cuda.Stream(), torch_gpu.cuda.Stream()
    FA = k.nonlinear_flux(*dA, t=3.0, stream=s1)
    FB = k.nonlinear_flux(*dB, t=4.0, stream=s2)
    torch_gpu.cuda.synchronize()
    assert rel_err_joint(_host(FA), o.nonlinear_flux(*A, t=3.0)) <= TOL
    assert rel_err_joint(_host(FB), o.nonlinear_flux(*B, t=4.0)) <= TOL


def test_validation_before_any_write(torch_gpu):
    import wvamd
    k = _kernel(wvamd.make_config(16, 16, 9, 5))
    lib = wvamd.library()
    A = [k.to_device(a) for a in _rand_state(k, 61)]
    e = ctypes.c_double()
    assert lib.wv_total_energy_device(k._h, None, k._coef(A[1]), k._coef(A[2]), ctypes.byref(e), None) == 3
    with pytest.raises(wvamd.WVKernelError) as err:
        k.rk4_advance(A[0], A[0], A[2], 0.0, 60.0, 1)
    assert err.value.status == "overlappingArrays"
    # a flux output overlapping the state: nothing written, uvw included
    uvw = torch_gpu.full((3,) + k.spatial_shape, 7.0, dtype=torch_gpu.float64, device="cuda")
    with pytest.raises(wvamd.WVKernelError) as err:
        k.nonlinear_flux_with_advection_fields(*A, out=(A[0], k.empty_coefficients(1)[0],
                                                        k.empty_coefficients(1)[0]), uvw=uvw)
    assert err.value.status == "overlappingArrays"
    torch_gpu.cuda.synchronize()
    assert bool((uvw == 7.0).all())
    with pytest.raises(wvamd.WVKernelError) as err:
        k.transform_state_field_derivatives(*A, 4)
    assert err.value.status == "invalidConfiguration"


@pytest.mark.parametrize("hydrostatic", [True, False], ids=["hydrostatic", "nonhydrostatic"])
def test_flux_triad_energy_flux_matches_closed_form(torch_gpu, hydrostatic):
    """The reference's closed-form three-wave triad (UnitTests/TestNonlinearFlux.m:68-262)
    through the HIP flux: FP64 at 1e-12 relative per wave (the reference's
    bar is 1e-7), mixed precision at 1e-4 of the largest of the three."""
    import triad_kat
    import wvamd
    cfg = triad_kat.config(wvamd.make_config, hydrostatic)
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-4), ("mixed-split16", 1e-4)):
        k = _kernel(cfg, prec)
        Ap, idx, E = triad_kat.setup(cfg, k.modes(), k.coriolis_frequency())
        Z = np.zeros_like(Ap)
        F = _host(k.nonlinear_flux(k.to_device(Ap), k.to_device(Z), k.to_device(Z), t=0.0, t0=0.0))
        Ep = triad_kat.energy_flux(F[0], Ap, k.coefficient_table("ApmTE"), idx)
        err = np.max(np.abs(Ep - E) / (np.abs(E) if prec == "fp64" else np.max(np.abs(E))))
        print(f"[gpu] {prec} triad energy flux rel err {err:.2e}", flush=True)
        assert err <= tol, (prec, Ep, E)
