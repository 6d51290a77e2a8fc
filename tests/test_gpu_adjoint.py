"""GPU checks of the adjoint (vector-Jacobian) entry points.

The reference has no adjoint, so these are validated against the forward
path, which is itself pinned to the reference:
  dot-product tests  <J dy, L> = <dy, J^T L> with <a, b> = Re sum conj(a) b.
                     The flux is exactly quadratic in the state, so its
                     tangent-linear is exact as (F(y+dy) - F(y-dy))/2; the RK4
                     tangent-linear is built stage by stage from that.
                     FP64 tolerance 1e-12 (1e-10 over many steps).
  finite differences of the cost J = 1/2 |(u, v, eta)(T) - target|^2 along
                     seeded random directions (central differences).
  consistency        the checkpointed adjoint equals the step-by-step one.
"""
import numpy as np
import pytest
from conftest import rel_err_joint

pytestmark = pytest.mark.gpu


def _ip(a, b):
    return float(sum((x.conj() * y).real.sum().item() for x, y in zip(a, b)))


def _norm(a):
    return float(np.sqrt(sum((x.abs() ** 2).sum().item() for x in a)))


def _kernel(kind, shape, precision="fp64"):
    import wvamd
    Nx, Ny, Nz, Nj = shape
    if kind == "stratified":
        from wvamd import stratification as S
        cfg = wvamd.make_config(Nx, Ny, Nz, Nj, Lx=100e3, Lz=4000.0, isHydrostatic=True)
        st = S.Stratification(*S.exponential()[:1], Lz=4000.0, g=cfg["g"], dN2=S.exponential()[1])
        return wvamd.WVTransformHydrostaticKernel(cfg, S.vertical_operators(st, Nz, Nj), precision)
    cfg = wvamd.make_config(Nx, Ny, Nz, Nj, Lx=50e3, isHydrostatic=(kind == "hydrostatic"))
    return wvamd.WVTransformConstantStratificationKernel(cfg, precision)


def _random(k, seed, amp):
    rng = np.random.default_rng(seed)
    return [k.to_device(amp * (rng.standard_normal(k.spectral_shape) + 1j * rng.standard_normal(k.spectral_shape)))
            for _ in range(3)]


def _flux(k, y, t, t0=0.0):
    return k.nonlinear_flux(*y, t=t, t0=t0)


def _tlm_flux(k, y, dy, t, t0=0.0):
    """Exact tangent-linear of the quadratic flux: (F(y+dy) - F(y-dy)) / 2."""
    fp = _flux(k, [a + b for a, b in zip(y, dy)], t, t0)
    fm = _flux(k, [a - b for a, b in zip(y, dy)], t, t0)
    return [(a - b) / 2 for a, b in zip(fp, fm)]


def _tlm_rk4_step(k, y, dy, t, h, t0=0.0):
    """Tangent-linear of one RK4 step and the stepped state."""
    k1 = _flux(k, y, t, t0)
    d1 = _tlm_flux(k, y, dy, t, t0)
    s2 = [a + 0.5 * h * b for a, b in zip(y, k1)]
    e2 = [a + 0.5 * h * b for a, b in zip(dy, d1)]
    k2 = _flux(k, s2, t + 0.5 * h, t0)
    d2 = _tlm_flux(k, s2, e2, t + 0.5 * h, t0)
    s3 = [a + 0.5 * h * b for a, b in zip(y, k2)]
    e3 = [a + 0.5 * h * b for a, b in zip(dy, d2)]
    k3 = _flux(k, s3, t + 0.5 * h, t0)
    d3 = _tlm_flux(k, s3, e3, t + 0.5 * h, t0)
    s4 = [a + h * b for a, b in zip(y, k3)]
    e4 = [a + h * b for a, b in zip(dy, d3)]
    k4 = _flux(k, s4, t + h, t0)
    d4 = _tlm_flux(k, s4, e4, t + h, t0)
    ynew = [a + (h / 6) * (b1 + 2 * b2 + 2 * b3 + b4) for a, b1, b2, b3, b4 in zip(y, k1, k2, k3, k4)]
    dnew = [a + (h / 6) * (b1 + 2 * b2 + 2 * b3 + b4) for a, b1, b2, b3, b4 in zip(dy, d1, d2, d3, d4)]
    return ynew, dnew


CASES = [("nonhydrostatic", (16, 16, 9, 5)), ("hydrostatic", (16, 16, 9, 5)), ("nonhydrostatic", (32, 24, 17, 11)),
         ("stratified", (16, 16, 17, 12)), ("nonhydrostatic", (12, 10, 9, 8))]


@pytest.mark.parametrize("kind,shape", CASES, ids=[f"{k}-{'x'.join(map(str, s))}" for k, s in CASES])
def test_flux_vjp_dot_product(torch_gpu, kind, shape):
    k = _kernel(kind, shape)
    print(f"[gpu] flux vjp {kind} {shape}", flush=True)
    y, dy, lam = _random(k, 1, 1e-2), _random(k, 2, 1e-2), _random(k, 3, 1.0)
    jdy = _tlm_flux(k, y, dy, t=700.0, t0=100.0)
    jtl = k.flux_vjp(*y, *lam, t=700.0, t0=100.0)
    lhs, rhs = _ip(jdy, lam), _ip(dy, jtl)
    scale = _norm(jdy) * _norm(lam)
    print(f"[gpu] <J dy, l> = {lhs:.6e}  <dy, J^T l> = {rhs:.6e}  rel {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale


def test_flux_vjp_mixed_tracks_fp64(torch_gpu):
    k64 = _kernel("nonhydrostatic", (32, 24, 17, 11))
    k32 = _kernel("nonhydrostatic", (32, 24, 17, 11), "mixed")
    y, lam = _random(k64, 1, 1e-2), _random(k64, 3, 1.0)
    g64 = [t.cpu().numpy() for t in k64.flux_vjp(*y, *lam, t=300.0)]
    g32 = [t.cpu().numpy() for t in k32.flux_vjp(*y, *lam, t=300.0)]
    err = rel_err_joint(g32, g64)
    print(f"[gpu] mixed flux vjp vs fp64 rel err {err:.2e}", flush=True)
    assert err <= 1e-4


@pytest.mark.parametrize("kind", ["nonhydrostatic", "hydrostatic", "stratified"])
def test_inverse_transform_vjp_dot_product(torch_gpu, kind):
    k = _kernel(kind, (16, 16, 17, 12) if kind == "stratified" else (24, 16, 9, 6))
    dy = _random(k, 5, 1.0)
    fb = torch_gpu.as_tensor(np.random.default_rng(6).standard_normal((4,) + k.spatial_shape)).cuda()
    ldy = k.transform_wave_vortex_to_uvweta(*dy, t=50.0)
    ltf = k.transform_wave_vortex_to_uvweta_vjp(fb, t=50.0)
    lhs = float((ldy * fb).sum().item())
    rhs = _ip(dy, ltf)
    scale = float(ldy.norm().item() * fb.norm().item())
    print(f"[gpu] transform vjp {kind}: rel {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale


@pytest.mark.parametrize("kind", ["nonhydrostatic", "stratified"])
def test_rk4_step_vjp_dot_product(torch_gpu, kind):
    k = _kernel(kind, (16, 16, 17, 12) if kind == "stratified" else (16, 16, 9, 5))
    y, dy, lam = _random(k, 11, 1e-3), _random(k, 12, 1e-3), _random(k, 13, 1.0)
    h, t = 200.0, 40.0
    _, dnew = _tlm_rk4_step(k, y, dy, t, h)
    lb = [a.clone() for a in lam]
    k.rk4_step_vjp(*y, t, h, *lb)
    lhs, rhs = _ip(dnew, lam), _ip(dy, lb)
    scale = _norm(dnew) * _norm(lam)
    print(f"[gpu] rk4 step vjp {kind}: rel {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale


def _flow(k, seed, urms=0.05):
    """Seeded |k|^-2 flow of the benchmark's kind, scaled to rms horizontal
    speed urms, and its 0.5-CFL time step."""
    from wvamd import flows
    md = k.modes()
    A = flows.random_flow_shape(md["Kh"], k.Nj, k.coefficient_table("ApmTE"), k.coefficient_table("A0TE"),
                                seed=seed, jmax=k.Nj)
    f = k.transform_wave_vortex_to_uvw(*[k.to_device(a) for a in A]).cpu().numpy()
    sc = flows.scale_factor_for_rms(f, k.config["Lz"], urms)
    A = [a * sc for a in A]
    f = k.transform_wave_vortex_to_uvw(*[k.to_device(a) for a in A]).cpu().numpy()
    dt = flows.cfl_time_step(f, md["k"], md["l"], k.coefficient_table("omega"), k.config["Lz"], 0.5)
    return [k.to_device(a) for a in A], dt


def test_checkpointed_adjoint_equals_stepwise(torch_gpu):
    k = _kernel("nonhydrostatic", (16, 16, 9, 5))
    y0, h = _flow(k, 21)
    lam = _random(k, 22, 1.0)
    t, n = 0.0, 6
    states = [[a.clone() for a in y0]]
    y = [a.clone() for a in y0]
    for i in range(n):
        k.rk4_advance(*y, t=t + i * h, h=h, nsteps=1)
        states.append([a.clone() for a in y])
    lb = [a.clone() for a in lam]
    for i in reversed(range(n)):
        k.rk4_step_vjp(*states[i], t + i * h, h, *lb)
    lc = [a.clone() for a in lam]
    k.rk4_adjoint(*y0, t, h, n, 4, *lc)
    err = rel_err_joint([a.cpu().numpy() for a in lc], [a.cpu().numpy() for a in lb])
    assert err <= 1e-14, err


def _cost(k, y0, t, h, n, target, mask):
    y = [a.clone() for a in y0]
    k.rk4_advance(*y, t=t, h=h, nsteps=n)
    f = k.transform_wave_vortex_to_uvweta(*y, t=t + n * h)
    r = (f - target) * mask
    return 0.5 * float((r * r).sum().item()), y, r


def _gradient(k, y0, t, h, n, target, mask, interval):
    J, yT, r = _cost(k, y0, t, h, n, target, mask)
    lam = k.transform_wave_vortex_to_uvweta_vjp(r.contiguous(), t=t + n * h)
    k.rk4_adjoint(*y0, t, h, n, interval, *lam)
    return J, lam


def _dot_test(k, y0, t, h, n, interval, seed):
    """n-step dot-product test against the exact tangent-linear; the
    perturbation is a flow of the same kind as the state (the tangent-linear
    is exact for any size of it)."""
    dy = _flow(k, seed)[0]
    lam = _random(k, seed + 1, 1.0)
    y, dt = [a.clone() for a in y0], dy
    for i in range(n):
        y, dt = _tlm_rk4_step(k, y, dt, t + i * h, h)
    lb = [a.clone() for a in lam]
    k.rk4_adjoint(*y0, t, h, n, interval, *lb)
    lhs, rhs = _ip(dt, lam), _ip(dy, lb)
    return abs(lhs - rhs) / (_norm(dt) * _norm(lam))


def _fd_test(k, y0, t, h, n, interval, target, mask, grad, ndir=8, seed=40, eps=1e-4):
    """Central differences of J along seeded directions that are flows of the
    same kind as the state (eps times a 5 cm/s rms flow)."""
    worst = 0.0
    for i in range(ndir):
        d = _flow(k, seed + i)[0]
        Jp = _cost(k, [a + eps * b for a, b in zip(y0, d)], t, h, n, target, mask)[0]
        Jm = _cost(k, [a - eps * b for a, b in zip(y0, d)], t, h, n, target, mask)[0]
        fd = (Jp - Jm) / (2 * eps)
        an = _ip(d, grad)
        rel = abs(fd - an) / abs(an)
        assert np.isfinite(rel), (fd, an)
        worst = max(worst, rel)
    return worst


def _target_and_mask(torch_gpu, k, y1, t, h, n):
    mask = torch_gpu.ones((4,) + k.spatial_shape, dtype=torch_gpu.float64, device="cuda")
    mask[2] = 0.0  # J over (u, v, eta)
    yt = [a.clone() for a in y1]
    k.rk4_advance(*yt, t=t, h=h, nsteps=n)
    return k.transform_wave_vortex_to_uvweta(*yt, t=t + n * h), mask


@pytest.mark.parametrize("kind,shape,n,interval", [("nonhydrostatic", (16, 16, 9, 5), 8, 3),
                                                   ("stratified", (32, 32, 17, 16), 10, 4)])
def test_cost_gradient_matches_finite_differences_and_tlm(torch_gpu, kind, shape, n, interval):
    k = _kernel(kind, shape)
    y0, h = _flow(k, 31)
    y1, _ = _flow(k, 32)
    t = 0.0
    target, mask = _target_and_mask(torch_gpu, k, y1, t, h, n)
    J, grad = _gradient(k, y0, t, h, n, target, mask, interval)
    worst = _fd_test(k, y0, t, h, n, interval, target, mask, grad)
    print(f"[gpu] {kind} {n}-step cost gradient vs FD worst rel {worst:.2e}", flush=True)
    assert worst <= 1e-6
    rel = _dot_test(k, y0, t, h, n, interval, 50)
    print(f"[gpu] {kind} {n}-step dot-product rel {rel:.2e}", flush=True)
    assert rel <= 1e-10


def test_c4_adjoint_256x256x65_200_steps(torch_gpu):
    """The C4 workload: exponential stratification, 256x256 horizontally, 65
    quadrature points / 64 modes, 200 RK4 steps checkpointed every 20, cost
    1/2 |(u, v, eta)(T) - target|^2 with the target the forward run of a
    second seeded state.  FP64 gradient against central differences along 8
    seeded directions and the dot-product test over the full window; the
    mixed-precision gradient's error against FP64 is reported."""
    import time
    import wvamd
    from wvamd import stratification as S
    cfg = wvamd.make_config(256, 256, 65, 64, Lx=200e3, Lz=4000.0, isHydrostatic=True)
    st = S.Stratification(*S.exponential()[:1], Lz=4000.0, g=cfg["g"], dN2=S.exponential()[1])
    ops = S.vertical_operators(st, 65, 64)
    k = wvamd.WVTransformHydrostaticKernel(cfg, ops, "fp64")
    y0, h = _flow(k, 4001, urms=0.1)
    y1, _ = _flow(k, 4002, urms=0.1)
    t, n, interval = 0.0, 200, 20
    target, mask = _target_and_mask(torch_gpu, k, y1, t, h, n)
    torch_gpu.cuda.synchronize()
    t0 = time.perf_counter()
    J, grad = _gradient(k, y0, t, h, n, target, mask, interval)
    torch_gpu.cuda.synchronize()
    print(f"[gpu] C4 fp64: J={J:.6e}, forward+adjoint {time.perf_counter() - t0:.2f} s (dt={h:.1f}s)", flush=True)
    worst = _fd_test(k, y0, t, h, n, interval, target, mask, grad)
    print(f"[gpu] C4 fp64 gradient vs FD (8 directions) worst rel {worst:.2e}", flush=True)
    assert worst <= 1e-8
    rel = _dot_test(k, y0, t, h, n, interval, 4100)
    print(f"[gpu] C4 fp64 {n}-step dot-product rel {rel:.2e}", flush=True)
    assert rel <= 1e-10
    k32 = wvamd.WVTransformHydrostaticKernel(cfg, ops, "mixed")
    _, grad32 = _gradient(k32, y0, t, h, n, target, mask, interval)
    err = rel_err_joint([a.cpu().numpy() for a in grad32], [a.cpu().numpy() for a in grad])
    print(f"[gpu] C4 mixed gradient vs fp64 rel err {err:.2e}", flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("kind", ["nonhydrostatic", "stratified"])
def test_adjoint_gradient_matches_fd_of_cpu_reference(torch_gpu, kind):
    """The GPU adjoint's gradient of J = 1/2 |(u, v, eta)(T) - target|^2
    against central differences of the same cost evaluated entirely on the
    CPU by the FP64 oracle (liboracle for constant N, the numpy oracle for the
    exponential stratification): forward RK4, inverse transform and cost."""
    from hydrostatic import HydrostaticOracle
    from oracle import Oracle
    shape = (16, 16, 9, 5) if kind == "nonhydrostatic" else (16, 16, 17, 12)
    k = _kernel(kind, shape)
    y0, h = _flow(k, 61)
    y1, _ = _flow(k, 62)
    t, n, interval = 0.0, 6, 4
    target, mask = _target_and_mask(torch_gpu, k, y1, t, h, n)
    J, grad = _gradient(k, y0, t, h, n, target, mask, interval)
    tgt, msk = target.cpu().numpy(), mask.cpu().numpy()
    if kind == "nonhydrostatic":
        o = Oracle(k.config)

        def fields(y):
            Ap, Am, A0, tt = o.rk4(*y, t=t, dt=h, nsteps=n)
            return o.inverse_uvweta(Ap, Am, A0, t=tt)
    else:
        ho = HydrostaticOracle(k.config, k.operators, Oracle(k.config).modes())

        def fields(y):
            Ap, Am, A0, tt = ho.rk4(*y, t=t, dt=h, nsteps=n)
            return ho.fields(Ap, Am, A0, t=tt)

    def cost(y):
        r = (fields(y) - tgt) * msk
        return 0.5 * float((r * r).sum())

    yh = [a.cpu().numpy() for a in y0]
    assert abs(cost(yh) - J) <= 1e-10 * abs(J)
    eps, worst = 1e-4, 0.0
    for i in range(4):
        d = [a.cpu().numpy() for a in _flow(k, 70 + i)[0]]
        fd = (cost([a + eps * b for a, b in zip(yh, d)]) - cost([a - eps * b for a, b in zip(yh, d)])) / (2 * eps)
        an = _ip([k.to_device(x) for x in d], grad)
        worst = max(worst, abs(fd - an) / abs(an))
    print(f"[gpu] {kind} gradient vs FD of the CPU reference worst rel {worst:.2e}", flush=True)
    assert worst <= 1e-7
