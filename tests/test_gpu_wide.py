"""GPU parity at horizontal lengths above 512 points, where the x kernels
take a row's transforms in several in-LDS FFT calls of at most 4096 points
and the generic x transforms take fewer rows per workgroup, and at Ny = 1024,
where the y tiles are 4 columns wide (FP32 rows of 32 bytes: the tiles are
renumbered onto XCDs).

Bars: FP64 flux, inverse and forward transforms and APV against the oracle at
1e-12 relative (max|diff| / max|ref|, the inverse at 1e-13); the mixed flux
at 1e-4 against the oracle (FP64 at Nx = 2048 needs more LDS than a
workgroup has and is rejected); the scalar advection against its numpy
restatement at 1e-12; the mixed flux VJP by the dot-product test at 1e-5.
"""
import numpy as np
import pytest
from conftest import rel_err, rel_err_joint

from oracle import Oracle

pytestmark = pytest.mark.gpu

SHAPES = [(1024, 24, 9, False), (1024, 20, 9, True)]


def _state(o, seed, scale=1e-3):
    rng = np.random.default_rng(seed)
    return [scale * (rng.standard_normal((o.Nkl, o.Nj)) + 1j * rng.standard_normal((o.Nkl, o.Nj)))
            for _ in range(3)]


@pytest.mark.parametrize("Nx,Ny,Nz,hydro", SHAPES, ids=[f"{a}x{b}x{c}{'h' if h else ''}" for a, b, c, h in SHAPES])
def test_fp64_at_nx_1024_matches_oracle(torch_gpu, Nx, Ny, Nz, hydro):
    import wvamd
    cfg = wvamd.make_config(Nx, Ny, Nz, None, Lx=400e3, Ly=20e3, isHydrostatic=hydro)
    print(f"[gpu] wide fp64 {Nx}x{Ny}x{Nz} hydro={hydro}", flush=True)
    k = wvamd.WVTransformConstantStratificationKernel(cfg, "fp64")
    o = Oracle(cfg)
    A = _state(o, Nx + Ny)
    dev = [k.to_device(a) for a in A]
    F = [t.cpu().numpy() for t in k.nonlinear_flux(*dev, t=450.0, t0=0.0)]
    ferr = rel_err_joint(F, o.nonlinear_flux(*A, t=450.0, t0=0.0))
    fields = k.transform_wave_vortex_to_uvweta(*dev, t=450.0).cpu().numpy()
    ref = o.inverse_uvweta(*A, t=450.0)
    ierr = float(np.abs(fields - ref).max() / np.abs(ref).max())
    apv = k.apv(*dev, t=450.0).cpu().numpy()
    aerr = rel_err(apv, o.apv(*A, t=450.0))
    # forward transform on the fields of a physical flow (white-noise
    # coefficients make fields whose forward transform amplifies rounding
    # by about 1e6, in the oracle too): against the oracle and round trip
    from wvamd import flows
    md = k.modes()
    B = flows.random_flow_shape(md["Kh"], k.Nj, k.coefficient_table("ApmTE"), k.coefficient_table("A0TE"),
                                seed=5, jmax=k.Nj)
    sel = [0, 1, 3] if hydro else [0, 1, 2, 3]
    f = k.transform_wave_vortex_to_uvweta(*[k.to_device(b) for b in B], t=7.0)[sel].contiguous()
    back = [t.cpu().numpy() for t in k.transform_to_wave_vortex(f, t=7.0)]
    berr = rel_err_joint(back, o.forward(f.cpu().numpy(), t=7.0))
    rerr = rel_err_joint(back, B)
    print(f"[gpu]   flux {ferr:.2e} inverse {ierr:.2e} forward {berr:.2e} round trip {rerr:.2e} apv {aerr:.2e}",
          flush=True)
    assert ferr <= 1e-12 and ierr <= 1e-13 and berr <= 1e-12 and rerr <= 1e-12 and aerr <= 1e-12


@pytest.mark.parametrize("Nx,Ny", [(1024, 24), (2048, 16)])
def test_mixed_flux_at_wide_grids_matches_oracle(torch_gpu, Nx, Ny):
    import wvamd
    cfg = wvamd.make_config(Nx, Ny, 9, None, Lx=400e3, Ly=10e3)
    print(f"[gpu] wide mixed {Nx}x{Ny}x9", flush=True)
    k = wvamd.WVTransformConstantStratificationKernel(cfg, "mixed-split16")
    o = Oracle(cfg)
    A = _state(o, Nx)
    F = [t.cpu().numpy() for t in k.nonlinear_flux(*[k.to_device(a) for a in A], t=450.0, t0=0.0)]
    err = rel_err_joint(F, o.nonlinear_flux(*A, t=450.0, t0=0.0))
    print(f"[gpu]   mixed flux rel err {err:.2e}", flush=True)
    assert err <= 1e-4
    if Nx == 2048:
        with pytest.raises(wvamd.WVKernelError) as e:
            wvamd.WVTransformConstantStratificationKernel(cfg, "fp64")
        assert e.value.status == "unsupportedOperation"


def test_scalar_advection_at_nx_1024_matches_restatement(torch_gpu):
    import wvamd
    from scalar_advection import advect_f_grid_scalar
    cfg = wvamd.make_config(1024, 16, 9, Lx=400e3, Ly=8e3)
    print("[gpu] wide scalar advection 1024x16x9", flush=True)
    k = wvamd.WVTransformConstantStratificationKernel(cfg, "fp64")
    rng = np.random.default_rng(5)
    scalar = rng.standard_normal((9, 16, 1024))
    uvw = rng.standard_normal((3, 9, 16, 1024)) * 0.1
    rows = Oracle(cfg).modes()["row"]
    ds = torch_gpu.from_numpy(scalar).cuda()
    du = torch_gpu.from_numpy(uvw).cuda()
    for antialias in (True, False):
        got = k.advect_f_grid_scalar(ds, du, antialias).cpu().numpy()
        err = rel_err(got, advect_f_grid_scalar(cfg, scalar, uvw, antialias, rows))
        print(f"[gpu]   aa={antialias} rel err {err:.2e}", flush=True)
        assert err <= 1e-12


def test_mixed_flux_vjp_at_nx_1024_dot_product(torch_gpu):
    import wvamd
    cfg = wvamd.make_config(1024, 16, 9, None, Lx=400e3, Ly=8e3)
    print("[gpu] wide mixed flux vjp 1024x16x9", flush=True)
    k = wvamd.WVTransformConstantStratificationKernel(cfg, "mixed")
    rng = np.random.default_rng(9)

    def rnd(amp):
        return [k.to_device(amp * (rng.standard_normal(k.spectral_shape) + 1j * rng.standard_normal(k.spectral_shape)))
                for _ in range(3)]

    y, dy, lam = rnd(1e-2), rnd(1e-2), rnd(1.0)
    fp = k.nonlinear_flux(*[a + b for a, b in zip(y, dy)], t=700.0)
    fm = k.nonlinear_flux(*[a - b for a, b in zip(y, dy)], t=700.0)
    jdy = [(a - b) / 2 for a, b in zip(fp, fm)]
    jtl = k.flux_vjp(*y, *lam, t=700.0)
    lhs = sum(float((a.conj() * b).real.sum()) for a, b in zip(jdy, lam))
    rhs = sum(float((a.conj() * b).real.sum()) for a, b in zip(dy, jtl))
    scale = np.sqrt(sum(float((a.abs() ** 2).sum()) for a in jdy)) * np.sqrt(sum(float((a.abs() ** 2).sum()) for a in lam))
    print(f"[gpu]   <J dy, l> {lhs:.6e} <dy, J^T l> {rhs:.6e} rel {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * scale
    k64 = wvamd.WVTransformConstantStratificationKernel(cfg, "fp64")
    with pytest.raises(wvamd.WVKernelError) as e:
        k64.flux_vjp(*y, *lam, t=700.0)
    assert e.value.status == "unsupportedOperation"


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed-split16", 1e-4)])
def test_flux_at_ny_1024_matches_oracle(torch_gpu, prec, tol):
    import wvamd
    cfg = wvamd.make_config(16, 1024, 9, None, Lx=8e3, Ly=400e3)
    print(f"[gpu] tall {prec} 16x1024x9", flush=True)
    k = wvamd.WVTransformConstantStratificationKernel(cfg, prec)
    o = Oracle(cfg)
    A = _state(o, 1024)
    F = [t.cpu().numpy() for t in k.nonlinear_flux(*[k.to_device(a) for a in A], t=450.0, t0=0.0)]
    err = rel_err_joint(F, o.nonlinear_flux(*A, t=450.0, t0=0.0))
    print(f"[gpu]   flux rel err {err:.2e}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("Nx,Ny", [(768, 12), (527, 12)])
def test_fp64_rounds_of_whole_transforms_match_oracle(torch_gpu, Nx, Ny):
    """Runtime lengths whose 8 transforms per row exceed one in-LDS FFT call:
    768 = 3 x 256 (mixed-radix passes) and 527 = 17 x 31 (direct DFT), run in
    rounds of whole transforms."""
    import wvamd
    cfg = wvamd.make_config(Nx, Ny, 9, None, Lx=300e3, Ly=10e3)
    print(f"[gpu] rounds fp64 {Nx}x{Ny}x9", flush=True)
    k = wvamd.WVTransformConstantStratificationKernel(cfg, "fp64")
    o = Oracle(cfg)
    A = _state(o, Nx + 1)
    dev = [k.to_device(a) for a in A]
    F = [t.cpu().numpy() for t in k.nonlinear_flux(*dev, t=450.0, t0=0.0)]
    ferr = rel_err_joint(F, o.nonlinear_flux(*A, t=450.0, t0=0.0))
    fields = k.transform_wave_vortex_to_uvweta(*dev, t=450.0).cpu().numpy()
    ref = o.inverse_uvweta(*A, t=450.0)
    ierr = float(np.abs(fields - ref).max() / np.abs(ref).max())
    aerr = rel_err(k.apv(*dev, t=450.0).cpu().numpy(), o.apv(*A, t=450.0))
    print(f"[gpu]   flux {ferr:.2e} inverse {ierr:.2e} apv {aerr:.2e}", flush=True)
    assert ferr <= 1e-12 and ierr <= 1e-13 and aerr <= 1e-12
