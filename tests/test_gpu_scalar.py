"""GPU advectFGridScalar (wv_advect_f_grid_scalar_device,
WVTransformConstantStratificationKernel.cpp:1032-1138) through the C-ABI:
against the reference kernel's own outputs (tests/golden/scalar_*.npz) and
against the numpy restatement (oracle/scalar_advection.py) at larger grids.

Tolerance: 1e-12 relative (max|diff| / max|ref|), the FP64 bar of the flux.
The path is FP64 in either precision.
"""
import numpy as np
import pytest
from conftest import config_from_golden, golden_files, load_golden, rel_err

from oracle import Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
SCALAR_GOLDEN = golden_files("scalar_")


def _real(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def _kernel(cfg, precision="fp64"):
    import wvamd
    print(f"[gpu] scalar advection {cfg['Nx']}x{cfg['Ny']}x{cfg['Nz']} hydro={cfg['isHydrostatic']} "
          f"aa={cfg['shouldAntialias']} {precision}", flush=True)
    return wvamd.WVTransformConstantStratificationKernel(cfg, precision)


@pytest.mark.parametrize("path", SCALAR_GOLDEN, ids=[p.split("/")[-1][:-4] for p in SCALAR_GOLDEN])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_scalar_advection_matches_reference_golden(torch_gpu, path, precision):
    g = load_golden(path)
    k = _kernel(config_from_golden(g), precision)
    scalar, uvw = _real(torch_gpu, g["scalar"]), _real(torch_gpu, g["uvw"])
    for aa, key in ((False, "rhs"), (True, "rhs_antialiased")):
        got = k.advect_f_grid_scalar(scalar, uvw, aa).cpu().numpy()
        err = rel_err(got, g[key])
        print(f"[gpu] {key} rel err {err:.2e}", flush=True)
        assert err <= TOL, (key, err)


@pytest.mark.parametrize("N,Nz,hydro,aa", [(64, 33, False, True), (256, 33, False, True), (96, 17, True, False),
                                             (128, 17, False, True), (512, 9, False, True)])
def test_scalar_advection_matches_oracle(torch_gpu, N, Nz, hydro, aa):
    """Larger grids against the numpy restatement: x lengths 64, 128, 256 and
    512 (the transforms compiled for their length) and 96 (runtime length);
    calling twice checks that the kept scratch holds no state."""
    import wvamd
    from scalar_advection import advect_f_grid_scalar
    cfg = wvamd.make_config(N, N, Nz, Lx=100e3, isHydrostatic=hydro, shouldAntialias=aa)
    k = _kernel(cfg)
    rng = np.random.default_rng(N + Nz)
    scalar = rng.standard_normal((Nz, N, N))
    uvw = rng.standard_normal((3, Nz, N, N)) * 0.1
    rows = Oracle(cfg).modes()["row"]
    ds, du = _real(torch_gpu, scalar), _real(torch_gpu, uvw)
    for antialias in (True, False):
        want = advect_f_grid_scalar(cfg, scalar, uvw, antialias, rows)
        for _ in range(2):
            got = k.advect_f_grid_scalar(ds, du, antialias).cpu().numpy()
            err = rel_err(got, want)
            print(f"[gpu] {N}x{N}x{Nz} aa={antialias} rel err {err:.2e}", flush=True)
            assert err <= TOL


def test_scalar_advection_rejects_null_and_overlap(torch_gpu):
    import ctypes as C
    import wvamd
    k = _kernel(wvamd.make_config(16, 16, 9, Lx=50e3))
    lib = wvamd.library()
    f64 = dict(dtype=torch_gpu.float64, device="cuda:0")
    scalar, out = torch_gpu.zeros(k.spatial_shape, **f64), torch_gpu.zeros(k.spatial_shape, **f64)
    uvw = torch_gpu.zeros((3,) + k.spatial_shape, **f64)

    def status(sp, up, op):
        rc = lib.wv_advect_f_grid_scalar_device(k._h, sp, up, 1, op, None)
        return wvamd.STATUS_NAMES[rc]

    p = lambda t, off=0: C.c_void_p(t.data_ptr() + off)
    assert status(None, p(uvw), p(out)) == "invalidPointer"
    assert status(p(scalar), None, p(out)) == "invalidPointer"
    assert status(p(scalar), p(uvw), p(scalar)) == "overlappingArrays"
    assert status(p(scalar), p(uvw), p(uvw, 8)) == "overlappingArrays"
    assert status(p(scalar), p(uvw), p(out)) == "success"
    with pytest.raises(wvamd.WVKernelError) as e:
        k.advect_f_grid_scalar(scalar, uvw, True, out=scalar)
    assert e.value.status == "overlappingArrays"
