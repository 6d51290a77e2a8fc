"""GPU parity of the HIP path (libwvamd.so) against the golden vectors written by
the reference kernel and against the CPU oracle, through the C-ABI.

Tolerances (relative, max|diff| / max|ref| over Fp, Fm, F0 jointly):
  FP64 path     1e-12 per flux and per RK4 step (the reference's own
                compiled-vs-MATLAB bar, wvCompiledBackendBuild.m:356-398)
  mixed path    1e-4 per flux on white-noise coefficients (FP32 fields, spectra,
                FFTs and vertical MFMA; white noise maximizes cancellation in the
                products -- realistic flows are checked through energy instead)
"""
import os

import numpy as np
import pytest
from conftest import GOLDEN, config_from_golden, golden_files, load_golden, rel_err, rel_err_joint

pytestmark = pytest.mark.gpu

TABLES = ["omega", "fWaveScale", "UApField", "VApField", "WApField", "NApField", "UA0Field", "VA0Field",
          "NA0Field", "A0FromVorticity", "A0FromBuoyancy", "ApmDProjection", "ApmNProjection", "ApmDScaled",
          "z", "h0", "verticalWavenumber", "Fg", "Gg", "inertialScale", "gWaveScale", "apmWProjectionPrefactor"]
GOLDEN_TINY = [p for p in golden_files("golden_c") if "flow" not in os.path.basename(p)]
IDS = [os.path.basename(p)[:-4] for p in GOLDEN_TINY]
FP64_TOL = 1e-12
MIXED_TOL = 1e-4


def _kernel(cfg, precision="fp64"):
    import wvamd
    print(f"[gpu] create {cfg['Nx']}x{cfg['Ny']}x{cfg['Nz']} Nj={cfg['Nj']} hydro={cfg['isHydrostatic']} "
          f"{precision}", flush=True)
    return wvamd.WVTransformConstantStratificationKernel(cfg, precision)


def _dev(k, arrays):
    return [k.to_device(a) for a in arrays]


def _host(ts):
    return [t.cpu().numpy() for t in ts]


@pytest.mark.parametrize("path", GOLDEN_TINY, ids=IDS)
def test_descriptor_bitwise(torch_gpu, path):
    g = load_golden(path)
    k = _kernel(config_from_golden(g))
    md = k.modes()
    assert np.array_equal(md["kMode"], g["kMode"]) and np.array_equal(md["lMode"], g["lMode"])
    assert np.array_equal(md["row"], g["row"]) and np.array_equal(md["conj"], g["conj"])
    for name in TABLES:
        assert np.array_equal(k.table(name), g["T_" + name]), name
    assert k.coriolis_frequency() == float(g["coriolis"])


@pytest.mark.parametrize("path", golden_files("descriptor_"),
                         ids=[os.path.basename(p)[:-4] for p in golden_files("descriptor_")])
def test_descriptor_large(torch_gpu, path):
    g = load_golden(path)
    k = _kernel(config_from_golden(g))
    md = k.modes()
    assert np.array_equal(md["kMode"], g["kMode"]) and np.array_equal(md["lMode"], g["lMode"])
    for name in TABLES:
        t = k.table(name)
        if "T_" + name in g:
            assert np.array_equal(t, g["T_" + name]), name
        else:
            assert np.array_equal(t[g["S_" + name + "_idx"]], g["S_" + name]), name


@pytest.mark.parametrize("path", GOLDEN_TINY, ids=IDS)
def test_flux_fp64_matches_reference_golden(torch_gpu, path):
    g = load_golden(path)
    k = _kernel(config_from_golden(g))
    F = _host(k.nonlinear_flux(*_dev(k, (g["Ap"], g["Am"], g["A0"])), t=float(g["t"]), t0=float(g["t0"])))
    torch_gpu.cuda.synchronize()
    err = rel_err_joint(F, (g["Fp"], g["Fm"], g["F0"]))
    print(f"[gpu] fp64 flux rel err {err:.2e}", flush=True)
    assert err <= FP64_TOL


@pytest.mark.parametrize("precision", ["mixed", "mixed-split16"])
@pytest.mark.parametrize("path", GOLDEN_TINY, ids=IDS)
def test_flux_mixed_matches_reference_golden(torch_gpu, path, precision):
    g = load_golden(path)
    k = _kernel(config_from_golden(g), precision)
    F = _host(k.nonlinear_flux(*_dev(k, (g["Ap"], g["Am"], g["A0"])), t=float(g["t"]), t0=float(g["t0"])))
    err = rel_err_joint(F, (g["Fp"], g["Fm"], g["F0"]))
    print(f"[gpu] {precision} flux rel err {err:.2e}", flush=True)
    assert err <= MIXED_TOL


@pytest.mark.parametrize("path", [p for p in GOLDEN_TINY if "inv_uvweta" in np.load(p).files],
                         ids=[i for p, i in zip(GOLDEN_TINY, IDS) if "inv_uvweta" in np.load(p).files])
def test_inverse_transforms_match_reference_golden(torch_gpu, path):
    g = load_golden(path)
    k = _kernel(config_from_golden(g))
    A = _dev(k, (g["Ap"], g["Am"], g["A0"]))
    t, t0 = float(g["t"]), float(g["t0"])
    assert rel_err(k.transform_wave_vortex_to_uvweta(*A, t=t, t0=t0).cpu().numpy(), g["inv_uvweta"]) <= 1e-13
    assert rel_err(k.transform_wave_vortex_to_uvw(*A, t=t, t0=t0).cpu().numpy(), g["inv_uvw"]) <= 1e-13


def test_zero_state_gives_exactly_zero_flux(torch_gpu):
    import wvamd
    for prec in ("fp64", "mixed", "mixed-split16"):
        k = _kernel(wvamd.make_config(16, 16, 9, 5), prec)
        Z = k.empty_coefficients()
        F = _host(k.nonlinear_flux(*Z, t=0.5))
        assert all(np.all(x == 0) for x in F)


def test_flux_rejects_overlap_and_bad_shapes(torch_gpu):
    import wvamd
    k = _kernel(wvamd.make_config(8, 8, 5, 3))
    A = k.empty_coefficients()
    with pytest.raises(wvamd.WVKernelError) as e:
        k.nonlinear_flux(*A, out=(A[0], k.empty_coefficients(1)[0], k.empty_coefficients(1)[0]))
    assert e.value.status == "overlappingArrays"
    bad = torch_gpu.zeros((k.Nkl + 1, k.Nj), dtype=torch_gpu.complex128, device="cuda")
    with pytest.raises(wvamd.WVKernelError) as e:
        k.nonlinear_flux(bad, A[1], A[2])
    assert e.value.status == "invalidShape"
    with pytest.raises(wvamd.WVKernelError) as e:
        k.nonlinear_flux(*A, t=float("nan"))
    assert e.value.status == "invalidConfiguration"


def test_create_rejects_invalid_configuration(torch_gpu):
    import wvamd
    cfg = wvamd.make_config(8, 8, 5, 3)
    cfg["Nj"] = 5
    with pytest.raises(wvamd.WVKernelError) as e:
        wvamd.WVTransformConstantStratificationKernel(cfg)
    assert e.value.status == "invalidConfiguration"
    cfg = wvamd.make_config(8, 8, 5, 3)
    cfg["contractVersion"] = 3
    with pytest.raises(wvamd.WVKernelError):
        wvamd.WVTransformConstantStratificationKernel(cfg)


def test_flux_fp64_64x64x33_flow_matches_reference(torch_gpu):
    path = os.path.join(GOLDEN, "golden_c64x64x33_flow.npz")
    if not os.path.exists(path):
        pytest.skip("64x64x33 reference flux fixture not generated")
    g = load_golden(path)
    k = _kernel(config_from_golden(g))
    F = _host(k.nonlinear_flux(*_dev(k, (g["Ap"], g["Am"], g["A0"])), t=float(g["t"]), t0=float(g["t0"])))
    assert rel_err_joint(F, (g["Fp"], g["Fm"], g["F0"])) <= FP64_TOL
    for prec in ("mixed", "mixed-split16"):
        km = _kernel(config_from_golden(g), prec)
        Fm_ = _host(km.nonlinear_flux(*_dev(km, (g["Ap"], g["Am"], g["A0"])), t=float(g["t"]), t0=float(g["t0"])))
        assert rel_err_joint(Fm_, (g["Fp"], g["Fm"], g["F0"])) <= MIXED_TOL, prec


@pytest.mark.parametrize("path", [p for p in GOLDEN_TINY if "fwd_fields" in np.load(p).files],
                         ids=[i for p, i in zip(GOLDEN_TINY, IDS) if "fwd_fields" in np.load(p).files])
def test_forward_transform_matches_reference_golden(torch_gpu, path):
    g = load_golden(path)
    k = _kernel(config_from_golden(g))
    fields = torch_gpu.as_tensor(np.ascontiguousarray(g["fwd_fields"])).to("cuda")
    out = _host(k.transform_to_wave_vortex(fields, t=float(g["t"]), t0=float(g["t0"])))
    for mine, ref in zip(out, (g["fwd_Ap"], g["fwd_Am"], g["fwd_A0"])):
        assert rel_err(mine, ref) <= 1e-13


def test_forward_inverts_inverse_64x64x33(torch_gpu):
    import wvamd
    from wvamd import flows
    for hydro in (False, True):
        k = _kernel(wvamd.make_config(64, 64, 33, 21, Lx=50e3, isHydrostatic=hydro))
        md = k.modes()
        A = flows.random_flow_shape(md["Kh"], k.Nj, k.coefficient_table("ApmTE"), k.coefficient_table("A0TE"),
                                    seed=5, jmax=k.Nj)
        dev = _dev(k, A)
        f = k.transform_wave_vortex_to_uvweta(*dev, t=7.0)
        f = f[[0, 1, 3]].contiguous() if hydro else f
        back = _host(k.transform_to_wave_vortex(f, t=7.0))
        assert rel_err_joint(back, A) <= 1e-12


@pytest.mark.parametrize("path", [p for p in GOLDEN_TINY if "fAll" in np.load(p).files],
                         ids=[i for p, i in zip(GOLDEN_TINY, IDS) if "fAll" in np.load(p).files])
def test_fg_all_derivatives_match_reference_golden(torch_gpu, path):
    g = load_golden(path)
    k = _kernel(config_from_golden(g))
    Apm, A0 = _dev(k, (g["Ap"], g["A0"]))
    assert rel_err(k.transform_to_spatial_domain_with_all_derivatives(True, Apm, A0).cpu().numpy(), g["fAll"]) <= 1e-13
    assert rel_err(k.transform_to_spatial_domain_with_all_derivatives(False, Apm, A0).cpu().numpy(), g["gAll"]) <= 1e-13


# Horizontal lengths that are not powers of two: the in-LDS FFTs run
# mixed-radix passes (composite 15, 18, 6, 12, 10, 14 and 9 where they apply,
# 3, 5, 7, 11, 13, then 16 and 8/4/2) for lengths whose prime factors are at
# most 13 (144 = 9 x 16, 288 = 18 x 16, 18 in one pass; 384 = 3 x 16 x 8,
# 320 = 5 x 16 x 4), and a direct DFT for the rest (17 x 19).
ODD_SHAPES = [(96, 80, 17, False), (45, 36, 9, False), (26, 22, 9, True), (48, 39, 9, False), (17, 19, 9, False),
              (144, 18, 9, False), (288, 36, 9, True), (384, 20, 9, False), (320, 24, 9, True)]


@pytest.mark.parametrize("Nx,Ny,Nz,hydro", ODD_SHAPES, ids=[f"{a}x{b}x{c}{'h' if h else ''}" for a, b, c, h in ODD_SHAPES])
def test_flux_and_transforms_fp64_non_power_of_two_match_oracle(torch_gpu, Nx, Ny, Nz, hydro):
    import wvamd
    from oracle import Oracle
    cfg = wvamd.make_config(Nx, Ny, Nz, None, Lx=60e3, Ly=50e3, isHydrostatic=hydro)
    k = _kernel(cfg)
    o = Oracle(cfg)
    rng = np.random.default_rng(Nx * 1000 + Ny)
    A = [1e-3 * (rng.standard_normal(k.spectral_shape) + 1j * rng.standard_normal(k.spectral_shape)) for _ in range(3)]
    F = _host(k.nonlinear_flux(*_dev(k, A), t=450.0, t0=0.0))
    err = rel_err_joint(F, o.nonlinear_flux(*A, t=450.0, t0=0.0))
    fields = k.transform_wave_vortex_to_uvweta(*_dev(k, A), t=450.0).cpu().numpy()
    ref = o.inverse_uvweta(*A, t=450.0)
    ferr = float(np.abs(fields - ref).max() / np.abs(ref).max())
    print(f"[gpu] {Nx}x{Ny}x{Nz} hydro={hydro} fp64 flux rel err {err:.2e}, inverse transform {ferr:.2e}", flush=True)
    assert err <= FP64_TOL and ferr <= 1e-13


def test_flux_mixed_non_power_of_two_matches_fp64(torch_gpu):
    import wvamd
    cfg = wvamd.make_config(96, 80, 17, None, Lx=60e3, Ly=50e3)
    k64, k32 = _kernel(cfg), _kernel(cfg, "mixed-split16")
    rng = np.random.default_rng(96)
    A = [1e-3 * (rng.standard_normal(k64.spectral_shape) + 1j * rng.standard_normal(k64.spectral_shape))
         for _ in range(3)]
    F64 = _host(k64.nonlinear_flux(*_dev(k64, A), t=450.0, t0=0.0))
    F32 = _host(k32.nonlinear_flux(*_dev(k32, A), t=450.0, t0=0.0))
    err = rel_err_joint(F32, F64)
    print(f"[gpu] 96x80x17 mixed-split16 flux rel err {err:.2e}", flush=True)
    assert err <= MIXED_TOL
