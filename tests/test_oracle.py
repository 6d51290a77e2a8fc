"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the checker.

* golden vectors written by the reference kernel itself (tests/golden/make_golden.py
  runs CompiledKernel + its direct-DFT engine from oracle/_ref);
* mathematical invariants: zero state -> exactly zero flux
  (TestWVKernelContract.cpp:395-399), spectral energy == spatially integrated
  energy, energy conservation of the nonlinear flux
  (UnitTests/TestNonlinearFlux.m:7-18), forward(inverse(A)) == A on retained
  modes, the RK4 closed form for y' = lambda y.
"""
import os

import numpy as np
import pytest
from conftest import GOLDEN, config_from_golden, golden_files, load_golden, rel_err, rel_err_joint

from oracle import Oracle, Reference, make_config

TABLES = ["omega", "fWaveScale", "UApField", "VApField", "WApField", "NApField", "UA0Field", "VA0Field",
          "NA0Field", "A0FromVorticity", "A0FromBuoyancy", "ApmDProjection", "ApmNProjection", "ApmDScaled",
          "z", "h0", "verticalWavenumber", "Fg", "Gg", "inertialScale", "gWaveScale", "apmWProjectionPrefactor"]

GOLDEN_TINY = [p for p in golden_files("golden_c") if "flow" not in os.path.basename(p)]


def _ids(paths):
    return [os.path.basename(p)[:-4] for p in paths]


@pytest.mark.parametrize("path", GOLDEN_TINY, ids=_ids(GOLDEN_TINY))
def test_oracle_descriptor_matches_reference_bitwise(path):
    g = load_golden(path)
    o = Oracle(config_from_golden(g))
    md = o.modes()
    assert np.array_equal(md["kMode"], g["kMode"]) and np.array_equal(md["lMode"], g["lMode"])
    assert np.array_equal(md["row"], g["row"]) and np.array_equal(md["conj"], g["conj"])
    assert md["kMode"][0] == 0 and md["lMode"][0] == 0  # zero mode first (TestWVKernelContract.cpp:268)
    for name in TABLES:
        assert np.array_equal(o.table(name), g["T_" + name]), name
    assert o.coriolis() == float(g["coriolis"])


@pytest.mark.parametrize("path", golden_files("descriptor_"), ids=_ids(golden_files("descriptor_")))
def test_oracle_descriptor_large_grids(path):
    g = load_golden(path)
    o = Oracle(config_from_golden(g))
    md = o.modes()
    assert o.Nkl == g["kMode"].size
    assert np.array_equal(md["kMode"], g["kMode"]) and np.array_equal(md["lMode"], g["lMode"])
    assert np.array_equal(md["row"], g["row"]) and np.array_equal(md["conj"], g["conj"])
    for name in TABLES:
        t = o.table(name)
        if "T_" + name in g:
            assert np.array_equal(t, g["T_" + name]), name
        else:
            assert np.array_equal(t[g["S_" + name + "_idx"]], g["S_" + name]), name
            cs = g["C_" + name]
            mine = np.array([np.sum(t), np.sum(t * t), np.sum(t * np.cos(0.001 * np.arange(t.size)))])
            assert np.allclose(mine, cs, rtol=1e-13, atol=0), name


def test_descriptor_counts_match_survey():
    # Nkl measured with the reference descriptor dump (survey table of configurations)
    assert Oracle(make_config(64, 64, 33, 21, Lx=50e3)).Nkl == 717
    assert Oracle(make_config(256, 256, 129, 85, Lx=200e3)).Nkl == 11439


@pytest.mark.parametrize("path", GOLDEN_TINY, ids=_ids(GOLDEN_TINY))
def test_oracle_flux_matches_reference(path):
    g = load_golden(path)
    o = Oracle(config_from_golden(g))
    F = o.nonlinear_flux(g["Ap"], g["Am"], g["A0"], t=float(g["t"]), t0=float(g["t0"]))
    assert rel_err_joint(F, (g["Fp"], g["Fm"], g["F0"])) <= 1e-12
    assert float(g["zero_flux_maxabs"]) == 0.0
    Z = np.zeros_like(g["Ap"])
    assert all(np.all(x == 0) for x in o.nonlinear_flux(Z, Z, Z, t=0.5))


@pytest.mark.parametrize("path", [p for p in GOLDEN_TINY if "inv_uvweta" in np.load(p).files],
                         ids=_ids([p for p in GOLDEN_TINY if "inv_uvweta" in np.load(p).files]))
def test_oracle_transforms_match_reference(path):
    g = load_golden(path)
    o = Oracle(config_from_golden(g))
    t, t0 = float(g["t"]), float(g["t0"])
    assert rel_err(o.inverse_uvweta(g["Ap"], g["Am"], g["A0"], t=t, t0=t0), g["inv_uvweta"]) <= 1e-13
    assert rel_err(o.inverse_uvw(g["Ap"], g["Am"], g["A0"], t=t, t0=t0), g["inv_uvw"]) <= 1e-13
    fwd = o.forward(g["fwd_fields"], t=t, t0=t0)
    for mine, ref in zip(fwd, (g["fwd_Ap"], g["fwd_Am"], g["fwd_A0"])):
        assert rel_err(mine, ref) <= 1e-13
    assert rel_err(o.fg_all(True, g["Ap"], g["A0"]), g["fAll"]) <= 1e-13
    assert rel_err(o.fg_all(False, g["Ap"], g["A0"]), g["gAll"]) <= 1e-13


def test_oracle_flux_matches_reference_64x64x33_flow():
    path = os.path.join(GOLDEN, "golden_c64x64x33_flow.npz")
    if not os.path.exists(path):
        pytest.skip("64x64x33 reference flux fixture not generated")
    g = load_golden(path)
    o = Oracle(config_from_golden(g))
    F = o.nonlinear_flux(g["Ap"], g["Am"], g["A0"], t=float(g["t"]), t0=float(g["t0"]))
    assert rel_err_joint(F, (g["Fp"], g["Fm"], g["F0"])) <= 1e-12


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("hydro", [False, True])
def test_oracle_matches_live_reference_random(hydro):
    cfg = make_config(10, 12, 9, 6, Lx=20e3, Ly=17e3, isHydrostatic=hydro, latitude=-40.0)
    o, r = Oracle(cfg), Reference(cfg)
    rng = np.random.default_rng(11)
    A = [rng.standard_normal((o.Nkl, o.Nj)) + 1j * rng.standard_normal((o.Nkl, o.Nj)) for _ in range(3)]
    assert rel_err_joint(o.nonlinear_flux(*A, t=100.0, t0=-3.0), r.nonlinear_flux(*A, t=100.0, t0=-3.0)) <= 1e-12


def _random_flow(o, seed=3, jmax=None):
    from wvamd import flows
    md = o.modes()
    Apm = o.table("ApmTE").reshape(o.Nkl, o.Nj)
    A0t = o.table("A0TE").reshape(o.Nkl, o.Nj)
    return flows.random_flow_shape(md["Kh"], o.Nj, Apm, A0t, seed=seed, jmax=o.Nj if jmax is None else jmax)


@pytest.mark.parametrize("hydro", [False, True])
def test_oracle_energy_spectral_equals_spatial(hydro):
    o = Oracle(make_config(32, 32, 17, 11, Lx=50e3, isHydrostatic=hydro))
    Ap, Am, A0 = _random_flow(o)
    es = o.energy_spatial(o.inverse_uvweta(Ap, Am, A0))
    assert abs(o.energy(Ap, Am, A0) - es) <= 1e-13 * es


@pytest.mark.parametrize("hydro", [False, True])
def test_oracle_flux_conserves_energy(hydro):
    # dE/dt = 2 sum TE * Re(conj(A) F) vanishes for the antialiased nonlinear
    # flux up to round-off (UnitTests/TestNonlinearFlux.m:7-18 bound 100 eps).
    o = Oracle(make_config(32, 32, 17, 11, Lx=50e3, isHydrostatic=hydro))
    Ap, Am, A0 = _random_flow(o)
    Fp, Fm, F0 = o.nonlinear_flux(Ap, Am, A0)
    Apm = o.table("ApmTE").reshape(o.Nkl, o.Nj)
    A0t = o.table("A0TE").reshape(o.Nkl, o.Nj)
    terms = np.concatenate([(Apm * np.real(np.conj(Ap) * Fp)).ravel(), (Apm * np.real(np.conj(Am) * Fm)).ravel(),
                            (A0t * np.real(np.conj(A0) * F0)).ravel()])
    assert abs(terms.sum()) <= 100 * np.finfo(float).eps * np.abs(terms).sum()


def test_oracle_forward_inverts_inverse():
    cfg = make_config(16, 16, 9, 5, Lx=50e3)
    o = Oracle(cfg)
    Ap, Am, A0 = _random_flow(o)
    back = o.forward(o.inverse_uvweta(Ap, Am, A0, t=10.0), t=10.0)
    for x, y in zip(back, (Ap, Am, A0)):
        assert rel_err(x, y) <= 1e-12


@pytest.mark.parametrize("lam", [-0.3 + 2.0j, 0.1 - 1.0j, 1.5j])
def test_oracle_rk4_linear_closed_form(lam):
    # classic RK4 on y' = lam y: y_{n+1} = y_n (1 + z + z^2/2 + z^3/6 + z^4/24), z = lam h
    h, n, y0 = 0.05, 40, 0.7 - 0.2j
    z = lam * h
    exact = y0 * (1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24) ** n
    got = Oracle.rk4_linear_test(lam, y0, h, n)
    assert abs(got - exact) <= 1e-13 * abs(exact)


def test_oracle_rk4_matches_stage_arithmetic():
    # The flux-driven RK4 equals the explicit stage formula of WVFixedStepRK4::step.
    cfg = make_config(8, 8, 5, 3, Lx=15e3)
    o = Oracle(cfg)
    rng = np.random.default_rng(5)
    y = [1e-3 * (rng.standard_normal((o.Nkl, o.Nj)) + 1j * rng.standard_normal((o.Nkl, o.Nj))) for _ in range(3)]
    h, t = 30.0, 4.0
    k1 = o.nonlinear_flux(*y, t=t)
    s = [a + 0.5 * h * b for a, b in zip(y, k1)]
    k2 = o.nonlinear_flux(*s, t=t + 0.5 * h)
    s = [a + 0.5 * h * b for a, b in zip(y, k2)]
    k3 = o.nonlinear_flux(*s, t=t + 0.5 * h)
    s = [a + h * b for a, b in zip(y, k3)]
    k4 = o.nonlinear_flux(*s, t=t + h)
    expect = [a + (h / 6.0) * (b + 2 * c + 2 * d + e) for a, b, c, d, e in zip(y, k1, k2, k3, k4)]
    got = o.rk4(*y, t=t, dt=h, nsteps=1)
    for a, b in zip(got[:3], expect):
        assert rel_err(a, b) <= 1e-14
    assert got[3] == t + h


# --- advectFGridScalar (numpy restatement in oracle/scalar_advection.py) ------
SCALAR_GOLDEN = golden_files("scalar_")


@pytest.mark.parametrize("path", SCALAR_GOLDEN, ids=_ids(SCALAR_GOLDEN))
def test_scalar_advection_oracle_matches_reference(path):
    """The restatement against the reference kernel's own advectFGridScalar
    (tests/golden/scalar_*.npz: even and odd grids, with and without the 2/3
    rule), both with and without the scalar antialias."""
    from scalar_advection import advect_f_grid_scalar
    g = load_golden(path)
    cfg = config_from_golden(g)
    rows = Oracle(cfg).modes()["row"]
    for aa, key in ((False, "rhs"), (True, "rhs_antialiased")):
        got = advect_f_grid_scalar(cfg, g["scalar"], g["uvw"], aa, rows)
        assert rel_err(got, g[key]) <= 1e-13, key


def test_scalar_advection_oracle_closed_forms():
    """phi = cos(2 pi x / Lx) cos(pi z / Lz) advected by u = 1 gives
    -d/dx phi = (2 pi / Lx) sin(2 pi x / Lx) cos(pi z / Lz); by w = 1,
    -d/dz phi = (pi / Lz) cos(2 pi x / Lx) sin(pi z / Lz) (both band-limited,
    so the antialias leaves them unchanged)."""
    from scalar_advection import advect_f_grid_scalar
    cfg = make_config(16, 8, 9, Lx=50e3, Ly=40e3)
    Nx, Ny, Nz, Lx, Lz = 16, 8, 9, cfg["Lx"], cfg["Lz"]
    x = np.arange(Nx) * Lx / Nx
    z = np.arange(Nz) * Lz / (Nz - 1)
    X, Z = np.meshgrid(x, z)  # [Nz, Nx]
    phi = np.broadcast_to((np.cos(2 * np.pi * X / Lx) * np.cos(np.pi * Z / Lz))[:, None, :], (Nz, Ny, Nx))
    rows = Oracle(cfg).modes()["row"]
    ones, zeros = np.ones((Nz, Ny, Nx)), np.zeros((Nz, Ny, Nx))
    want_u = np.broadcast_to(((2 * np.pi / Lx) * np.sin(2 * np.pi * X / Lx) * np.cos(np.pi * Z / Lz))[:, None, :],
                             (Nz, Ny, Nx))
    want_w = np.broadcast_to(((np.pi / Lz) * np.cos(2 * np.pi * X / Lx) * np.sin(np.pi * Z / Lz))[:, None, :],
                             (Nz, Ny, Nx))
    for aa in (False, True):
        got_u = advect_f_grid_scalar(cfg, phi, np.stack([ones, zeros, zeros]), aa, rows)
        got_w = advect_f_grid_scalar(cfg, phi, np.stack([zeros, zeros, ones]), aa, rows)
        assert rel_err(got_u, want_u) <= 1e-13
        assert rel_err(got_w, want_w) <= 1e-12


CONTRACT = golden_files("contract_")


@pytest.mark.parametrize("path", CONTRACT, ids=_ids(CONTRACT))
def test_oracle_state_derivatives_and_flux_using_fields_match_reference(path):
    """transformStateFieldDerivatives (...Kernel.cpp:719-742) for u, v, w, eta and
    nonlinearFluxUsingAdvectionFields (:1160-1171) with seeded advecting fields,
    against the reference kernel's own outputs."""
    g = load_golden(path)
    o = Oracle(config_from_golden(g))
    t, t0 = float(g["t"]), float(g["t0"])
    for f in range(4):
        d = o.state_field_derivatives(g["Ap"], g["Am"], g["A0"], f, t=t, t0=t0)
        assert rel_err(d, g["derivatives"][f]) <= 1e-13, f
    F = o.nonlinear_flux_using_advection_fields(g["Ap"], g["Am"], g["A0"], g["uvw"], t=t, t0=t0)
    assert rel_err_joint(F, (g["Fp"], g["Fm"], g["F0"])) <= 1e-13
    # the given fields are what the products use: the state's own fields give
    # a different flux
    own = o.nonlinear_flux(g["Ap"], g["Am"], g["A0"], t=t, t0=t0)
    assert rel_err_joint(own, (g["Fp"], g["Fm"], g["F0"])) > 1e-3


@pytest.mark.parametrize("hydrostatic", [True, False], ids=["hydrostatic", "nonhydrostatic"])
def test_oracle_triad_energy_flux_matches_closed_form(hydrostatic):
    """The reference's closed-form three-wave triad (UnitTests/TestNonlinearFlux.m:68-262):
    energy flux into each wave of the triad; the reference's bar is RelTol
    1e-7, observed 5e-14, checked at 1e-12."""
    import triad_kat
    cfg = triad_kat.config(make_config, hydrostatic)
    o = Oracle(cfg)
    Ap, idx, E = triad_kat.setup(cfg, o.modes(), o.coriolis())
    Z = np.zeros_like(Ap)
    Fp, _, _ = o.nonlinear_flux(Ap, Z, Z, t=0.0, t0=0.0)
    ApmTE = o.table("ApmTE").reshape(o.Nkl, o.Nj)
    Ep = triad_kat.energy_flux(Fp, Ap, ApmTE, idx)
    assert np.all(np.abs(Ep - E) <= 1e-12 * np.abs(E)), (Ep, E)
    # a mistake in the flux fails it: the flux of a wave with the wrong phase
    Fp2, _, _ = o.nonlinear_flux(Ap * np.exp(0.3j), Z, Z, t=0.0, t0=0.0)
    assert not np.all(np.abs(triad_kat.energy_flux(Fp2, Ap, ApmTE, idx) - E) <= 1e-7 * np.abs(E))
