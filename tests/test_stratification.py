"""CPU checks of the arbitrary-stratification host eigenproblem
(wvamd.stratification) and of the numpy oracle of the hydrostatic flux
(oracle/hydrostatic.py).

Pins:
  constant N      modes, depths and quadrature grid equal the closed forms of
                  the constant-N transform (cos/sin modes with the reference's
                  normalization, WVKernelTypes.cpp:203-212); the oracle's flux
                  with those operators equals the reference kernel's
                  hydrostatic flux (golden vectors) and the C++ oracle's
  exponential N^2 N0^2 exp(2z/b): eigenvalues from the closed-form Bessel
                  determinant J0(sD) Y0(s0) - J0(s0) Y0(sD) = 0, s = (N0 b/c) e^{z/b},
                  and mode shapes G ~ J0(s) Y0(s0) - Y0(s) J0(s0)
  operators       PF PFinv = I, QG QGinv = I, the quadrature weights integrate
                  F_j to Lz delta_j0 (verticalProjectionOperatorsWithRigidLid.m:35-37);
                  the dense diffZF/diffZG operators applied to synthesized
                  fields equal -(N^2/g) G and F/h synthesis (the form the GPU uses)
"""
import os

import numpy as np
import pytest
from conftest import GOLDEN, config_from_golden, load_golden, rel_err_joint
from scipy.optimize import brentq
from scipy.special import j0, y0

from wvamd import stratification as S

G_ = 9.81


def _constant(cfg):
    N2, dN2 = S.constant(cfg["N0"])
    return S.Stratification(N2, Lz=cfg["Lz"], g=cfg["g"], dN2=dN2)


def test_constant_n_modes_match_closed_form():
    N0, Lz, Nz, Nj = 5.2e-3, 1300.0, 17, 11
    st = S.Stratification(*S.constant(N0)[:1], Lz=Lz, g=G_, dN2=S.constant(N0)[1])
    m = np.arange(1, Nz) * np.pi / Lz
    h = N0 ** 2 / (G_ * m ** 2)
    assert np.max(np.abs(st.h[:Nz - 1] - h) / h) < 1e-10
    ops = S.vertical_operators(st, Nz, Nj)
    z = ops["z"]
    assert np.max(np.abs(z - np.linspace(-Lz, 0.0, Nz))) < 1e-9 * Lz
    j = np.arange(1, Nj)
    sign = (-1.0) ** j
    Gg = sign * np.sqrt(2.0 * G_ / (Lz * N0 ** 2))
    Fg = sign * h[:Nj - 1] * m[:Nj - 1] * np.sqrt(2.0 * G_ / (Lz * N0 ** 2))
    F = ops["PFinv"][:, 1:] * ops["P"][None, 1:]
    G = ops["QGinv"][:, 1:] * ops["Q"][None, 1:]
    Fe = Fg[None, :] * np.cos(np.outer(z + Lz, m[:Nj - 1]))
    Ge = Gg[None, :] * np.sin(np.outer(z + Lz, m[:Nj - 1]))
    assert np.max(np.abs(F - Fe)) < 1e-9 * np.max(np.abs(Fe))
    assert np.max(np.abs(G - Ge)) < 1e-9 * np.max(np.abs(Ge))
    assert np.allclose(ops["PFinv"][:, 0] * ops["P"][0], 1.0)


def test_exponential_stratification_matches_bessel_solution():
    N0, b, Lz = 5.2e-3, 1300.0, 4000.0
    N2, dN2 = S.exponential(N0, b)
    st = S.Stratification(N2, Lz=Lz, g=G_, dN2=dN2)

    def det(c):
        s0 = N0 * b / c
        sD = s0 * np.exp(-Lz / b)
        return j0(sD) * y0(s0) - j0(s0) * y0(sD)

    zz = np.linspace(-Lz, 0.0, 41)
    for j in (1, 2, 3, 8, 20, 40, 64):
        c = np.sqrt(G_ * st.h[j - 1])
        root = brentq(det, 0.999 * c, 1.001 * c, xtol=1e-18)
        assert abs(root ** 2 / G_ - st.h[j - 1]) <= 1e-10 * st.h[j - 1], j
        if j <= 8:
            s0 = N0 * b / root
            s = s0 * np.exp(zz / b)
            Gb = j0(s) * y0(s0) - y0(s) * j0(s0)
            # normalization (1/g) int N^2 G^2 dz = 1 and F(0) > 0, by quadrature
            zf = np.linspace(-Lz, 0.0, 40001)
            sf = s0 * np.exp(zf / b)
            Gf = j0(sf) * y0(s0) - y0(sf) * j0(s0)
            nrm = np.sqrt(np.trapezoid(N2(zf) * Gf ** 2, zf) / G_)
            dG0 = np.gradient(Gf, zf)[-1]
            Gb *= np.sign(dG0) / nrm
            mine = st.G(zz, j)[:, j - 1]
            assert np.max(np.abs(mine - Gb)) <= 1e-6 * np.max(np.abs(Gb)), j


@pytest.mark.parametrize("kind", ["constant", "exponential"])
def test_projection_operator_invariants(kind):
    Lz, Nz, Nj = 4000.0, 33, 24
    N2, dN2 = S.constant() if kind == "constant" else S.exponential()
    st = S.Stratification(N2, Lz=Lz, g=G_, dN2=dN2)
    ops = S.vertical_operators(st, Nz, Nj)
    assert np.max(np.abs(ops["PF"] @ ops["PFinv"] - np.eye(Nj))) < 1e-12
    assert np.max(np.abs(ops["QG"][1:] @ ops["QGinv"][:, 1:] - np.eye(Nj - 1))) < 1e-12
    F = ops["PFinv"] * ops["P"][None, :]
    integrals = ops["w"] @ F
    assert abs(integrals[0] - Lz) < 1e-10 * Lz
    assert np.max(np.abs(integrals[1:])) < 1e-9 * Lz
    # dense diffZF / diffZG (MATLAB) applied to synthesized fields == the
    # synthesis matrices the GPU uses for d/dz
    from hydrostatic import HydrostaticOracle
    from oracle import Oracle, make_config
    cfg = make_config(8, 8, Nz, Nj, Lz=Lz, isHydrostatic=True)
    ho = HydrostaticOracle(cfg, ops, Oracle(cfg).modes())
    G = ops["QGinv"] * ops["Q"][None, :]
    c = np.random.default_rng(1).standard_normal(Nj)
    lhs = ho.DzF @ (F @ c)
    rhs = -(ops["N2"] / G_) * (G @ c)
    assert np.max(np.abs(lhs - rhs)) <= 1e-11 * np.max(np.abs(rhs))
    lhs = ho.DzG @ (G @ c)
    Fh = F[:, 1:] / ops["h"][None, 1:]
    rhs = Fh @ c[1:]
    assert np.max(np.abs(lhs - rhs)) <= 1e-11 * np.max(np.abs(rhs))


@pytest.mark.parametrize("name", ["golden_c16x16x9_h", "golden_c6x5x7_h", "golden_c15x13x9_h", "golden_c7x5x7_h_noaa"])
def test_hydrostatic_oracle_reduces_to_reference_constant_n(name):
    path = os.path.join(GOLDEN, name + ".npz")
    g = load_golden(path)
    cfg = config_from_golden(g)
    from hydrostatic import HydrostaticOracle
    from oracle import Oracle
    st = _constant(cfg)
    ops = S.vertical_operators(st, cfg["Nz"], cfg["Nj"])
    ho = HydrostaticOracle(cfg, ops, Oracle(cfg).modes())
    F = ho.nonlinear_flux(g["Ap"], g["Am"], g["A0"], t=float(g["t"]), t0=float(g["t0"]))
    err = rel_err_joint(F, (g["Fp"], g["Fm"], g["F0"]))
    assert err <= 1e-9, err


def test_hydrostatic_oracle_forward_inverts_inverse():
    from hydrostatic import HydrostaticOracle
    from oracle import Oracle, make_config
    from wvamd import flows
    cfg = make_config(16, 16, 17, 12, Lx=100e3, Lz=4000.0, isHydrostatic=True)
    st = S.Stratification(*S.exponential()[:1], Lz=cfg["Lz"], g=cfg["g"], dN2=S.exponential()[1])
    ops = S.vertical_operators(st, cfg["Nz"], cfg["Nj"])
    md = Oracle(cfg).modes()
    ho = HydrostaticOracle(cfg, ops, md)
    A = flows.random_flow_shape(md["Kh"], cfg["Nj"], ho.ApmTE.ravel(), ho.A0TE.ravel(), seed=9, jmax=cfg["Nj"])
    f = ho.fields(*A, t=250.0)
    B = ho.forward(f[0], f[1], f[3], t=250.0)
    assert rel_err_joint(B, A) <= 1e-12
