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
                  and mode shapes G ~ J0(s) Y0(s0) - Y0(s) J0(s0); the
                  hydrostatic oracle's advection products (incl. the eta dLnN2
                  term, dLnN2 = 2/b) equal the closed-form products of a
                  Bessel-mode flow
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


def _bessel_mode(N0, b, Lz, h_guess):
    """Closed-form rigid-lid mode of N^2 = N0^2 exp(2z/b): the eigen speed c
    (root of J0(sD) Y0(s0) - J0(s0) Y0(sD)), and callables G(z), G'(z), G''(z)
    of G = J0(s) Y0(s0) - Y0(s) J0(s0), s = (N0 b / c) e^{z/b} (ds/dz = s/b,
    J0' = -J1, J1' = J0 - J1/s)."""
    from scipy.special import j1, y1

    def det(c):
        s0 = N0 * b / c
        sD = s0 * np.exp(-Lz / b)
        return j0(sD) * y0(s0) - j0(s0) * y0(sD)

    c0 = np.sqrt(G_ * h_guess)
    c = brentq(det, 0.999 * c0, 1.001 * c0, xtol=1e-18)
    s0 = N0 * b / c

    def G(z):
        s = s0 * np.exp(z / b)
        return j0(s) * y0(s0) - y0(s) * j0(s0)

    def dG(z):
        s = s0 * np.exp(z / b)
        return -(s / b) * (j1(s) * y0(s0) - y1(s) * j0(s0))

    def d2G(z):
        s = s0 * np.exp(z / b)
        return -(s / b) ** 2 * G(z)

    return c * c / G_, G, dG, d2G


def test_exponential_n2_advection_products_match_bessel_closed_form():
    """Pins the arbitrary-N^2 flux of oracle/hydrostatic.py before projection:
    a flow of j = 1 and j = 2 modes of N^2 = N0^2 exp(2z/b) at three
    wavevectors, whose u, v, w, eta and their x, y, z derivatives are written
    in closed form (Bessel functions), gives
        Fu = -(u u_x + v u_y + w u_z),  Fv = -(u v_x + v v_y + w v_z),
        Feta = -(u eta_x + v eta_y + w (eta_z + eta dLnN2)),  dLnN2 = 2/b
    (Forcing/WVNonlinearAdvection.m:103-107), with the vertical structure
    and its derivatives from the Bessel solution, not from the operators.
    Observed agreement is 1e-11 (checked at 1e-9); a sign flip of the
    eta dLnN2 term (a mistake planted in the oracle's input) misses by 1.8."""
    from hydrostatic import HydrostaticOracle
    from oracle import Oracle, make_config
    N0, b, Lz = 5.2e-3, 1300.0, 4000.0
    Nx, Ny, Nz, Nj = 16, 16, 65, 12
    cfg = make_config(Nx, Ny, Nz, Nj, Lx=100e3, Ly=80e3, Lz=Lz, isHydrostatic=True)
    N2f, dN2f = S.exponential(N0, b)
    st = S.Stratification(N2f, Lz=Lz, g=G_, dN2=dN2f)
    ops = S.vertical_operators(st, Nz, Nj)
    z = ops["z"]
    # d ln N^2 / dz of the exponential is exactly 2/b
    assert np.max(np.abs(ops["dLnN2"] - 2.0 / b)) <= 1e-13 * (2.0 / b)
    md = Oracle(cfg).modes()
    ho = HydrostaticOracle(cfg, ops, md)

    # closed-form F_j = h_j G_j', G_j, with the operators' normalization and sign
    Fcl, Gcl, dFcl, dGcl = {}, {}, {}, {}
    for j in (1, 2):
        h, G, dG, d2G = _bessel_mode(N0, b, Lz, st.h[j - 1])
        assert abs(h - ops["h"][j]) <= 1e-10 * h
        Gop = ho.G[:, j]
        alpha = (Gop @ G(z)) / (G(z) @ G(z))
        Gcl[j], dGcl[j] = alpha * G(z), alpha * dG(z)
        Fcl[j], dFcl[j] = h * dGcl[j], h * alpha * d2G(z)
        assert np.max(np.abs(Gop - Gcl[j])) <= 1e-6 * np.max(np.abs(Gcl[j])), j
        assert np.max(np.abs(ho.F[:, j] - Fcl[j])) <= 1e-6 * np.max(np.abs(Fcl[j])), j
        # d/dz F = -(N^2/g) G: the Bessel equation, so the closed forms agree
        assert np.max(np.abs(dFcl[j] + N2f(z) / G_ * Gcl[j])) <= 1e-12 * np.max(np.abs(dFcl[j]))

    # a state of two waves, a geostrophic mode and a mean density anomaly
    # (K = 0, horizontally uniform eta: it enters Feta only through
    # w (eta_z + eta dLnN2)), each scaled to a physical amplitude
    km, lm = md["kMode"], md["lMode"]
    pick = [int(np.flatnonzero((km == a) & (lm == c))[0]) for a, c in ((1, 0), (1, 2), (0, 1), (0, 0))]
    parts = []
    for arr, im, j, phase, target in ((0, pick[0], 1, 0.3, ("u", 0.05)), (1, pick[1], 2, 1.1, ("u", 0.04)),
                                      (2, pick[2], 1, -0.7, ("u", 0.02)), (2, pick[3], 2, 0.0, ("eta", 30.0))):
        A = [np.zeros((ho.Nkl, Nj), complex) for _ in range(3)]
        A[arr][im, j] = np.exp(1j * phase)
        f = ho.fields(*A, t=0.0)
        A[arr] *= target[1] / np.max(np.abs(f[0 if target[0] == "u" else 3]))
        parts.append(A)
    Ap, Am, A0 = (sum(p[a] for p in parts) for a in range(3))
    t = 500.0
    u, v, w, n = ho._spectra(Ap, Am, A0, t)

    x = np.arange(Nx) * cfg["Lx"] / Nx
    y = np.arange(Ny) * cfg["Ly"] / Ny
    X, Y = np.meshgrid(x, y)  # [Ny, Nx]

    def synth(spec, prof, dk=0.0, dl=0.0):
        """sum over the picked modes of spec e^{i(kx+ly)} + c.c. (the real part
        alone at K = 0) times the vertical profile."""
        out = np.zeros((Nz, Ny, Nx))
        for im in pick:
            for j in (1, 2):
                if spec[im, j] == 0:
                    continue
                k, l = md["k"][im], md["l"][im]
                fac = (1j * k) ** dk * (1j * l) ** dl
                wgt = 1.0 if (km[im] == 0 and lm[im] == 0) else 2.0
                plane = wgt * np.real(fac * spec[im, j] * np.exp(1j * (k * X + l * Y)))
                out += prof[j][:, None, None] * plane[None]
        return out

    U, Ux, Uy, Uz = synth(u, Fcl), synth(u, Fcl, 1), synth(u, Fcl, 0, 1), synth(u, dFcl)
    V, Vx, Vy, Vz = synth(v, Fcl), synth(v, Fcl, 1), synth(v, Fcl, 0, 1), synth(v, dFcl)
    W = synth(w, Gcl)
    E, Ex, Ey, Ez = synth(n, Gcl), synth(n, Gcl, 1), synth(n, Gcl, 0, 1), synth(n, dGcl)
    Fu = -(U * Ux + V * Uy + W * Uz)
    Fv = -(U * Vx + V * Vy + W * Vz)
    Fe = -(U * Ex + V * Ey + W * (Ez + E * (2.0 / b)))
    # the eta dLnN2 term is an O(1) part of Feta for this flow
    assert np.max(np.abs(W * E * (2.0 / b))) > 0.2 * np.max(np.abs(Fe))

    mine = ho.spatial_products(Ap, Am, A0, t)
    for a, ref, name in zip(mine, (Fu, Fv, Fe), ("Fu", "Fv", "Feta")):
        err = np.max(np.abs(a - ref)) / np.max(np.abs(ref))
        assert err <= 1e-9, (name, err)
    wrong = HydrostaticOracle(cfg, dict(ops, dLnN2=-ops["dLnN2"]), md).spatial_products(Ap, Am, A0, t)
    assert np.max(np.abs(wrong[2] - Fe)) / np.max(np.abs(Fe)) > 0.1
