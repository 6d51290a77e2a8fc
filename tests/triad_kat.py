"""Closed-form three-wave triad: the known-answer test of the reference's
UnitTests/TestNonlinearFlux.m:68-262 (testNonlinearWaveTriad), restated.

Three waves of the constant-stratification transform at Lx = Ly = 2 Lz =
4 km, 16 x 16 x 9 points, no antialiasing (:246-248):
  A: kMode 1, lMode 0, j 1;  B: kMode 1, lMode 1, j 1;  C: kMode 0, lMode 1, j 2,
each with a real, positive Ap (initWithWaveModes/addWaveModes with phi = 0 at
t = t0, FlowComponents/WVInternalGravityWaveMethods.m:185-220).  With
L = Lx / 2 pi, T = f / Omega of each wave, N2 = N0^2 / f^2 and the true
amplitudes A = 2 Ap h Aw, B = 2 Ap h Aw / sqrt(2), C = 2 Ap h Aw (:91-112),
the energy flux into each wave,
    Ep = 2 Apm_TE_factor Re(Fp conj(Ap))   (@WVTransform/energyFluxFromNonlinearFlux.m:37),
is, for the hydrostatic transform,
    E_A = -pi ABC/(16 L^3) (7Ta + Tb - 8Tc + (5 N2 - 2) Ta Tb Tc)
    E_B = -pi ABC/(16 L^3) (-3Ta + 3Tb + 6Tc - 6 N2 Ta Tb Tc)
    E_C = -pi ABC/(16 L^3) (-4Ta - 4Tb + 2Tc + (N2 + 2) Ta Tb Tc)
and for the nonhydrostatic one
    E_A = -pi ABC/(16 L^3) (7Ta - 4Tc + (5 N2 - 2) Ta Tb Tc)
    E_B = -pi ABC/(16 L^3) (-5Ta + 3Tb + 2Tc - 6 N2 Ta Tb Tc)
    E_C = -pi ABC/(16 L^3) (-2Ta - 3Tb + 2Tc + (N2 + 2) Ta Tb Tc)
(:234-246).  Equivalent depths and frequencies are the closed forms
h = N0^2 / (g m^2) (hydrostatic), (N0^2 - f^2) / (g (m^2 + K^2))
(nonhydrostatic), m = j pi / Lz, Omega = sqrt(g h K^2 + f^2).  The
reference checks these at RelTol 1e-7 (:83-84).
"""
import numpy as np

N0 = 5.2e-3
LXYZ = (4e3, 4e3, 2e3)
WAVES = ((1, 0, 1), (1, 1, 1), (0, 1, 2))  # (kMode, lMode, j) of A, B, C


def config(make_config, hydrostatic):
    Lx, Ly, Lz = LXYZ
    return make_config(16, 16, 9, 8, Lx=Lx, Ly=Ly, Lz=Lz, N0=N0, isHydrostatic=hydrostatic, shouldAntialias=False)


def setup(cfg, modes, f):
    """(Ap, indices, expected Ep at the three waves)."""
    g, Lx, Lz = cfg["g"], cfg["Lx"], cfg["Lz"]
    hydro = bool(cfg["isHydrostatic"])
    Nkl, Nj = len(modes["k"]), int(cfg["Nj"])
    L = Lx / (2.0 * np.pi)
    Aw = np.sqrt(2.0 * g / Lz) / N0 if hydro else np.sqrt(2.0 * g / ((N0 * N0 - f * f) * Lz))
    Ap = np.zeros((Nkl, Nj), complex)
    idx, T, amp = [], [], []
    for n, (km, lm, j) in enumerate(WAVES):
        im = int(np.flatnonzero((modes["kMode"] == km) & (modes["lMode"] == lm))[0])
        K = modes["Kh"][im]
        m = j * np.pi / Lz
        h = N0 * N0 / (g * m * m) if hydro else (N0 * N0 - f * f) / (g * (m * m + K * K))
        Om = np.sqrt(g * h * K * K + f * f)
        # true amplitude 0.05 L (u ~ 5 cm/s), Ap from A = 2 Ap h Aw (/ sqrt 2 for B)
        true = 0.05 * L
        a = true / (2.0 * h * Aw) * (np.sqrt(2.0) if n == 1 else 1.0)
        Ap[im, j] = a
        idx.append((im, j))
        T.append(f / Om)
        amp.append(true)
    A, B, C = amp
    Ta, Tb, Tc = T
    N2 = N0 * N0 / (f * f)
    pre = -np.pi * (A * B * C / 16.0 / L ** 3)
    if hydro:
        E = (pre * (7 * Ta + Tb - 8 * Tc + (5 * N2 - 2) * Ta * Tb * Tc),
             pre * (-3 * Ta + 3 * Tb + 6 * Tc - 6 * N2 * Ta * Tb * Tc),
             pre * (-4 * Ta - 4 * Tb + 2 * Tc + (N2 + 2) * Ta * Tb * Tc))
    else:
        E = (pre * (7 * Ta - 4 * Tc + (5 * N2 - 2) * Ta * Tb * Tc),
             pre * (-5 * Ta + 3 * Tb + 2 * Tc - 6 * N2 * Ta * Tb * Tc),
             pre * (-2 * Ta - 3 * Tb + 2 * Tc + (N2 + 2) * Ta * Tb * Tc))
    return Ap, idx, np.array(E)


def energy_flux(Fp, Ap, ApmTE, idx):
    """Ep = 2 Apm_TE_factor Re(Fp conj(Ap)) at the given (mode, j) entries."""
    return np.array([2.0 * ApmTE[im, j] * np.real(Fp[im, j] * np.conj(Ap[im, j])) for im, j in idx])
