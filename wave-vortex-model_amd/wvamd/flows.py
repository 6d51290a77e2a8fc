"""Seeded synthetic wave-vortex initial states and the CFL time step.

The random state is defined by this module (MATLAB's ``rng`` twister used by
the reference's ``initWithRandomFlow``, @WVTransform/addRandomFlow.m:50-91, is
not reproducible outside MATLAB):

* every draw comes from a counter-based SplitMix64 hash of
  (seed, array id, coefficient index), so the state is identical on any
  machine and independent of evaluation order;
* coefficient magnitudes are ``(Kh/Kh_min)**slope`` (slope -2 by default,
  ``geo_slope`` for the geostrophic A0, default the same; the
  Kh = 0 column uses the Kh_min value), with uniform random phases, for
  vertical modes ``j <= jmax`` and only where the mode exists (its
  total-energy factor is nonzero);
* self-conjugate constraints of the k = l = 0 column: ``Am = conj(Ap)``
  (@WVTransformConstantStratification/transformUVEtaToWaveVortex.m:38-39) and
  real ``A0``;
* wave (Ap, Am) and geostrophic (A0) parts are scaled to carry equal total
  energy, with the closed-form total-energy factors;
* ``scale_to_rms_speed`` then rescales all three arrays so that the
  volume-averaged horizontal speed ``sqrt(<u^2 + v^2>)`` (trapezoidal
  weights in z, or the given quadrature weights, uniform in x and y) equals
  the requested value.

Arrays are numpy complex128 with shape (Nkl, Nj), i.e. the [Nj, Nkl]
column-major coefficient layout of CompiledKernel.
"""
from __future__ import annotations

import numpy as np

_MASK = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _MASK
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n uniform doubles in [0, 1) for (seed, stream); entry i depends only on i."""
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        bits = _splitmix64(key ^ (idx * np.uint64(0xD1B54A32D192ED03)))
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def random_flow_shape(Kh: np.ndarray, Nj: int, ApmTE: np.ndarray, A0TE: np.ndarray,
                      seed: int, jmax: int, slope: float = -2.0, geo_slope: float | None = None):
    """Unit-scaled Ap, Am, A0 with equal wave and geostrophic total energy.

    Kh: [Nkl] horizontal wavenumber magnitude per WV mode.
    ApmTE, A0TE: total-energy factors with shape (Nkl, Nj).
    """
    Nkl = Kh.shape[0]
    M = Nkl * Nj
    kh = np.asarray(Kh, np.float64)
    khmin = kh[kh > 0].min() if np.any(kh > 0) else 1.0
    jj = np.repeat(np.arange(Nj)[None, :], Nkl, axis=0)
    active = jj <= jmax
    out = []
    for stream, te in ((1, ApmTE), (2, ApmTE), (3, A0TE)):
        sl = slope if stream < 3 or geo_slope is None else geo_slope
        amp_mode = np.where(kh > 0, (np.maximum(kh, khmin) / khmin) ** sl, 1.0)
        amp = np.repeat(amp_mode[:, None], Nj, axis=1)
        ph = 2.0 * np.pi * uniform(seed, stream, M).reshape(Nkl, Nj)
        a = amp * np.exp(1j * ph)
        a[(~active) | (np.asarray(te).reshape(Nkl, Nj) <= 0)] = 0.0
        out.append(a)
    Ap, Am, A0 = out
    zero = np.where(kh == 0)[0]
    if zero.size:
        i0 = zero[0]
        Am[i0, :] = np.conj(Ap[i0, :])
        A0[i0, :] = A0[i0, :].real
    apmte = np.asarray(ApmTE).reshape(Nkl, Nj)
    a0te = np.asarray(A0TE).reshape(Nkl, Nj)
    ew = float(np.sum(apmte * (np.abs(Ap) ** 2 + np.abs(Am) ** 2)))
    eg = float(np.sum(a0te * np.abs(A0) ** 2))
    if ew > 0 and eg > 0:
        sw, sg = 1.0 / np.sqrt(ew), 1.0 / np.sqrt(eg)
        Ap, Am, A0 = Ap * sw, Am * sw, A0 * sg
    return Ap, Am, A0


def trapezoid_weights(Nz: int, Lz: float) -> np.ndarray:
    dz = Lz / (Nz - 1)
    w = np.full(Nz, dz)
    w[0] = w[-1] = 0.5 * dz
    return w


def rms_speed(fields: np.ndarray, Lz: float, weights: np.ndarray | None = None) -> float:
    """sqrt(<u^2+v^2>) of fields shaped (nfield>=2, Nz, Ny, Nx); weights: the
    vertical quadrature weights (summing to Lz), trapezoidal by default."""
    u, v = fields[0], fields[1]
    Nz = u.shape[0]
    w = trapezoid_weights(Nz, Lz) if weights is None else np.asarray(weights, np.float64)
    s = ((u * u + v * v).mean(axis=(1, 2)) * w).sum() / Lz
    return float(np.sqrt(s))


def scale_factor_for_rms(fields: np.ndarray, Lz: float, target: float, weights: np.ndarray | None = None) -> float:
    r = rms_speed(fields, Lz, weights)
    return target / r if r > 0 else 1.0


def cfl_time_step(fields: np.ndarray, k: np.ndarray, l: np.ndarray, omega: np.ndarray,
                  Lz: float, cfl: float = 0.5, has_w: bool = True) -> float:
    """Fixed step for the given CFL number, following
    Integrators/WVModelFixedTimeStepMethods.m:50-95 (timeStepForCFL):
    advective dt = cfl*dx/max|u_h| with dx = pi/max(|k|,|l|)
    (@WVGeometryDoublyPeriodic/WVGeometryDoublyPeriodic.m:865-879), limited by
    the vertical advective dt = cfl/max|w/(1.5 dz)| over levels 2:end, and the
    oscillatory dt = cfl*2*pi/max|omega| over j >= 1; the smaller is returned.

    fields: (>=3, Nz, Ny, Nx) with u, v, w; omega shaped (Nkl, Nj).
    """
    u, v = fields[0], fields[1]
    uv_max = float(np.sqrt(u * u + v * v).max())
    dx = np.pi / max(float(np.abs(k).max()), float(np.abs(l).max()))
    adv = cfl * dx / uv_max if uv_max > 0 else np.inf
    Nz = u.shape[0]
    if has_w:
        dz = Lz / (Nz - 1)
        W = fields[2][1:]
        ratio = np.abs(W / (1.5 * dz)).max()
        if ratio > 0:
            adv = min(adv, cfl / ratio)
    om = np.asarray(omega)
    osc = np.inf
    if om.shape[1] > 1:
        osc = cfl * 2.0 * np.pi / float(np.abs(om[:, 1:]).max())
    return float(min(adv, osc))
