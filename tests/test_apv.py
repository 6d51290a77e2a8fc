"""Available potential vorticity (Operations/APVOperation.m:15-26) on the CPU.

The oracle forms APV from its state-field derivatives (themselves pinned to the
reference kernel's transformStateFieldDerivatives outputs in
tests/golden/contract_*.npz).  Here the formula and the derivative
conventions are pinned independently: the oracle's spatial fields u, v, w,
eta (pinned by the inverse-transform goldens) are differentiated on the grid
by their own Fourier / cosine / sine series -- x and y by FFT, z by a least
squares fit of cos(m_j z) (u, v) or sin(m_j z) (w, eta) on the Nz points --
and APV = zeta_z - zeta_x eta_x - zeta_y eta_y - (zeta_z + f) eta_z is formed
from those, with f = 2 Omega sin(latitude).  A sign error or a swapped
component in the APV composition, or a wrong family in a z derivative, fails
this test.  Where the reference kernel is built (oracle/_ref), APV from its
own derivative fields is compared as well.
"""
import numpy as np
import pytest

from oracle import Oracle, Reference, make_config


def _state(o, seed, scale=1e-3):
    rng = np.random.default_rng(seed)
    return [scale * (rng.standard_normal((o.Nkl, o.Nj)) + 1j * rng.standard_normal((o.Nkl, o.Nj)))
            for _ in range(3)]


def _dxy(f, L, axis):
    n = f.shape[axis]
    k = 2 * np.pi / L * np.fft.fftfreq(n, 1.0 / n)
    if n % 2 == 0:
        k[n // 2] = 0.0
    shape = [1] * f.ndim
    shape[axis] = n
    return np.real(np.fft.ifft(1j * k.reshape(shape) * np.fft.fft(f, axis=axis), axis=axis))


def _dz(f, z, Lz, cosine):
    """d/dz of a (Nz, Ny, Nx) field that is a cosine (or sine) series in
    m_j = j pi / Lz on the grid z."""
    Nz = len(z)
    flat = f.reshape(Nz, -1)
    if cosine:
        m = np.arange(Nz) * np.pi / Lz
        B = np.cos(np.outer(z, m))
        coef = np.linalg.solve(B, flat)
        return (-np.sin(np.outer(z, m)) * m) @ coef
    m = np.arange(1, Nz - 1) * np.pi / Lz
    B = np.sin(np.outer(z[1:-1], m))
    coef = np.linalg.solve(B, flat[1:-1])
    return ((np.cos(np.outer(z, m)) * m) @ coef).reshape(f.shape)


def _apv_from_grid(cfg, fields):
    u, v, w, eta = fields
    Lx, Ly, Lz, Nz = cfg["Lx"], cfg["Ly"], cfg["Lz"], cfg["Nz"]
    z = np.linspace(-Lz, 0.0, Nz)
    f = 2.0 * cfg["rotationRate"] * np.sin(np.deg2rad(cfg["latitude"]))
    uy = _dxy(u, Ly, 1)
    vx = _dxy(v, Lx, 2)
    wx, wy = _dxy(w, Lx, 2), _dxy(w, Ly, 1)
    ex, ey = _dxy(eta, Lx, 2), _dxy(eta, Ly, 1)
    uz = _dz(u, z, Lz, True).reshape(u.shape)
    vz = _dz(v, z, Lz, True).reshape(v.shape)
    ez = _dz(eta, z, Lz, False)
    zx, zy, zz = wy - vz, uz - wx, vx - uy
    return zz - zx * ex - zy * ey - (zz + f) * ez


CASES = [
    dict(Nx=16, Ny=16, Nz=9, Nj=5),
    dict(Nx=32, Ny=24, Nz=17, Nj=11, Ly=40e3),
    dict(Nx=16, Ny=16, Nz=17, Nj=10, isHydrostatic=True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['Nx']}x{c['Ny']}x{c['Nz']}{'h' if c.get('isHydrostatic') else ''}")
def test_oracle_apv_matches_grid_spectral_differentiation(case):
    cfg = make_config(**case)
    o = Oracle(cfg)
    assert np.allclose(o.table("z"), np.linspace(-cfg["Lz"], 0.0, cfg["Nz"]), rtol=0, atol=1e-9 * cfg["Lz"])
    A = _state(o, 7)
    t, t0 = 123.0, 11.0
    mine = o.apv(*A, t=t, t0=t0)
    fields = o.inverse_uvweta(*A, t=t, t0=t0)
    ref = _apv_from_grid(cfg, fields)
    scale = np.max(np.abs(ref))
    assert scale > 0
    err = np.max(np.abs(mine - ref)) / scale
    assert err < 1e-10, err
    # the nonlinear terms are not negligible in this state (a dropped term
    # would show): APV differs from zeta_z alone
    zz = _dxy(fields[1], cfg["Lx"], 2) - _dxy(fields[0], cfg["Ly"], 1)
    assert np.max(np.abs(ref - zz)) > 1e-3 * scale


@pytest.mark.skipif(not Reference.available(), reason="reference kernel not built here")
def test_oracle_apv_matches_reference_kernel_derivatives():
    cfg = make_config(32, 32, 17, 11)
    o, r = Oracle(cfg), Reference(cfg)
    A = _state(o, 9)
    a, b = o.apv(*A, t=50.0), r.apv(*A, t=50.0)
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) < 1e-12
