"""advectFGridScalar restated in numpy (FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU entry
point wv_advect_f_grid_scalar_device; nothing in the library uses it.

It follows WVTransformConstantStratificationKernel.cpp:1032-1138 (and the
scalar antialias of :1008-1030): the scalar's half spectrum (unnormalized
r2c along x, then along y) scaled by 1/(Nx Ny); d/dx = i k (k = 0 at the x
Nyquist column), d/dy = i l; d/dz by the DCT-I of the column (REDFT00, scaled
by 1/(Nz - 1), last coefficient halved), coefficient j times -pi j / Lz
(j = 0 and Nz - 1 dropped), then halved and synthesized by the DST-I of the
interior (RODFT00 over z = 1 ... Nz - 2); imaginary parts of the
self-conjugate rows cleared; unnormalized c2r of the three derivatives;
rhs = -(u dx + v dy + w dz).  The antialias keeps only the half-spectrum rows
that store a wave-vortex mode or complete one (the rows of kx = 0 and of the x
Nyquist column mirrored in ky), scaled by 1/(Nx Ny).

Arrays are [Nz, Ny, Nx] (the reference's [Nx, Ny, Nz], x fastest).
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sf


def keep_rows(Nx: int, Ny: int, storage_rows) -> np.ndarray:
    """[Ny, Nx/2 + 1] mask of the rows the scalar antialias keeps (:467-471):
    every mode's storage row, and for kx = 0 (and the x Nyquist column of an
    even Nx) its mirror in ky."""
    nxh = Nx // 2 + 1
    rows = np.asarray(storage_rows, dtype=np.int64)
    keep = np.zeros(nxh * Ny, dtype=bool)
    keep[rows] = True
    k, l = rows % nxh, rows // nxh
    edge = (k == 0) | ((Nx % 2 == 0) & (k == Nx // 2))
    keep[k[edge] + nxh * ((Ny - l[edge]) % Ny)] = True
    return keep.reshape(Ny, nxh)


def _forward(f: np.ndarray, Nx: int, Ny: int) -> np.ndarray:
    return np.fft.fft(np.fft.rfft(f, axis=-1), axis=-2)


def _inverse(X: np.ndarray, Nx: int) -> np.ndarray:
    # unnormalized c2r (imaginary parts of the DC and Nyquist bins ignored)
    return np.fft.irfft(np.fft.ifft(X, axis=-2, norm="forward"), n=Nx, axis=-1, norm="forward")


def advect_f_grid_scalar(cfg: dict, scalar, uvw, antialias: bool, storage_rows) -> np.ndarray:
    Nx, Ny, Nz = int(cfg["Nx"]), int(cfg["Ny"]), int(cfg["Nz"])
    Lx, Ly, Lz = float(cfg["Lx"]), float(cfg["Ly"]), float(cfg["Lz"])
    phi = np.asarray(scalar, dtype=np.float64).reshape(Nz, Ny, Nx)
    U, V, W = np.asarray(uvw, dtype=np.float64).reshape(3, Nz, Ny, Nx)
    nxh = Nx // 2 + 1
    half = _forward(phi, Nx, Ny) / (Nx * Ny)  # :1063-1075
    k = 2.0 * np.pi * np.arange(nxh) / Lx
    if Nx % 2 == 0:
        k[Nx // 2] = 0.0
    lIndex = np.arange(Ny)
    l = 2.0 * np.pi * np.where(lIndex < (Ny + 1) // 2, lIndex, lIndex - Ny) / Ly
    dx = half * (1j * k)[None, None, :]
    dy = half * (1j * l)[None, :, None]
    # :1078-1094
    c = sf.dct(half, type=1, axis=0) / (Nz - 1)
    c[Nz - 1] *= 0.5
    c[0] = 0.0
    c[Nz - 1] = 0.0
    c *= -(np.pi * np.arange(Nz) / Lz)[:, None, None]
    c[1:Nz - 1] *= 0.5
    dz = np.zeros_like(half)
    if Nz > 2:
        dz[1:Nz - 1] = sf.dst(c[1:Nz - 1], type=1, axis=0)
    # :1095-1100
    ks = [0] + ([Nx // 2] if Nx % 2 == 0 else [])
    ls = [0] + ([Ny // 2] if Ny % 2 == 0 else [])
    for a in (dx, dy, dz):
        for kk in ks:
            for ll in ls:
                a[:, ll, kk] = a[:, ll, kk].real
    rhs = -(U * _inverse(dx, Nx) + V * _inverse(dy, Nx) + W * _inverse(dz, Nx))
    if antialias:  # :1008-1030
        X = _forward(rhs, Nx, Ny) / (Nx * Ny)
        X[:, ~keep_rows(Nx, Ny, storage_rows)] = 0.0
        rhs = _inverse(X, Nx)
    return rhs
