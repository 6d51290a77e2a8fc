"""Seeded synthetic workloads at the benchmark configurations (shared by the
benchmark and the GPU tests).

arbitrary_stratification_512: the hydrostatic transform of N^2 = N0^2
exp(2z/b) (N0 = 5.2e-3 s^-1, b = 1300 m), Lz = 4000 m, 512 x 512 points
(Lx = Ly = 400 km), 65 quadrature points, 64 vertical modes; a flow with
|k|^-2 wave and |k|^-3 geostrophic amplitudes (equal wave and geostrophic
energy), rms horizontal speed 0.1 m/s over the quadrature, and the 0.5-CFL
fixed time step.
"""
from __future__ import annotations

import numpy as np

from . import flows
from . import stratification as S

ARBITRARY_STRATIFICATION_512 = dict(Nx=512, Ny=512, Nz=65, Nj=64, Lx=400e3, Lz=4000.0, N0=5.2e-3, b=1300.0,
                                    seed=3001, urms=0.1)


def exponential_operators(Nz: int, Nj: int, Lz: float, N0: float = 5.2e-3, b: float = 1300.0, g: float = 9.81):
    N2, dN2 = S.exponential(N0, b)
    return S.vertical_operators(S.Stratification(N2, Lz=Lz, g=g, dN2=dN2), Nz, Nj)


def arbitrary_stratification_512(wvamd_module, precision: str = "fp64", params: dict | None = None):
    """(config, operators, kernel, [Ap, Am, A0] host arrays, dt) of the
    512 x 512 x 65 arbitrary-stratification workload (params overrides the
    defaults, e.g. a smaller grid)."""
    c = dict(ARBITRARY_STRATIFICATION_512, **(params or {}))
    cfg = wvamd_module.make_config(c["Nx"], c["Ny"], c["Nz"], c["Nj"], Lx=c["Lx"], Lz=c["Lz"], N0=c["N0"],
                                   isHydrostatic=True)
    ops = exponential_operators(c["Nz"], c["Nj"], c["Lz"], c["N0"], c["b"], cfg["g"])
    k = wvamd_module.WVTransformHydrostaticKernel(cfg, ops, precision)
    md = k.modes()
    A = flows.random_flow_shape(md["Kh"], c["Nj"], k.coefficient_table("ApmTE"), k.coefficient_table("A0TE"),
                                seed=c["seed"], jmax=c["Nj"], slope=-2.0, geo_slope=-3.0)
    f = k.transform_wave_vortex_to_uvw(*[k.to_device(a) for a in A]).cpu().numpy()
    s = flows.scale_factor_for_rms(f, c["Lz"], c["urms"], ops["w"])
    A = [np.asarray(a) * s for a in A]
    dt = flows.cfl_time_step(f * s, md["k"], md["l"], k.coefficient_table("omega"), c["Lz"], 0.5)
    return cfg, ops, k, A, dt
