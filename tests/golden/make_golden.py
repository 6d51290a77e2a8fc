"""Write golden vectors from the REFERENCE kernel (oracle/_ref/libwvref.so).

Run in the container that has the reference tree (after ``make -C oracle``):

    python tests/golden/make_golden.py            # tiny grids (seconds)
    python tests/golden/make_golden.py --c0       # + the 64x64x33 flux (~4 min)

Inputs are seeded numpy draws (or the seeded synthetic flow of
wvamd.flows for the 64x64x33 case); every output comes from the reference's
own WVTransformConstantStratificationKernel with its direct-DFT engine.
Files are written next to this script as compressed .npz (no pickles).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "wave-vortex-model_amd"))

from oracle import Oracle, Reference, make_config  # noqa: E402
from wvamd import flows  # noqa: E402

TABLES = ["omega", "fWaveScale", "UApField", "VApField", "WApField", "NApField", "UA0Field", "VA0Field",
          "NA0Field", "A0FromVorticity", "A0FromBuoyancy", "ApmDProjection", "ApmNProjection", "ApmDScaled",
          "z", "h0", "verticalWavenumber", "Fg", "Gg", "inertialScale", "gWaveScale", "apmWProjectionPrefactor"]

# (name, Nx, Ny, Nz, Nj, Lx, Ly, hydrostatic, antialias)
TINY = [
    ("c6x5x7_nh", 6, 5, 7, 4, 15e3, 12e3, False, True),
    ("c6x5x7_h", 6, 5, 7, 4, 15e3, 12e3, True, True),
    ("c8x6x7_nh_noaa", 8, 6, 7, 5, 15e3, 12e3, False, False),
    ("c7x5x7_h_noaa", 7, 5, 7, 6, 15e3, 12e3, True, False),
    ("c16x16x9_nh", 16, 16, 9, 5, 50e3, 50e3, False, True),
    ("c16x16x9_h", 16, 16, 9, 5, 50e3, 50e3, True, True),
    ("c12x10x9_nh_maxNj", 12, 10, 9, 8, 15e3, 12e3, False, True),
    ("c8x8x5_nh_Nj1", 8, 8, 5, 1, 15e3, 15e3, False, True),
    ("c15x13x9_h", 15, 13, 9, 5, 15e3, 12e3, True, True),
    ("c32x32x17_nh", 32, 32, 17, 11, 50e3, 50e3, False, True),
]

DESCRIPTOR_ONLY = [
    ("d16x12x9_nh", 16, 12, 9, 5, 15e3, 12e3, False, True),
    ("d64x64x33_nh", 64, 64, 33, 21, 50e3, 50e3, False, True),
    ("d256x256x129_nh", 256, 256, 129, 85, 200e3, 200e3, False, True),
]


def cfg_of(row):
    _, Nx, Ny, Nz, Nj, Lx, Ly, hyd, aa = row
    return make_config(Nx, Ny, Nz, Nj, Lx=Lx, Ly=Ly, isHydrostatic=hyd, shouldAntialias=aa)


def cfg_arrays(cfg):
    keys = ["Nx", "Ny", "Nz", "Nj", "Lx", "Ly", "Lz", "N0", "rho0", "g", "planetaryRadius",
            "rotationRate", "latitude", "isHydrostatic", "shouldAntialias"]
    return {"cfg_" + k: np.array(float(cfg[k])) for k in keys}


def descriptor_payload(r: Reference, full: bool):
    md = r.modes()
    out = {"kMode": md["kMode"].astype(np.int32), "lMode": md["lMode"].astype(np.int32),
           "row": md["row"].astype(np.int64), "conj": md["conj"], "coriolis": np.array(r.coriolis())}
    rng = np.random.default_rng(7)
    for name in TABLES:
        t = r.table(name)
        if full or t.size <= 4096:
            out["T_" + name] = t
        else:
            idx = np.sort(rng.choice(t.size, 512, replace=False))
            out["S_" + name + "_idx"] = idx
            out["S_" + name] = t[idx]
            out["C_" + name] = np.array([np.sum(t), np.sum(t * t), np.sum(t * np.cos(0.001 * np.arange(t.size)))])
    return out


def rand_coef(rng, r):
    sh = (r.Nkl, r.Nj)
    return [rng.standard_normal(sh) + 1j * rng.standard_normal(sh) for _ in range(3)]


def write(name, payload):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **payload)
    print("wrote", os.path.relpath(path, ROOT), f"{os.path.getsize(path) / 1024:.1f} KiB")


def tiny_cases():
    for row in TINY:
        cfg = cfg_of(row)
        r = Reference(cfg)
        rng = np.random.default_rng(1000 + TINY.index(row))
        Ap, Am, A0 = rand_coef(rng, r)
        t, t0 = 37.5, 2.0
        t_start = time.time()
        Fp, Fm, F0 = r.nonlinear_flux(Ap, Am, A0, t=t, t0=t0)
        Nx, Ny, Nz = r.spatial
        inv = r.inverse_uvweta(Ap, Am, A0, t=t, t0=t0)
        inv3 = r.inverse_uvw(Ap, Am, A0, t=t, t0=t0)
        nf = 3 if cfg["isHydrostatic"] else 4
        fields = rng.standard_normal((nf, Nz, Ny, Nx))
        fwd = r.forward(fields, t=t, t0=t0)
        fAll = r.fg_all(True, Ap, A0)
        gAll = r.fg_all(False, Ap, A0)
        zero = r.nonlinear_flux(*(np.zeros_like(Ap) for _ in range(3)), t=0.5, t0=0.0)
        payload = dict(cfg_arrays(cfg), t=np.array(t), t0=np.array(t0), Ap=Ap, Am=Am, A0=A0, Fp=Fp, Fm=Fm, F0=F0,
                       zero_flux_maxabs=np.array(max(np.abs(z).max() for z in zero)))
        if Nx * Ny * Nz <= 4096:  # spatial outputs only for the smallest grids, to keep fixtures small
            payload.update(inv_uvweta=inv, inv_uvw=inv3, fwd_fields=fields, fwd_Ap=fwd[0], fwd_Am=fwd[1],
                           fwd_A0=fwd[2], fAll=fAll, gAll=gAll)
        payload.update(descriptor_payload(r, full=True))
        print(row[0], f"Nkl={r.Nkl} ({time.time() - t_start:.2f}s)")
        write("golden_" + row[0], payload)


def descriptor_cases():
    for row in DESCRIPTOR_ONLY:
        cfg = cfg_of(row)
        r = Reference(cfg)
        payload = dict(cfg_arrays(cfg))
        payload.update(descriptor_payload(r, full=False))
        print(row[0], "Nkl", r.Nkl)
        write("descriptor_" + row[0], payload)


def c0_flux():
    """Reference flux at the 64x64x33 configuration on the seeded synthetic flow."""
    cfg = make_config(64, 64, 33, 21, Lx=50e3, Ly=50e3)
    r = Reference(cfg)
    o = Oracle(cfg)
    Kh = o.modes()["Kh"]
    ApmTE = o.table("ApmTE").reshape(o.Nkl, o.Nj)
    A0TE = o.table("A0TE").reshape(o.Nkl, o.Nj)
    Ap, Am, A0 = flows.random_flow_shape(Kh, o.Nj, ApmTE, A0TE, seed=4001, jmax=16)
    fields = r.inverse_uvweta(Ap, Am, A0)
    s = flows.scale_factor_for_rms(fields, cfg["Lz"], 0.05)
    Ap, Am, A0 = Ap * s, Am * s, A0 * s
    t0 = time.time()
    Fp, Fm, F0 = r.nonlinear_flux(Ap, Am, A0, t=0.0, t0=0.0)
    print(f"c0 reference flux {time.time() - t0:.1f}s")
    write("golden_c64x64x33_flow", dict(cfg_arrays(cfg), t=np.array(0.0), t0=np.array(0.0), Ap=Ap, Am=Am, A0=A0,
                                        Fp=Fp, Fm=Fm, F0=F0, seed=np.array(4001), jmax=np.array(16)))


# advectFGridScalar fixtures: even and odd grids, with and without the 2/3 rule
SCALAR = [("s16x16x9_nh", 16, 16, 9, False, True), ("s15x13x9_h", 15, 13, 9, True, True),
          ("s16x12x9_nh_noaa", 16, 12, 9, False, False)]


def scalar_cases():
    for name, Nx, Ny, Nz, hydro, aa in SCALAR:
        cfg = make_config(Nx, Ny, Nz, Lx=50e3, Ly=40e3, isHydrostatic=hydro, shouldAntialias=aa)
        r = Reference(cfg)
        rng = np.random.default_rng(7000 + SCALAR.index((name, Nx, Ny, Nz, hydro, aa)))
        scalar = rng.standard_normal((Nz, Ny, Nx))
        uvw = rng.standard_normal((3, Nz, Ny, Nx)) * 0.1
        rhs = r.advect_f_grid_scalar(scalar, uvw, False)
        rhs_aa = r.advect_f_grid_scalar(scalar, uvw, True)
        write("scalar_" + name, dict(cfg_arrays(cfg), scalar=scalar, uvw=uvw, rhs=rhs, rhs_antialiased=rhs_aa))


# transformStateFieldDerivatives and nonlinearFluxUsingAdvectionFields
# fixtures (the advecting fields are seeded draws, not the state's own)
CONTRACT = [("k8x6x7_nh_noaa", 8, 6, 7, 5, False, False), ("k6x5x7_h", 6, 5, 7, 4, True, True),
            ("k16x16x9_nh", 16, 16, 9, 5, False, True)]


def contract_cases():
    for i, (name, Nx, Ny, Nz, Nj, hydro, aa) in enumerate(CONTRACT):
        cfg = make_config(Nx, Ny, Nz, Nj, Lx=15e3, Ly=12e3, isHydrostatic=hydro, shouldAntialias=aa)
        r = Reference(cfg)
        rng = np.random.default_rng(9000 + i)
        Ap, Am, A0 = rand_coef(rng, r)
        t, t0 = 21.0, 3.0
        deriv = np.stack([r.state_field_derivatives(Ap, Am, A0, f, t=t, t0=t0) for f in range(4)])
        uvw = rng.standard_normal((3, Nz, Ny, Nx)) * 0.1
        Fp, Fm, F0 = r.nonlinear_flux_using_advection_fields(Ap, Am, A0, uvw, t=t, t0=t0)
        write("contract_" + name, dict(cfg_arrays(cfg), t=np.array(t), t0=np.array(t0), Ap=Ap, Am=Am, A0=A0,
                                       derivatives=deriv, uvw=uvw, Fp=Fp, Fm=Fm, F0=F0))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--c0", action="store_true", help="also write the 64x64x33 reference flux (slow)")
    ap.add_argument("--scalar", action="store_true", help="write only the advectFGridScalar fixtures")
    ap.add_argument("--contract", action="store_true",
                    help="write only the transformStateFieldDerivatives / UsingAdvectionFields fixtures")
    args = ap.parse_args()
    if not Reference.available():
        sys.exit("oracle/_ref/libwvref.so is missing: run `make -C oracle ref` where the reference tree exists")
    if args.scalar:
        scalar_cases()
        sys.exit(0)
    if args.contract:
        contract_cases()
        sys.exit(0)
    tiny_cases()
    descriptor_cases()
    if args.c0:
        c0_flux()
