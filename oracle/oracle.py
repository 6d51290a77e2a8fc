"""ctypes bindings for the CPU checker libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg.  ``Oracle`` wraps liboracle.so (the FP64 CPU
restatement of CompiledKernel + WVFixedStepRK4); ``Reference`` wraps
_ref/libwvref.so (the reference kernel itself, built in place from the
reference sources with its direct-DFT engine; only present where that tree
was available at build time).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libwvref.so")
REF_FAST_SO = os.path.join(HERE, "_ref", "libwvref_fast.so")


class Config(C.Structure):
    _fields_ = [
        ("Nx", C.c_size_t), ("Ny", C.c_size_t), ("Nz", C.c_size_t), ("Nj", C.c_size_t),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double), ("N0", C.c_double),
        ("rho0", C.c_double), ("g", C.c_double), ("planetaryRadius", C.c_double),
        ("rotationRate", C.c_double), ("latitude", C.c_double),
        ("isHydrostatic", C.c_int32), ("shouldAntialias", C.c_int32),
    ]


def make_config(Nx, Ny, Nz, Nj=None, Lx=50e3, Ly=None, Lz=1300.0, N0=5.2e-3, rho0=1025.0,
                g=9.81, planetaryRadius=6.371e6, rotationRate=7.2921e-5, latitude=33.0,
                isHydrostatic=False, shouldAntialias=True):
    if Nj is None:  # WVGeometryDoublyPeriodicStratifiedConstant.m:61-68
        Nj = (2 * (Nz - 1)) // 3 if shouldAntialias and Nz - 1 > 3 else Nz - 1
    if Ly is None:
        Ly = Lx
    return dict(Nx=Nx, Ny=Ny, Nz=Nz, Nj=Nj, Lx=Lx, Ly=Ly, Lz=Lz, N0=N0, rho0=rho0, g=g,
                planetaryRadius=planetaryRadius, rotationRate=rotationRate, latitude=latitude,
                isHydrostatic=bool(isHydrostatic), shouldAntialias=bool(shouldAntialias))


def _cfg_struct(cfg: dict) -> Config:
    s = Config()
    for name, _ in Config._fields_:
        v = cfg[name]
        setattr(s, name, int(v) if name in ("isHydrostatic", "shouldAntialias") else v)
    return s


_P = C.POINTER(C.c_double)


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


class _Lib:
    prefix = ""
    path = ""

    def __init__(self, cfg: dict):
        self.lib = self._load()
        self.cfg = dict(cfg)
        msg = C.create_string_buffer(512)
        self.h = getattr(self.lib, self.prefix + "create")(C.byref(_cfg_struct(cfg)), msg, 512)
        if not self.h:
            raise ValueError(msg.value.decode())
        self.Nkl = int(getattr(self.lib, self.prefix + "nkl")(self.h))
        self.Nj = cfg["Nj"]
        self.M = self.Nkl * self.Nj
        self.spatial = (cfg["Nx"], cfg["Ny"], cfg["Nz"])

    _cache: dict = {}

    @classmethod
    def _load(cls):
        if cls.path not in _Lib._cache:
            if not os.path.exists(cls.path):
                raise FileNotFoundError(cls.path)
            lib = C.CDLL(cls.path)
            p = cls.prefix
            getattr(lib, p + "create").restype = C.c_void_p
            getattr(lib, p + "create").argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
            getattr(lib, p + "destroy").argtypes = [C.c_void_p]
            getattr(lib, p + "nkl").restype = C.c_size_t
            getattr(lib, p + "nkl").argtypes = [C.c_void_p]
            getattr(lib, p + "coriolis").restype = C.c_double
            getattr(lib, p + "coriolis").argtypes = [C.c_void_p]
            getattr(lib, p + "table").restype = C.c_long
            getattr(lib, p + "table").argtypes = [C.c_void_p, C.c_char_p, _P]
            getattr(lib, p + "modes").argtypes = [C.c_void_p] + [C.c_void_p] * 9
            for fn in ("nonlinear_flux", "inverse_uvweta", "inverse_uvw"):
                getattr(lib, p + fn).restype = C.c_int
                getattr(lib, p + fn).argtypes = [C.c_void_p, C.c_double, C.c_double] + [_P] * (6 if fn == "nonlinear_flux" else 4)
            for fn in ("forward_uvweta", "forward_uveta"):
                getattr(lib, p + fn).restype = C.c_int
                getattr(lib, p + fn).argtypes = [C.c_void_p, _P, C.c_double, C.c_double, _P, _P, _P]
            getattr(lib, p + "fg_all_derivatives").restype = C.c_int
            getattr(lib, p + "fg_all_derivatives").argtypes = [C.c_void_p, C.c_int, _P, _P, _P]
            getattr(lib, p + "nonlinear_flux_using_advection_fields").restype = C.c_int
            getattr(lib, p + "nonlinear_flux_using_advection_fields").argtypes = (
                [C.c_void_p, C.c_double, C.c_double] + [_P] * 7)
            getattr(lib, p + "state_field_derivatives").restype = C.c_int
            getattr(lib, p + "state_field_derivatives").argtypes = (
                [C.c_void_p, C.c_double, C.c_double, _P, _P, _P, C.c_int, _P])
            if hasattr(lib, p + "advect_f_grid_scalar"):
                getattr(lib, p + "advect_f_grid_scalar").restype = C.c_int
                getattr(lib, p + "advect_f_grid_scalar").argtypes = [C.c_void_p, _P, _P, C.c_int, _P]
            if hasattr(lib, p + "engine_compare"):
                lib.wvr_engine_compare.restype = C.c_double
                lib.wvr_engine_compare.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_size_t, C.c_uint]
                lib.wvr_set_threads.argtypes = [C.c_int]
            if p == "wvo_":
                lib.wvo_rk4_advance.restype = C.c_int
                lib.wvo_rk4_advance.argtypes = [C.c_void_p, _P, _P, _P, _P, C.c_double, C.c_double, C.c_int]
                lib.wvo_rk4_linear_test.argtypes = [C.c_double, C.c_double, _P, C.c_double, C.c_int]
                lib.wvo_energy.restype = C.c_double
                lib.wvo_energy.argtypes = [C.c_void_p, _P, _P, _P]
                lib.wvo_energy_spatial.restype = C.c_double
                lib.wvo_energy_spatial.argtypes = [C.c_void_p, _P]
                lib.wvo_set_threads.argtypes = [C.c_int]
            _Lib._cache[cls.path] = lib
        return _Lib._cache[cls.path]

    def close(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.prefix + "destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    # descriptor -----------------------------------------------------------
    def modes(self):
        n = self.Nkl
        out = dict(kMode=np.zeros(n, np.int64), lMode=np.zeros(n, np.int64), k=np.zeros(n), l=np.zeros(n),
                   Kh=np.zeros(n), cosAlpha=np.zeros(n), sinAlpha=np.zeros(n), row=np.zeros(n, np.int64),
                   conj=np.zeros(n, np.uint8))
        self._fn("modes")(self.h, *[out[k].ctypes.data for k in
                                    ("kMode", "lMode", "k", "l", "Kh", "cosAlpha", "sinAlpha", "row", "conj")])
        return out

    def table(self, name: str) -> np.ndarray:
        buf = np.zeros(2 * self.M + 2 * self.cfg["Nz"] + 16)
        n = self._fn("table")(self.h, name.encode(), _ptr(buf))
        if n < 0:
            raise KeyError(name)
        return buf[:n].copy()

    def coriolis(self) -> float:
        return float(self._fn("coriolis")(self.h))

    # helpers --------------------------------------------------------------
    def _coef(self, a=None):
        if a is None:
            return np.zeros((self.Nkl, self.Nj), np.complex128)
        a = np.ascontiguousarray(np.asarray(a, np.complex128).reshape(self.Nkl, self.Nj))
        return a

    def _fields(self, nf):
        Nx, Ny, Nz = self.spatial
        return np.zeros((nf, Nz, Ny, Nx))

    # operations -----------------------------------------------------------
    def nonlinear_flux(self, Ap, Am, A0, t=0.0, t0=0.0):
        Ap, Am, A0 = self._coef(Ap), self._coef(Am), self._coef(A0)
        Fp, Fm, F0 = self._coef(), self._coef(), self._coef()
        rc = self._fn("nonlinear_flux")(self.h, t, t0, *map(_ptr, (Ap, Am, A0, Fp, Fm, F0)))
        if rc:
            raise RuntimeError(f"nonlinear_flux failed: {rc}")
        return Fp, Fm, F0

    def nonlinear_flux_using_advection_fields(self, Ap, Am, A0, uvw, t=0.0, t0=0.0):
        Ap, Am, A0 = self._coef(Ap), self._coef(Am), self._coef(A0)
        Nx, Ny, Nz = self.spatial
        uvw = np.ascontiguousarray(uvw, dtype=np.float64).reshape(3, Nz, Ny, Nx)
        Fp, Fm, F0 = self._coef(), self._coef(), self._coef()
        rc = self._fn("nonlinear_flux_using_advection_fields")(self.h, t, t0, *map(_ptr, (Ap, Am, A0, uvw, Fp, Fm, F0)))
        if rc:
            raise RuntimeError(f"nonlinear_flux_using_advection_fields failed: {rc}")
        return Fp, Fm, F0

    def state_field_derivatives(self, Ap, Am, A0, field, t=0.0, t0=0.0):
        Ap, Am, A0 = self._coef(Ap), self._coef(Am), self._coef(A0)
        out = self._fields(3)
        rc = self._fn("state_field_derivatives")(self.h, t, t0, *map(_ptr, (Ap, Am, A0)), int(field), _ptr(out))
        if rc:
            raise RuntimeError(f"state_field_derivatives failed: {rc}")
        return out

    def apv(self, Ap, Am, A0, t=0.0, t0=0.0):
        """Available potential vorticity, Operations/APVOperation.m:15-26:
        zeta_z - zeta_x eta_x - zeta_y eta_y - (zeta_z + f) eta_z with
        zeta_x = w_y - v_z, zeta_y = u_z - w_x, zeta_z = v_x - u_y
        (@WVTransform/classDefinedOperationForKnownVariable.m:126-137) and
        eta_true = eta (the constant-N background as the no-motion profile,
        for which EtaTrueOperation.m's inversion returns z - eta), from this
        library's state-field derivatives.  Returns (Nz, Ny, Nx)."""
        du, dv, dw, de = (self.state_field_derivatives(Ap, Am, A0, f, t, t0) for f in range(4))
        zx = dw[1] - dv[2]
        zy = du[2] - dw[0]
        zz = dv[0] - du[1]
        return zz - zx * de[0] - zy * de[1] - (zz + self.coriolis()) * de[2]

    def inverse_uvweta(self, Ap, Am, A0, t=0.0, t0=0.0):
        Ap, Am, A0 = self._coef(Ap), self._coef(Am), self._coef(A0)
        out = self._fields(4)
        rc = self._fn("inverse_uvweta")(self.h, t, t0, *map(_ptr, (Ap, Am, A0, out)))
        if rc:
            raise RuntimeError(f"inverse failed: {rc}")
        return out

    def inverse_uvw(self, Ap, Am, A0, t=0.0, t0=0.0):
        Ap, Am, A0 = self._coef(Ap), self._coef(Am), self._coef(A0)
        out = self._fields(3)
        rc = self._fn("inverse_uvw")(self.h, t, t0, *map(_ptr, (Ap, Am, A0, out)))
        if rc:
            raise RuntimeError(f"inverse failed: {rc}")
        return out

    def forward(self, fields, t=0.0, t0=0.0):
        fields = np.ascontiguousarray(fields, dtype=np.float64)
        Ap, Am, A0 = self._coef(), self._coef(), self._coef()
        fn = "forward_uveta" if self.cfg["isHydrostatic"] else "forward_uvweta"
        rc = self._fn(fn)(self.h, _ptr(fields), t, t0, *map(_ptr, (Ap, Am, A0)))
        if rc:
            raise RuntimeError(f"forward failed: {rc}")
        return Ap, Am, A0

    def advect_f_grid_scalar(self, scalar, uvw, antialias):
        """advectFGridScalar (...Kernel.cpp:1032-1138); arrays [Nz, Ny, Nx]
        (the reference's [Nx, Ny, Nz], x fastest) and [3, Nz, Ny, Nx]."""
        Nx, Ny, Nz = self.spatial
        scalar = np.ascontiguousarray(scalar, dtype=np.float64).reshape(Nz, Ny, Nx)
        uvw = np.ascontiguousarray(uvw, dtype=np.float64).reshape(3, Nz, Ny, Nx)
        out = np.zeros((Nz, Ny, Nx))
        rc = self._fn("advect_f_grid_scalar")(self.h, _ptr(scalar), _ptr(uvw), int(bool(antialias)), _ptr(out))
        if rc:
            raise RuntimeError(f"advect_f_grid_scalar failed: {rc}")
        return out

    def fg_all(self, f_family, Apm, A0):
        Apm, A0 = self._coef(Apm), self._coef(A0)
        out = self._fields(4)
        rc = self._fn("fg_all_derivatives")(self.h, int(bool(f_family)), _ptr(Apm), _ptr(A0), _ptr(out))
        if rc:
            raise RuntimeError(f"fg_all failed: {rc}")
        return out


class Oracle(_Lib):
    prefix = "wvo_"
    path = ORACLE_SO

    def rk4(self, Ap, Am, A0, t, dt, nsteps, t0=0.0):
        Ap, Am, A0 = (self._coef(x).copy() for x in (Ap, Am, A0))
        tt = np.array([t], np.float64)
        rc = self.lib.wvo_rk4_advance(self.h, _ptr(Ap), _ptr(Am), _ptr(A0), _ptr(tt), t0, dt, int(nsteps))
        if rc:
            raise RuntimeError("rk4 failed")
        return Ap, Am, A0, float(tt[0])

    def energy(self, Ap, Am, A0) -> float:
        Ap, Am, A0 = self._coef(Ap), self._coef(Am), self._coef(A0)
        return float(self.lib.wvo_energy(self.h, _ptr(Ap), _ptr(Am), _ptr(A0)))

    def energy_spatial(self, fields) -> float:
        fields = np.ascontiguousarray(fields, dtype=np.float64)
        return float(self.lib.wvo_energy_spatial(self.h, _ptr(fields)))

    @classmethod
    def rk4_linear_test(cls, lam: complex, y0: complex, dt: float, nsteps: int) -> complex:
        lib = cls._load()
        y = np.array([y0.real, y0.imag])
        lib.wvo_rk4_linear_test(lam.real, lam.imag, _ptr(y), dt, int(nsteps))
        return complex(y[0], y[1])

    @classmethod
    def set_threads(cls, n: int):
        cls._load().wvo_set_threads(int(n))


class Reference(_Lib):
    prefix = "wvr_"
    path = REF_SO

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)


class ReferenceFast(Reference):
    """The reference kernel over the build-authored fast engine
    (fast_fft_engine.cpp): the reference's CPU path at production sizes."""
    path = REF_FAST_SO

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_FAST_SO)

    @classmethod
    def set_threads(cls, n: int):
        """Threads of the fast engine for kernels created afterwards (0: OpenMP default)."""
        cls._load().wvr_set_threads(int(n))

    @classmethod
    def engine_compare(cls, kind: str, n0: int, n1: int = 1, batch: int = 3, seed: int = 1) -> float:
        """max|fast - direct| / max|direct| of one plan of the given kind
        ("r2c", "c2r", "dct1", "dst1") against WVReferenceFFTEngine."""
        k = {"r2c": 0, "c2r": 1, "dct1": 2, "dst1": 3}[kind]
        return float(cls._load().wvr_engine_compare(k, n0, n1, batch, seed))
