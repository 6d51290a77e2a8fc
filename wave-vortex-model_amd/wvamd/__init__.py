"""Python host bindings of libwvamd.so (the MI355X wave-vortex flux / RK4 library).

The class ``WVTransformConstantStratificationKernel`` mirrors the reference
C++ contract of the same name
(CompiledKernel/include/WaveVortexKernel/WVTransformConstantStratificationKernel.hpp:47-110):
``create`` from a configuration, the descriptor accessors, ``nonlinear_flux``,
the inverse transforms, and the fixed-step RK4 of
PortableRuntime/src/WVRungeKutta.cpp.  Failures raise ``WVKernelError`` with
the reference's status code (WVKernelTypes.hpp:22-35).

Arrays are torch tensors on the GPU (device entry points, state resident in
HBM) or numpy arrays on the host (host entry points, copied in and out).
Coefficients are complex128 with shape (Nkl, Nj) -- the reference's
column-major [Nj, Nkl] -- and spatial bundles float64 with shape
(Nfield, Nz, Ny, Nx) -- the reference's [Nx, Ny, Nz, Nfield].

The library must be built (``make -C wave-vortex-model_amd/csrc``); there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwvamd.so")

STATUS_NAMES = [
    "success", "invalidConfiguration", "invalidShape", "invalidPointer", "overlappingArrays",
    "sizeOverflow", "allocationFailure", "fftPlanFailure", "fftExecutionFailure", "numericalFailure",
    "unsupportedOperation", "reentrantExecution",
]
PRECISIONS = {"fp64": 0, "mixed": 1, "mixed-split16": 2}
CONTRACT_VERSION = 4


class WVKernelError(RuntimeError):
    def __init__(self, code: int, message: str):
        self.code = int(code)
        self.status = STATUS_NAMES[self.code] if 0 <= self.code < len(STATUS_NAMES) else str(code)
        super().__init__(f"{self.status}: {message}")


class Config(C.Structure):
    """wv_config / WVTransformConstantStratificationConfiguration (WVKernelTypes.hpp:129-146)."""
    _fields_ = [
        ("contractVersion", C.c_uint32),
        ("Nx", C.c_size_t), ("Ny", C.c_size_t), ("Nz", C.c_size_t), ("Nj", C.c_size_t),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double), ("N0", C.c_double),
        ("rho0", C.c_double), ("g", C.c_double), ("planetaryRadius", C.c_double),
        ("rotationRate", C.c_double), ("latitude", C.c_double),
        ("isHydrostatic", C.c_int32), ("shouldAntialias", C.c_int32),
    ]


class StageTiming(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("total_ms", C.c_double), ("launches", C.c_uint64),
                ("bytes_per_launch", C.c_double)]


def default_nj(Nz: int, antialias: bool = True) -> int:
    """Default vertical mode count: floor(2*(Nz-1)/3) when antialiased and
    Nz-1 > 3, else Nz-1
    (@WVGeometryDoublyPeriodicStratifiedConstant/WVGeometryDoublyPeriodicStratifiedConstant.m:61-68)."""
    maxNj = Nz - 1
    return (2 * maxNj) // 3 if antialias and maxNj > 3 else maxNj


def make_config(Nx, Ny, Nz, Nj=None, Lx=50e3, Ly=None, Lz=1300.0, N0=5.2e-3, rho0=1025.0, g=9.81,
                planetaryRadius=6.371e6, rotationRate=7.2921e-5, latitude=33.0, isHydrostatic=False,
                shouldAntialias=True) -> dict:
    return dict(Nx=int(Nx), Ny=int(Ny), Nz=int(Nz), Nj=int(default_nj(Nz, shouldAntialias) if Nj is None else Nj), Lx=float(Lx),
                Ly=float(Lx if Ly is None else Ly), Lz=float(Lz), N0=float(N0), rho0=float(rho0), g=float(g),
                planetaryRadius=float(planetaryRadius), rotationRate=float(rotationRate),
                latitude=float(latitude), isHydrostatic=bool(isHydrostatic), shouldAntialias=bool(shouldAntialias))


def _config_struct(cfg: dict) -> Config:
    s = Config()
    s.contractVersion = int(cfg.get("contractVersion", CONTRACT_VERSION))
    for name, _ in Config._fields_[1:]:
        v = cfg[name]
        setattr(s, name, int(bool(v)) if name in ("isHydrostatic", "shouldAntialias") else v)
    return s


_lib = None
EXPORTED_SYMBOLS = [
    "wv_create", "wv_destroy", "wv_last_error", "wv_contract_version", "wv_nkl", "wv_coriolis_frequency",
    "wv_descriptor_modes", "wv_descriptor_table", "wv_persistent_bytes", "wv_nonlinear_flux_device",
    "wv_nonlinear_flux", "wv_nonlinear_flux_with_advection_fields_device", "wv_advect_f_grid_scalar_device",
    "wv_transform_wave_vortex_to_uvweta_device", "wv_transform_wave_vortex_to_uvw_device",
    "wv_transform_uvweta_to_wave_vortex_device", "wv_transform_uveta_to_wave_vortex_device",
    "wv_transform_to_spatial_domain_with_f_all_derivatives_device",
    "wv_transform_to_spatial_domain_with_g_all_derivatives_device",
    "wv_rk4_advance_device", "wv_total_energy_device", "wv_set_stage_instrumentation", "wv_stage_timings",
    "wv_reset_stage_timings", "wv_create_hydrostatic", "wv_nonlinear_flux_vjp_device",
    "wv_transform_wave_vortex_to_uvweta_vjp_device", "wv_rk4_step_vjp_device", "wv_rk4_adjoint_device",
    "wv_scratch_bytes", "wv_metrics", "wv_engine_identifier", "wv_engine_library_identity",
    "wv_nonlinear_flux_schedule_identifier", "wv_phase_implementation_identifier",
    "wv_coefficient_arithmetic_mode_identifier", "wv_inverse_normalization_placement_identifier",
    "wv_optimization_implementation_identifier", "wv_coefficient_worker_count", "wv_phase_reservation_bytes",
    "wv_nonlinear_flux_using_advection_fields_device", "wv_prepare_scalar_advection",
    "wv_transform_state_field_derivatives_device", "wv_rk4_advance_to_time_device", "wv_rk4_counters",
    "wv_debug_inject_launch_failure", "wv_total_energy_no_alloc_device", "wv_apv_device",
]

IDENTIFIERS = ("engine_identifier", "engine_library_identity", "nonlinear_flux_schedule_identifier",
               "phase_implementation_identifier", "coefficient_arithmetic_mode_identifier",
               "inverse_normalization_placement_identifier", "optimization_implementation_identifier")

# WVDynamicalField (...Kernel.hpp:10)
DYNAMICAL_FIELDS = {"u": 0, "v": 1, "w": 2, "eta": 3}


class KernelMetrics(C.Structure):
    """wv_kernel_metrics / WVKernelMetrics (...Kernel.hpp:12-45), same members and order."""
    _fields_ = [(n, C.c_size_t) for n in (
        "descriptorBytes", "planBytes", "engineBytes", "kernelManagementBytes", "scratchCapacityBytes",
        "scratchHighWaterBytes", "halfSpectrumScratchCapacityBytes", "realScratchCapacityBytes", "planCount",
        "executionCount", "horizontalExecutionCount", "verticalExecutionCount", "nonlinearFluxCallCount",
        "nonlinearFluxPhaseEvaluationCount", "advectionVelocityReconstructionCount", "scalarAdvectionCount",
        "scalarAntialiasCount", "bytesCopied")] + [(n, C.c_double) for n in (
        "phaseSeconds", "reconstructionSeconds", "derivativeReconstructionSeconds", "productSeconds",
        "projectionSeconds", "coefficientAssemblySeconds", "derivativeCoefficientAssemblySeconds",
        "coefficientProjectionSeconds", "scalarForwardSeconds", "scalarDerivativeAssemblySeconds",
        "scalarVerticalDerivativeSeconds", "scalarInverseSeconds", "scalarProductSeconds", "scalarAntialiasSeconds")]


class VerticalOperators(C.Structure):
    """wv_vertical_operators: pointers to the arrays of an arbitrary stratification."""
    _fields_ = [(name, C.c_void_p) for name in ("z", "N2", "dLnN2", "h", "P", "Q", "PFinv", "PF", "QGinv", "QG")]


def library() -> C.CDLL:
    """Load libwvamd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C wave-vortex-model_amd/csrc`")
    lib = C.CDLL(LIB_PATH)
    vp, dp, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)
    lib.wv_create.restype = C.c_int
    lib.wv_create.argtypes = [C.POINTER(Config), C.c_int, C.c_int, C.POINTER(vp)]
    lib.wv_create_hydrostatic.restype = C.c_int
    lib.wv_create_hydrostatic.argtypes = [C.POINTER(Config), C.POINTER(VerticalOperators), C.c_int, C.c_int, C.POINTER(vp)]
    lib.wv_destroy.argtypes = [vp]
    lib.wv_last_error.restype = C.c_char_p
    lib.wv_contract_version.restype = C.c_uint32
    lib.wv_nkl.restype = C.c_size_t
    lib.wv_nkl.argtypes = [vp]
    lib.wv_coriolis_frequency.restype = C.c_double
    lib.wv_coriolis_frequency.argtypes = [vp]
    lib.wv_persistent_bytes.restype = C.c_size_t
    lib.wv_persistent_bytes.argtypes = [vp]
    lib.wv_descriptor_modes.restype = C.c_int
    lib.wv_descriptor_modes.argtypes = [vp] + [vp] * 9
    lib.wv_descriptor_table.restype = C.c_long
    lib.wv_descriptor_table.argtypes = [vp, C.c_char_p, vp]
    for fn in ("wv_nonlinear_flux_device",):
        getattr(lib, fn).restype = C.c_int
        getattr(lib, fn).argtypes = [vp, C.c_double, C.c_double] + [vp] * 6 + [vp]
    lib.wv_nonlinear_flux.restype = C.c_int
    lib.wv_nonlinear_flux.argtypes = [vp, C.c_double, C.c_double] + [vp] * 6
    lib.wv_nonlinear_flux_with_advection_fields_device.restype = C.c_int
    lib.wv_nonlinear_flux_with_advection_fields_device.argtypes = [vp, C.c_double, C.c_double] + [vp] * 7 + [vp]
    lib.wv_advect_f_grid_scalar_device.restype = C.c_int
    lib.wv_advect_f_grid_scalar_device.argtypes = [vp, vp, vp, C.c_int, vp, vp]
    for fn in ("wv_transform_wave_vortex_to_uvweta_device", "wv_transform_wave_vortex_to_uvw_device"):
        getattr(lib, fn).restype = C.c_int
        getattr(lib, fn).argtypes = [vp, C.c_double, C.c_double, vp, vp, vp, vp, vp]
    for fn in ("wv_transform_uvweta_to_wave_vortex_device", "wv_transform_uveta_to_wave_vortex_device"):
        getattr(lib, fn).restype = C.c_int
        getattr(lib, fn).argtypes = [vp, vp, C.c_double, C.c_double, vp, vp, vp, vp]
    for fn in ("wv_transform_to_spatial_domain_with_f_all_derivatives_device",
               "wv_transform_to_spatial_domain_with_g_all_derivatives_device"):
        getattr(lib, fn).restype = C.c_int
        getattr(lib, fn).argtypes = [vp, vp, vp, vp, vp]
    lib.wv_rk4_advance_device.restype = C.c_int
    lib.wv_rk4_advance_device.argtypes = [vp, vp, vp, vp, dp, C.c_double, C.c_double, C.c_int, vp]
    lib.wv_total_energy_device.restype = C.c_int
    lib.wv_total_energy_device.argtypes = [vp, vp, vp, vp, dp, vp]
    lib.wv_set_stage_instrumentation.restype = C.c_int
    lib.wv_set_stage_instrumentation.argtypes = [vp, C.c_int]
    lib.wv_stage_timings.restype = C.c_int
    lib.wv_stage_timings.argtypes = [vp, C.POINTER(StageTiming), C.c_int]
    lib.wv_reset_stage_timings.argtypes = [vp]
    lib.wv_nonlinear_flux_vjp_device.restype = C.c_int
    lib.wv_nonlinear_flux_vjp_device.argtypes = [vp, C.c_double, C.c_double] + [vp] * 9 + [vp]
    lib.wv_transform_wave_vortex_to_uvweta_vjp_device.restype = C.c_int
    lib.wv_transform_wave_vortex_to_uvweta_vjp_device.argtypes = [vp, vp, C.c_double, C.c_double, vp, vp, vp, vp]
    lib.wv_rk4_step_vjp_device.restype = C.c_int
    lib.wv_rk4_step_vjp_device.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, C.c_double, vp, vp, vp, vp]
    lib.wv_rk4_adjoint_device.restype = C.c_int
    lib.wv_rk4_adjoint_device.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                          vp, vp, vp, vp]
    lib.wv_scratch_bytes.restype = C.c_size_t
    lib.wv_scratch_bytes.argtypes = [vp]
    lib.wv_metrics.restype = C.c_int
    lib.wv_metrics.argtypes = [vp, C.POINTER(KernelMetrics)]
    for name in IDENTIFIERS:
        getattr(lib, "wv_" + name).restype = C.c_char_p
        getattr(lib, "wv_" + name).argtypes = [vp]
    for fn in ("wv_coefficient_worker_count", "wv_phase_reservation_bytes"):
        getattr(lib, fn).restype = C.c_size_t
        getattr(lib, fn).argtypes = [vp]
    lib.wv_nonlinear_flux_using_advection_fields_device.restype = C.c_int
    lib.wv_nonlinear_flux_using_advection_fields_device.argtypes = [vp, C.c_double, C.c_double] + [vp] * 7 + [vp]
    lib.wv_prepare_scalar_advection.restype = C.c_int
    lib.wv_prepare_scalar_advection.argtypes = [vp]
    lib.wv_transform_state_field_derivatives_device.restype = C.c_int
    lib.wv_transform_state_field_derivatives_device.argtypes = [vp, C.c_double, C.c_double, vp, vp, vp, C.c_int,
                                                                vp, vp]
    lib.wv_rk4_advance_to_time_device.restype = C.c_int
    lib.wv_rk4_advance_to_time_device.argtypes = [vp, vp, vp, vp, dp, C.c_double, C.c_double, C.c_double, vp]
    lib.wv_rk4_counters.restype = C.c_int
    lib.wv_rk4_counters.argtypes = [vp, C.POINTER(C.c_uint64), dp]
    lib.wv_total_energy_no_alloc_device.restype = C.c_int
    lib.wv_total_energy_no_alloc_device.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.wv_apv_device.restype = C.c_int
    lib.wv_apv_device.argtypes = [vp, C.c_double, C.c_double, vp, vp, vp, vp, vp]
    lib.wv_debug_inject_launch_failure.restype = C.c_int
    lib.wv_debug_inject_launch_failure.argtypes = [vp, C.c_long]
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise WVKernelError(rc, library().wv_last_error().decode(errors="replace"))


def _torch():
    import torch  # noqa: WPS433 (device memory and streams only)
    return torch


class WVTransformConstantStratificationKernel:
    """Constant-stratification wave-vortex kernel on one MI355X.

    precision: "fp64" (everything in FP64) or "mixed" (FP32 spatial fields,
    spectra, FFTs and vertical MFMA products; FP64 coefficients, RK4 state
    and reductions).
    """

    def __init__(self, config: dict, precision: str = "fp64", device: int = 0):
        lib = library()
        self.config = dict(config)
        self.precision = precision
        self.device = int(device)
        h = C.c_void_p()
        _check(self._create(lib, config, precision, h))
        self._h = h
        self._finish(lib, config)

    def _create(self, lib, config, precision, h):
        return lib.wv_create(C.byref(_config_struct(config)), PRECISIONS[precision], self.device, C.byref(h))

    def _finish(self, lib, config):
        h = self._h
        self.Nkl = int(lib.wv_nkl(h))
        self.Nj = int(config["Nj"])
        self.M = self.Nkl * self.Nj
        self.spectral_shape = (self.Nkl, self.Nj)
        self.spatial_shape = (int(config["Nz"]), int(config["Ny"]), int(config["Nx"]))

    @classmethod
    def create(cls, config: dict, precision: str = "fp64", device: int = 0):
        return cls(config, precision, device)

    def close(self):
        if getattr(self, "_h", None):
            library().wv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- descriptor --
    def modes(self) -> dict:
        n = self.Nkl
        out = dict(kMode=np.zeros(n, np.int64), lMode=np.zeros(n, np.int64), k=np.zeros(n), l=np.zeros(n),
                   Kh=np.zeros(n), cosAlpha=np.zeros(n), sinAlpha=np.zeros(n), row=np.zeros(n, np.int64),
                   conj=np.zeros(n, np.uint8))
        _check(library().wv_descriptor_modes(self._h, *[out[k].ctypes.data for k in (
            "kMode", "lMode", "k", "l", "Kh", "cosAlpha", "sinAlpha", "row", "conj")]))
        return out

    def table(self, name: str) -> np.ndarray:
        lib = library()
        n = lib.wv_descriptor_table(self._h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        buf = np.zeros(n)
        lib.wv_descriptor_table(self._h, name.encode(), buf.ctypes.data)
        return buf

    def coefficient_table(self, name: str) -> np.ndarray:
        t = self.table(name)
        if t.size == 2 * self.M:
            return t.view(np.complex128).reshape(self.Nkl, self.Nj)
        return t.reshape(self.Nkl, self.Nj)

    def coriolis_frequency(self) -> float:
        return float(library().wv_coriolis_frequency(self._h))

    def persistent_bytes(self) -> int:
        return int(library().wv_persistent_bytes(self._h))

    def scratch_bytes(self) -> int:
        """scratchBytes (...Kernel.hpp:70)."""
        return int(library().wv_scratch_bytes(self._h))

    def metrics(self) -> dict:
        """metrics() (...Kernel.cpp:522-537) as a dict of WVKernelMetrics members."""
        m = KernelMetrics()
        _check(library().wv_metrics(self._h, C.byref(m)))
        return {name: getattr(m, name) for name, _ in KernelMetrics._fields_}

    def identifiers(self) -> dict:
        """The identifier accessors (...Kernel.hpp:59-67)."""
        lib = library()
        out = {name: getattr(lib, "wv_" + name)(self._h).decode() for name in IDENTIFIERS}
        out["coefficient_worker_count"] = int(lib.wv_coefficient_worker_count(self._h))
        out["phase_reservation_bytes"] = int(lib.wv_phase_reservation_bytes(self._h))
        return out

    # --------------------------------------------------------- allocation --
    def empty_coefficients(self, n=3):
        torch = _torch()
        return [torch.zeros(self.spectral_shape, dtype=torch.complex128, device=f"cuda:{self.device}") for _ in range(n)]

    def to_device(self, a):
        torch = _torch()
        t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.complex128).reshape(self.spectral_shape))
        return t.to(f"cuda:{self.device}")

    def _dev(self, t, shape, dtype_name):
        torch = _torch()
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError("device entry points take CUDA/HIP torch tensors")
        if tuple(t.shape) != tuple(shape) or not t.is_contiguous() or str(t.dtype) != dtype_name:
            raise WVKernelError(2, f"expected a contiguous {dtype_name} tensor of shape {shape}, got "
                                   f"{tuple(t.shape)} {t.dtype}")
        return C.c_void_p(t.data_ptr())

    def _coef(self, t):
        return self._dev(t, self.spectral_shape, "torch.complex128")

    @staticmethod
    def _stream(stream):
        if stream is None:
            return C.c_void_p(_torch().cuda.current_stream().cuda_stream)
        if hasattr(stream, "cuda_stream"):  # a torch.cuda.Stream
            return C.c_void_p(stream.cuda_stream)
        return C.c_void_p(stream)

    # ---------------------------------------------------------------- flux --
    def nonlinear_flux(self, Ap, Am, A0, t=0.0, t0=0.0, out=None, stream=None):
        """nonlinearFlux (...Kernel.cpp:1140-1146) on device tensors; returns (Fp, Fm, F0)."""
        if out is None:
            out = self.empty_coefficients()
        Fp, Fm, F0 = out
        _check(library().wv_nonlinear_flux_device(self._h, float(t), float(t0), self._coef(Ap), self._coef(Am),
                                                  self._coef(A0), self._coef(Fp), self._coef(Fm), self._coef(F0),
                                                  self._stream(stream)))
        return Fp, Fm, F0

    def nonlinear_flux_host(self, Ap, Am, A0, t=0.0, t0=0.0):
        """Host-memory variant (copies in and out, synchronous)."""
        arrs = [np.ascontiguousarray(np.asarray(a, np.complex128).reshape(self.spectral_shape)) for a in (Ap, Am, A0)]
        outs = [np.zeros(self.spectral_shape, np.complex128) for _ in range(3)]
        _check(library().wv_nonlinear_flux(self._h, float(t), float(t0), *[a.ctypes.data for a in arrs],
                                           *[o.ctypes.data for o in outs]))
        return tuple(outs)

    def transform_wave_vortex_to_uvweta(self, Ap, Am, A0, t=0.0, t0=0.0, out=None, stream=None):
        torch = _torch()
        shape = (4,) + self.spatial_shape
        if out is None:
            out = torch.zeros(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        _check(library().wv_transform_wave_vortex_to_uvweta_device(
            self._h, float(t), float(t0), self._coef(Ap), self._coef(Am), self._coef(A0),
            self._dev(out, shape, "torch.float64"), self._stream(stream)))
        return out

    def transform_wave_vortex_to_uvw(self, Ap, Am, A0, t=0.0, t0=0.0, out=None, stream=None):
        torch = _torch()
        shape = (3,) + self.spatial_shape
        if out is None:
            out = torch.zeros(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        _check(library().wv_transform_wave_vortex_to_uvw_device(
            self._h, float(t), float(t0), self._coef(Ap), self._coef(Am), self._coef(A0),
            self._dev(out, shape, "torch.float64"), self._stream(stream)))
        return out

    def nonlinear_flux_with_advection_fields(self, Ap, Am, A0, t=0.0, t0=0.0, out=None, uvw=None, stream=None):
        """nonlinearFluxWithAdvectionFields (...Kernel.cpp:1148-1158): ((Fp, Fm, F0), uvw [3, Nz, Ny, Nx])."""
        torch = _torch()
        if out is None:
            out = self.empty_coefficients()
        shape = (3,) + self.spatial_shape
        if uvw is None:
            uvw = torch.zeros(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        _check(library().wv_nonlinear_flux_with_advection_fields_device(
            self._h, float(t), float(t0), self._coef(Ap), self._coef(Am), self._coef(A0),
            *[self._coef(o) for o in out], self._dev(uvw, shape, "torch.float64"), self._stream(stream)))
        return tuple(out), uvw

    def nonlinear_flux_using_advection_fields(self, Ap, Am, A0, uvw, t=0.0, t0=0.0, out=None, stream=None):
        """nonlinearFluxUsingAdvectionFields (...Kernel.cpp:1160-1171): the flux
        with the given advecting fields uvw [3, Nz, Ny, Nx] in the products."""
        if out is None:
            out = self.empty_coefficients()
        _check(library().wv_nonlinear_flux_using_advection_fields_device(
            self._h, float(t), float(t0), self._coef(Ap), self._coef(Am), self._coef(A0),
            self._dev(uvw, (3,) + self.spatial_shape, "torch.float64"), *[self._coef(o) for o in out],
            self._stream(stream)))
        return tuple(out)

    def apv(self, Ap, Am, A0, t=0.0, t0=0.0, out=None, stream=None):
        """Available potential vorticity (Operations/APVOperation.m:15-26,
        eta_true = eta for the constant-N background) as an (Nz, Ny, Nx)
        tensor."""
        torch = _torch()
        shape = self.spatial_shape
        if out is None:
            out = torch.zeros(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        _check(library().wv_apv_device(self._h, float(t), float(t0), self._coef(Ap), self._coef(Am),
                                       self._coef(A0), self._dev(out, shape, "torch.float64"),
                                       self._stream(stream)))
        return out

    def transform_state_field_derivatives(self, Ap, Am, A0, field, t=0.0, t0=0.0, out=None, stream=None):
        """transformStateFieldDerivatives (...Kernel.cpp:719-742): (d/dx, d/dy,
        d/dz) of field "u", "v", "w" or "eta" as a (3, Nz, Ny, Nx) tensor."""
        torch = _torch()
        shape = (3,) + self.spatial_shape
        if out is None:
            out = torch.zeros(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        f = DYNAMICAL_FIELDS[field] if isinstance(field, str) else int(field)
        _check(library().wv_transform_state_field_derivatives_device(
            self._h, float(t), float(t0), self._coef(Ap), self._coef(Am), self._coef(A0), f,
            self._dev(out, shape, "torch.float64"), self._stream(stream)))
        return out

    def prepare_scalar_advection(self):
        """prepareScalarAdvection (...Kernel.cpp:1004-1006)."""
        _check(library().wv_prepare_scalar_advection(self._h))

    def advect_f_grid_scalar(self, scalar, advection_fields, should_antialias=True, out=None, stream=None):
        """advectFGridScalar (...Kernel.cpp:1032-1138): -(u, v, w) . grad(scalar)
        of an F-grid scalar [Nz, Ny, Nx] (the reference's [Nx, Ny, Nz]) with the
        velocity bundle [3, Nz, Ny, Nx], antialiased when should_antialias."""
        torch = _torch()
        if out is None:
            out = torch.zeros(self.spatial_shape, dtype=torch.float64, device=f"cuda:{self.device}")
        _check(library().wv_advect_f_grid_scalar_device(
            self._h, self._dev(scalar, self.spatial_shape, "torch.float64"),
            self._dev(advection_fields, (3,) + self.spatial_shape, "torch.float64"), int(bool(should_antialias)),
            self._dev(out, self.spatial_shape, "torch.float64"), self._stream(stream)))
        return out

    def transform_to_wave_vortex(self, fields, t=0.0, t0=0.0, out=None, stream=None):
        """transformUVWEtaToWaveVortex (nonhydrostatic, 4 fields) or
        transformUVEtaToWaveVortex (hydrostatic, 3 fields), ...Kernel.cpp:583-684."""
        hydro = bool(self.config["isHydrostatic"])
        nf = 3 if hydro else 4
        if out is None:
            out = self.empty_coefficients()
        fn = (library().wv_transform_uveta_to_wave_vortex_device if hydro
              else library().wv_transform_uvweta_to_wave_vortex_device)
        _check(fn(self._h, self._dev(fields, (nf,) + self.spatial_shape, "torch.float64"), float(t), float(t0),
                  self._coef(out[0]), self._coef(out[1]), self._coef(out[2]), self._stream(stream)))
        return out

    def transform_to_spatial_domain_with_all_derivatives(self, f_family: bool, Apm, A0, out=None, stream=None):
        """transformToSpatialDomainWith{F,G}AllDerivatives (...Kernel.cpp:819-863):
        (value, d/dx, d/dy, d/dz) as a (4, Nz, Ny, Nx) float64 tensor."""
        torch = _torch()
        shape = (4,) + self.spatial_shape
        if out is None:
            out = torch.zeros(shape, dtype=torch.float64, device=f"cuda:{self.device}")
        fn = (library().wv_transform_to_spatial_domain_with_f_all_derivatives_device if f_family
              else library().wv_transform_to_spatial_domain_with_g_all_derivatives_device)
        _check(fn(self._h, self._coef(Apm), self._coef(A0), self._dev(out, shape, "torch.float64"),
                  self._stream(stream)))
        return out

    # ----------------------------------------------------------------- RK4 --
    def rk4_advance(self, Ap, Am, A0, t, h, nsteps, t0=0.0, stream=None) -> float:
        """nsteps fixed RK4 steps in place (WVFixedStepRK4::step); returns the new t."""
        tt = C.c_double(float(t))
        _check(library().wv_rk4_advance_device(self._h, self._coef(Ap), self._coef(Am), self._coef(A0), C.byref(tt),
                                               float(t0), float(h), int(nsteps), self._stream(stream)))
        return tt.value

    def rk4_advance_to_time(self, Ap, Am, A0, t, final_time, h, t0=0.0, stream=None) -> float:
        """WVFixedStepRK4::advanceToTime (WVRungeKutta.cpp:464-480) in place; returns final_time."""
        tt = C.c_double(float(t))
        _check(library().wv_rk4_advance_to_time_device(self._h, self._coef(Ap), self._coef(Am), self._coef(A0),
                                                       C.byref(tt), float(t0), float(final_time), float(h),
                                                       self._stream(stream)))
        return tt.value

    def rk4_counters(self) -> dict:
        n, h = C.c_uint64(), C.c_double()
        _check(library().wv_rk4_counters(self._h, C.byref(n), C.byref(h)))
        return {"stepCount": int(n.value), "lastStepSize": float(h.value)}

    def inject_launch_failure(self, n: int):
        """Test hook: the n-th next flux-pipeline launch fails (0 disarms)."""
        _check(library().wv_debug_inject_launch_failure(self._h, int(n)))

    def total_energy(self, Ap, Am, A0, stream=None) -> float:
        e = C.c_double()
        _check(library().wv_total_energy_device(self._h, self._coef(Ap), self._coef(Am), self._coef(A0), C.byref(e),
                                                self._stream(stream)))
        return e.value

    # ----------------------------------------------------- instrumentation --
    # --------------------------------------------------------------- adjoint --
    # Cotangents follow <a, b> = Re sum conj(a) b; the gradient of a real cost
    # with respect to a complex coefficient is dJ/dRe + i dJ/dIm.
    def flux_vjp(self, Ap, Am, A0, Lp, Lm, L0, t=0.0, t0=0.0, out=None, stream=None):
        """J(y)^T (Lp, Lm, L0) for the nonlinear flux at time t."""
        if out is None:
            out = self.empty_coefficients()
        _check(library().wv_nonlinear_flux_vjp_device(
            self._h, float(t), float(t0), *[self._coef(a) for a in (Ap, Am, A0, Lp, Lm, L0)],
            *[self._coef(o) for o in out], self._stream(stream)))
        return out

    def transform_wave_vortex_to_uvweta_vjp(self, fields_bar, t=0.0, t0=0.0, out=None, stream=None):
        """Transpose of transform_wave_vortex_to_uvweta: (4, Nz, Ny, Nx) field
        cotangents -> coefficient cotangents."""
        if out is None:
            out = self.empty_coefficients()
        shape = (4,) + self.spatial_shape
        _check(library().wv_transform_wave_vortex_to_uvweta_vjp_device(
            self._h, self._dev(fields_bar, shape, "torch.float64"), float(t), float(t0),
            *[self._coef(o) for o in out], self._stream(stream)))
        return out

    def rk4_step_vjp(self, Ap, Am, A0, t, h, Lp, Lm, L0, t0=0.0, stream=None):
        """In place: (Lp, Lm, L0) <- cotangent before one RK4 step from (Ap, Am, A0) at t."""
        _check(library().wv_rk4_step_vjp_device(
            self._h, *[self._coef(a) for a in (Ap, Am, A0)], float(t), float(t0), float(h),
            *[self._coef(a) for a in (Lp, Lm, L0)], self._stream(stream)))

    def rk4_adjoint(self, Ap, Am, A0, t, h, nsteps, interval, Lp, Lm, L0, t0=0.0, stream=None):
        """In place: (Lp, Lm, L0) <- cotangent of the state at t, given that of
        the state after nsteps RK4 steps; checkpoints every `interval` steps."""
        _check(library().wv_rk4_adjoint_device(
            self._h, *[self._coef(a) for a in (Ap, Am, A0)], float(t), float(t0), float(h), int(nsteps),
            int(interval), *[self._coef(a) for a in (Lp, Lm, L0)], self._stream(stream)))

    def total_energy_into(self, Ap, Am, A0, out, stream=None):
        """Total energy written to the one-element float64 device tensor out,
        allocating nothing (wv_total_energy_no_alloc_device)."""
        _check(library().wv_total_energy_no_alloc_device(self._h, self._coef(Ap), self._coef(Am), self._coef(A0),
                                                         self._dev(out, (1,), "torch.float64"), self._stream(stream)))
        return out

    def set_stage_instrumentation(self, enabled: bool):
        _check(library().wv_set_stage_instrumentation(self._h, int(bool(enabled))))

    def stage_timings(self) -> list:
        lib = library()
        arr = (StageTiming * 32)()
        n = lib.wv_stage_timings(self._h, arr, 32)
        if n < 0:
            _check(9)
        return [dict(name=arr[i].name.decode(), total_ms=arr[i].total_ms, launches=int(arr[i].launches),
                     bytes_per_launch=arr[i].bytes_per_launch) for i in range(min(n, 32))]

    def reset_stage_timings(self):
        library().wv_reset_stage_timings(self._h)


class WVTransformHydrostaticKernel(WVTransformConstantStratificationKernel):
    """Hydrostatic wave-vortex kernel of an arbitrary stratification N^2(z)
    (the reference's WVTransformHydrostatic), built from the vertical
    operators of wvamd.stratification.vertical_operators (or any dict with
    the same MATLAB-named arrays).  Same methods as the constant-N kernel;
    the fields of the inverse transform are (u, v, w, eta) and the forward
    transform takes (u, v, eta)."""

    _OPS = ("z", "N2", "dLnN2", "h", "P", "Q", "PFinv", "PF", "QGinv", "QG")

    def __init__(self, config: dict, operators: dict, precision: str = "fp64", device: int = 0):
        cfg = dict(config)
        if not cfg.get("isHydrostatic"):
            raise ValueError("the arbitrary-stratification transform is hydrostatic: set isHydrostatic")
        Nz, Nj = int(cfg["Nz"]), int(cfg["Nj"])
        shapes = {"z": (Nz,), "N2": (Nz,), "dLnN2": (Nz,), "h": (Nj,), "P": (Nj,), "Q": (Nj,),
                  "PFinv": (Nz, Nj), "PF": (Nj, Nz), "QGinv": (Nz, Nj), "QG": (Nj, Nz)}
        # MATLAB column-major layout
        self._ops_arrays = {}
        for name in self._OPS:
            a = np.asarray(operators[name], dtype=np.float64)
            if a.shape != shapes[name]:
                raise ValueError(f"operator {name} has shape {a.shape}, expected {shapes[name]}")
            self._ops_arrays[name] = np.asfortranarray(a).ravel(order="F").copy()
        self.operators = {k: np.asarray(operators[k]) for k in self._OPS}
        super().__init__(cfg, precision, device)

    def _create(self, lib, config, precision, h):
        ops = VerticalOperators(*[self._ops_arrays[n].ctypes.data for n in self._OPS])
        return lib.wv_create_hydrostatic(C.byref(_config_struct(config)), C.byref(ops), PRECISIONS[precision],
                                         self.device, C.byref(h))
