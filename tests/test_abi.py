"""CPU checks of the product boundary: the C-ABI library loads and exports
every entry point include/wv_hip.h declares; the Python bindings cover them;
host-only helpers (synthetic flows) behave.  No GPU compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest
from conftest import ROOT

import wvamd
from wvamd import flows

HEADER = os.path.join(ROOT, "include", "wv_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "wv_create" in syms and "wv_nonlinear_flux_device" in syms and "wv_rk4_advance_device" in syms


def test_library_exports_every_declared_symbol():
    lib = wvamd.library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_bindings_cover_header():
    assert sorted(wvamd.EXPORTED_SYMBOLS) == declared_symbols()


def test_contract_version_and_error_string():
    lib = wvamd.library()
    assert lib.wv_contract_version() == wvamd.CONTRACT_VERSION == 4
    assert isinstance(lib.wv_last_error(), bytes)


def test_create_rejects_bad_arguments_without_gpu_work():
    lib = wvamd.library()
    h = ctypes.c_void_p()
    assert lib.wv_create(None, 0, 0, ctypes.byref(h)) == 3  # invalidPointer
    cfg = wvamd._config_struct(wvamd.make_config(8, 8, 5))
    assert lib.wv_create(ctypes.byref(cfg), 7, 0, ctypes.byref(h)) == 1  # invalidConfiguration (precision)


def test_status_names_match_reference_codes():
    # WVKernelStatusCode order, WVKernelTypes.hpp:22-35
    assert wvamd.STATUS_NAMES[0] == "success" and wvamd.STATUS_NAMES[-1] == "reentrantExecution"
    assert len(wvamd.STATUS_NAMES) == 12


def test_default_nj():
    assert wvamd.default_nj(33) == 21 and wvamd.default_nj(129) == 85 and wvamd.default_nj(65) == 42


def test_random_flow_is_deterministic_and_split():
    rng = np.random.default_rng(0)
    Nkl, Nj = 40, 6
    Kh = np.sort(rng.uniform(0, 1, Nkl))
    Kh[0] = 0.0
    te = rng.uniform(0.5, 2.0, (Nkl, Nj))
    a = flows.random_flow_shape(Kh, Nj, te, te, seed=9, jmax=4)
    b = flows.random_flow_shape(Kh, Nj, te, te, seed=9, jmax=4)
    c = flows.random_flow_shape(Kh, Nj, te, te, seed=10, jmax=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    Ap, Am, A0 = a
    assert np.all(Ap[:, 5:] == 0) and np.all(A0[:, 5:] == 0)
    assert np.array_equal(Am[0], np.conj(Ap[0])) and np.all(A0[0].imag == 0)
    ew = np.sum(te * (np.abs(Ap) ** 2 + np.abs(Am) ** 2))
    eg = np.sum(te * np.abs(A0) ** 2)
    assert abs(ew - 1) < 1e-12 and abs(eg - 1) < 1e-12


def test_uniform_counter_rng_independent_of_length():
    assert np.array_equal(flows.uniform(5, 1, 10), flows.uniform(5, 1, 1000)[:10])
    u = flows.uniform(5, 1, 100000)
    assert 0 <= u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.01


def test_new_entry_points_reject_null_kernel_without_gpu_work():
    lib = wvamd.library()
    assert lib.wv_scratch_bytes(None) == 0 and lib.wv_phase_reservation_bytes(None) == 0
    m = wvamd.KernelMetrics()
    assert lib.wv_metrics(None, ctypes.byref(m)) == 3
    for name in wvamd.IDENTIFIERS:
        assert getattr(lib, "wv_" + name)(None) is None
    assert lib.wv_prepare_scalar_advection(None) == 3
    assert lib.wv_nonlinear_flux_using_advection_fields_device(None, 0.0, 0.0, *([None] * 8)) == 3
    assert lib.wv_transform_state_field_derivatives_device(None, 0.0, 0.0, None, None, None, 0, None, None) == 3
    t = ctypes.c_double(0.0)
    assert lib.wv_rk4_advance_to_time_device(None, None, None, None, ctypes.byref(t), 0.0, 1.0, 0.1, None) == 3
    assert lib.wv_rk4_counters(None, None, None) == 3
    assert lib.wv_debug_inject_launch_failure(None, 1) == 3
    e = ctypes.c_double()
    assert lib.wv_total_energy_device(None, None, None, None, ctypes.byref(e), None) == 3
    assert lib.wv_apv_device(None, 0.0, 0.0, None, None, None, None, None) == 3


def test_metrics_struct_mirrors_reference_members():
    # WVKernelMetrics (...Kernel.hpp:12-45): 18 counters then 14 timings
    names = [n for n, _ in wvamd.KernelMetrics._fields_]
    assert names[0] == "descriptorBytes" and names[8] == "planCount" and names[17] == "bytesCopied"
    assert names[18] == "phaseSeconds" and names[-1] == "scalarAntialiasSeconds" and len(names) == 32
    assert ctypes.sizeof(wvamd.KernelMetrics) == 32 * 8
