"""The drop-in C++ adapter builds against the contract's own headers and
links with the reference's PortableRuntime callers, unchanged.

Runs where the reference tree is mounted (the build reads the contract
headers and runtime sources where they lie); the GPU box runs the prebuilt
driver instead (tests/test_gpu_runtime.py)."""
import os
import subprocess

import pytest
from conftest import ROOT

REF = "/root/reference"
ADAPTER = os.path.join(ROOT, "wave-vortex-model_amd", "adapter")


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "PortableRuntime")), reason="reference tree not mounted")
def test_adapter_and_runtime_driver_build_and_link():
    r = subprocess.run(["make", "-j8", "-C", ADAPTER], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    exe = os.path.join(ROOT, "oracle", "_ref", "hip_runtime", "runtime_driver_hip")
    assert os.access(exe, os.X_OK)
    syms = subprocess.run(["nm", "-C", "--defined-only", exe], capture_output=True, text=True).stdout
    # the contract class is defined by the adapter (not the reference's CPU kernel)
    for member in ("WVTransformConstantStratificationKernel::create", "WVTransformConstantStratificationKernel::nonlinearFlux",
                   "WVTransformConstantStratificationKernel::advectFGridScalar", "WVFixedStepRK4::advanceToTime",
                   "WVAdaptiveRK23::advanceToTime"):
        assert member in syms, member
    assert "projectSingleFluxTargetImpl" not in syms  # no reference kernel internals linked in
    libs = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libwvamd.so" in libs
