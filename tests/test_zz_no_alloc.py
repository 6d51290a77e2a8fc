"""The allocation-free total energy (wv_total_energy_no_alloc_device: partial
sums in the library's static device scratch, result to a device pointer)
against wv_total_energy_device, through the C-ABI: two calls in a row, the
second with a smaller input, then the larger again.  Same kernels and the same
partial-sum layout, so the results are equal bit for bit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(k, seed):
    rng = np.random.default_rng(seed)
    return [k.to_device(1e-2 * (rng.standard_normal(k.spectral_shape) + 1j * rng.standard_normal(k.spectral_shape)))
            for _ in range(3)]


def test_no_alloc_energy_matches_allocating_entry_point(torch_gpu):
    import wvamd
    lib = wvamd.library()
    big = wvamd.WVTransformConstantStratificationKernel(wvamd.make_config(64, 64, 33, 21), "fp64")
    small = wvamd.WVTransformConstantStratificationKernel(wvamd.make_config(16, 16, 9, 5), "mixed")
    assert small.M < big.M
    stream = torch_gpu.cuda.Stream()
    for k, seed in ((big, 1), (small, 2), (big, 3)):
        print(f"[gpu] no-alloc energy M={k.M} seed={seed}", flush=True)
        A = _state(k, seed)
        ptrs = [ctypes.c_void_p(a.data_ptr()) for a in A]
        ref = ctypes.c_double()
        assert lib.wv_total_energy_device(k._h, *ptrs, ctypes.byref(ref), None) == 0
        out = torch_gpu.full((1,), float("nan"), dtype=torch_gpu.float64, device="cuda")
        torch_gpu.cuda.synchronize()
        rc = lib.wv_total_energy_no_alloc_device(k._h, *ptrs, ctypes.c_void_p(out.data_ptr()),
                                                 ctypes.c_void_p(stream.cuda_stream))
        assert rc == 0, lib.wv_last_error()
        stream.synchronize()
        assert out.item() == ref.value, (out.item(), ref.value)
        assert k.total_energy_into(*A, out).item() == ref.value
    ptrs = [ctypes.c_void_p(a.data_ptr()) for a in _state(small, 4)]
    assert lib.wv_total_energy_no_alloc_device(small._h, *ptrs, None, None) == 3  # invalidPointer
    assert lib.wv_total_energy_no_alloc_device(small._h, None, ptrs[1], ptrs[2], ptrs[0], None) == 3
