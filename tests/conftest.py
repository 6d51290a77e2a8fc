import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "wave-vortex-model_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def rel_err(a, b):
    """max|a-b| / max|b| -- the reference's relative error (wvCompiledBackendBuild.m:458-460)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))


def rel_err_joint(mine, refs):
    """Largest |a-b| over a tuple of arrays divided by the largest |b| over all of
    them: the flux outputs (Fp, Fm, F0) share one physical scale, and a single
    array can be pure round-off (e.g. Fp, Fm when only j = 0 is retained)."""
    den = max(float(np.max(np.abs(np.asarray(b)))) for b in refs)
    num = max(float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) for a, b in zip(mine, refs))
    return num / den if den > 0 else num


def golden_files(prefix="golden_"):
    return sorted(glob.glob(os.path.join(GOLDEN, prefix + "*.npz")))


def load_golden(path):
    return dict(np.load(path, allow_pickle=False))


def config_from_golden(g):
    keys = ["Nx", "Ny", "Nz", "Nj", "Lx", "Ly", "Lz", "N0", "rho0", "g", "planetaryRadius", "rotationRate",
            "latitude", "isHydrostatic", "shouldAntialias"]
    cfg = {}
    for k in keys:
        v = float(g["cfg_" + k])
        if k in ("Nx", "Ny", "Nz", "Nj"):
            cfg[k] = int(v)
        elif k in ("isHydrostatic", "shouldAntialias"):
            cfg[k] = bool(v)
        else:
            cfg[k] = v
    return cfg


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def torch_gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    return torch
