"""Write tests/golden/runtime_driver.npz: the reference PortableRuntime
(forcing engine, integration system, WVFixedStepRK4, WVAdaptiveRK23) over the
REFERENCE kernel, from oracle/_ref/runtime_driver_ref (built in place by
`make -C oracle runtime-ref` where the reference tree exists).  The GPU test
runs the same driver over the MI355X kernel and compares.

    python tests/golden/make_runtime_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests", "runtime"))
from driver_io import read_records  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "runtime_driver_ref")

if __name__ == "__main__":
    if not os.path.exists(DRIVER):
        sys.exit(f"{DRIVER} is missing: run `make -C oracle runtime-ref` where the reference tree exists")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out.bin")
        subprocess.run([DRIVER, path], check=True)
        rec = read_records(path)
    out = os.path.join(HERE, "runtime_driver.npz")
    np.savez_compressed(out, **rec)
    print("wrote", os.path.relpath(out, ROOT), f"{os.path.getsize(out) / 1024:.1f} KiB", sorted(rec))
