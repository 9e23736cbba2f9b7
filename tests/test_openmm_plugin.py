"""The C++ OpenMM kernel classes (lib/libgbi_openmm.so) on the CPU: plugin entry points,
and the driver's error path without a device (no compute)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "openmmcustomgbintegralforce_amd")
PLUGIN = os.path.join(PKG, "lib", "libgbi_openmm.so")
DRIVER = os.path.join(PKG, "lib", "gbi_openmm_driver")


def write_input(path, model, itype, rule, pos, q, R, method=None, cutoff=1.0, box=(2.0, 2.0, 2.0), exclusions=None):
    with open(path, "wb") as fh:
        np.array([model, itype, rule, len(q)], dtype=np.int32).tofile(fh)
        np.ascontiguousarray(pos, dtype=np.float64).tofile(fh)
        np.ascontiguousarray(q, dtype=np.float64).tofile(fh)
        np.ascontiguousarray(R, dtype=np.float64).tofile(fh)
        if method is not None:
            np.array([method], dtype=np.int32).tofile(fh)
            np.array([cutoff, *box], dtype=np.float64).tofile(fh)
            if exclusions is not None:
                ex = np.asarray(exclusions, dtype=np.int32).reshape(-1, 2)
                np.array([ex.shape[0]], dtype=np.int32).tofile(fh)
                ex.tofile(fh)


def test_plugin_exports_openmm_entry_points():
    lib = ctypes.CDLL(PLUGIN)
    for name in ("registerPlatforms", "registerKernelFactories", "registerCharmmHipKernelFactories"):
        assert hasattr(lib, name), name


def test_driver_reports_missing_device(tmp_path):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present: the GPU test covers the driver")
    inp, out = tmp_path / "in.bin", tmp_path / "out.txt"
    write_input(inp, 0, 0, 0, np.zeros((2, 3)) + [[0, 0, 0], [0.3, 0, 0]], [0.5, -0.5], [0.2, 0.2])
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "no HIP device" in r.stderr


def test_canonical_detection_by_integral_names():
    """The canonical Born expression names its integrals I4 and I7: a force whose integral
    named I4 has order 7 (and I7 order 4) is not the canonical one and goes to the generated
    kernels, as does one asking for energy parameter derivatives; declaring I7 before I4 is
    still canonical."""
    r = subprocess.run([DRIVER, "--canonical-check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = dict(ln.split() for ln in r.stdout.splitlines())
    assert got == {"canonical": "1", "swapped_names": "0", "reordered": "1", "with_param_derivative": "0"}
