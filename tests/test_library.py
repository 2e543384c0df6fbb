"""The C-ABI library loads and exports every function include/xfluid.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "xfluid.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(xf_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("xf_sim_create", "xf_sim_time_step", "xf_sim_frame_step", "xf_project", "xf_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "xfluid_amd", "libxfluid_hip.so"))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_params_default_without_gpu():
    import xfluid_amd as xf
    p = xf.default_params()
    assert p.flip_ratio == 0.95 and p.pcg_max_iter == 100 and p.mic_tau == 0.97 and p.iso == 0.4
    assert p.pressure_solver == xf.PRESSURE_VARIATIONAL


def test_no_device_fails_loudly():
    """Without a HIP device the library refuses to create a simulator (no CPU fallback)."""
    import pytest
    import xfluid_amd as xf
    if xf.device_count() > 0:
        pytest.skip("a HIP device is present")
    lib = ctypes.CDLL(os.path.join(ROOT, "xfluid_amd", "libxfluid_hip.so"))
    lib.xf_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    assert lib.xf_sim_create(8, 8, 8, ctypes.byref(h)) == -3
    assert b"no HIP device" in lib.xf_last_error()
    with pytest.raises(Exception):
        xf.FluidSimulator(8, 8, 8)


def test_bench_refuses_several_gpus():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
