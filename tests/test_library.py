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
