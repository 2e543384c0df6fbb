"""Code generation of the pipelined sweeps (CPU test: hipcc cross-compiles).

The record wave of xf_sweep_wave.hip publishes each step with an LDS store
written as inline assembly, after an explicit `s_waitcnt vmcnt(N)` that leaves
the later steps' LDS-DMA requests in flight.  If the compiler puts its own
full `s_waitcnt vmcnt(0)` in front of such a store (it does when a register the
store's value lands in may still be the target of an earlier load, e.g. a load
on a rarely taken path inside the loop), every step drains the requests in
flight and the sweep runs at about half speed with the same results.  This
test compiles the file for gfx950 and checks every publishing store of every
sweep kernel.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_record_publish_has_no_full_vmcnt_wait(tmp_path):
    src = os.path.join(ROOT, "xfluid_amd", "csrc", "xf_sweep_wave.hip")
    out = tmp_path / "sweep.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-w", f"-I{os.path.join(ROOT, 'include')}", "-S", "--cuda-device-only", src,
           "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    lines = out.read_text().split("\n")
    kernel, publishes, bad = None, 0, []
    for i, line in enumerate(lines):
        m = re.match(r"^(_Z\w*k_sweep_w8\w*):", line)
        if m:
            kernel = m.group(1)
        elif re.match(r"^_Z\w*:", line):
            kernel = None
        if kernel and "ds_write_b32" in line and "ASMSTART" in lines[i - 1]:
            publishes += 1
            before = [l.strip() for l in lines[max(0, i - 6):i]]
            if any(l.startswith("s_waitcnt") and "vmcnt(0)" in l for l in before):
                bad.append((kernel, i))
    assert publishes > 20, "no publishing stores found: the sweep kernels changed shape"
    assert not bad, f"full vmcnt wait before a publishing LDS store: {bad[:4]}"
