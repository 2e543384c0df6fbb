"""The allocation-free P2G adjoint (xf_sim_p2g_adjoint_noalloc) against the
allocating one (xf_sim_p2g_adjoint, device pointers), through the C ABI.

Two calls in a row share the GPU's static scratch: the first on a 24x20x16
grid with 6,000 particles, the second on a smaller 14x12x10 grid with 700
particles (so sums left over from the first call would corrupt the second if
the scratch were not cleared).  The weight and value sums use atomics in both
entry points, so the sums' order (hence the last bits) differs between runs:
gvel and gpos agree to 1e-12 relative.  A grid over the scratch's size is
rejected with an error before any work.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    import torch
    import xfluid_amd as xf
    if xf.device_count() < 1 or not torch.cuda.is_available():
        pytest.fail("no HIP device")
    return xf, torch


def _case(xf, torch, dims, n, seed, kernel):
    ni, nj, nk = dims
    dx = 0.1
    rng = np.random.default_rng(seed)
    s = xf.FluidSimulator(ni, nj, nk, xf.default_params(p2g_kernel=kernel))
    s.setCellWidth(dx)
    lo, hi = 0.5 * dx, np.array([ni, nj, nk]) * dx - 0.5 * dx
    pos = rng.uniform(lo, hi, (n, 3))
    vel = rng.standard_normal((n, 3))
    shapes = [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]
    gf = [rng.standard_normal(sh) for sh in shapes]
    gpos0 = rng.standard_normal((n, 3))
    D = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda").contiguous()  # noqa: E731
    out = []
    for fn in ("p2g_adjoint", "p2g_adjoint_noalloc"):
        gvel = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        gpos = D(gpos0)
        torch.cuda.synchronize()
        getattr(s, fn)(D(pos), D(vel), *(D(a) for a in gf), gvel, gpos)
        out.append((gvel.cpu().numpy(), gpos.cpu().numpy()))
    s.close()
    return out


def _close(a, b, what):
    err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
    assert err <= 1e-12, f"{what}: relative difference {err:.3e}"


@pytest.mark.parametrize("kernel", [0, 1])
def test_noalloc_matches_allocating_entry_point(env, kernel):
    xf, torch = env
    for dims, n, seed in (((24, 20, 16), 6000, 1), ((14, 12, 10), 700, 2)):
        say(f"[noalloc] p2g adjoint kernel={kernel} grid={dims} n={n}")
        (gv_a, gp_a), (gv_b, gp_b) = _case(xf, torch, dims, n, seed, kernel)
        assert np.abs(gv_a).max() > 0 and np.abs(gp_a - 0).max() > 0
        _close(gv_b, gv_a, f"gvel {dims}")
        _close(gp_b, gp_a, f"gpos {dims}")


def test_noalloc_rejects_oversized_grid(env):
    xf, torch = env
    say("[noalloc] oversized grid is rejected")
    s = xf.FluidSimulator(132, 128, 128, xf.default_params())  # 133*128*128 u faces > the scratch holds
    s.setCellWidth(0.01)
    t = lambda *sh: torch.zeros(sh, dtype=torch.float64, device="cuda")  # noqa: E731
    with pytest.raises(xf.XfError, match="static scratch"):
        s.p2g_adjoint_noalloc(t(4, 3), t(4, 3), t(128, 128, 133), t(128, 129, 132), t(129, 128, 132), t(4, 3),
                              t(4, 3))
    s.close()
