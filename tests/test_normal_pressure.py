"""The NormalPressureSolver operator (pressuresolver.cpp:50-347), the other
pressure solver behind the reference's projection seam (FluidSimulator.h:214-216).

Pinned against the reference's own NormalPressureSolver (oracle/_ref): matrix
and right-hand side bit for bit, the PCG pressure (same MIC(0) PCG, same
operation order) bit for bit, and the velocities after its applyPressure bit
for bit when given the same pressure.  Without the reference build, the
closed-form checks below still pin the matrix: rows only for fluid cells,
diagonal = in-range non-solid neighbours, off-diagonals -1 between fluid
neighbours, A symmetric with zero row sums away from air, and a divergence-free
projection (the reference's own post-check, pressuresolver.cpp:197-209).
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import FIELD, STAGE, default_params, oracle, ref_available
from scenes import oracle_dam_break, ref_like, small_params

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref/libxfref.so not built")


def _same(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(np.asarray(a, float) - np.asarray(b, float))
        raise AssertionError(f"{what}: {np.count_nonzero(d)} differ, max {d.max():.3e}")


def _state(dims=(14, 11, 9), dx=0.1, walls=True, seed=4):
    """A dam-break state two substeps in, with forces applied (non-trivial u, v, w)."""
    p = small_params(pressure_solver=1)
    o = oracle_dam_break(*dims, dx, (0.5, 0.6, 1.0), seed, p)
    for _ in range(2):
        o.time_step(1.0 / 60)
    o.stage(STAGE["MATERIAL"]); o.stage(STAGE["SAVE"]); o.stage(STAGE["FORCES"], 1.0 / 60)
    if not walls:  # fluid reaching the grid boundary (outside counts as solid)
        m = o.get(FIELD["MATERIAL"])
        m[m == 2] = 1
        o.set(FIELD["MATERIAL"], m)
    return o


def _assemble(o):
    O = oracle()
    N = o.ni * o.nj * o.nk
    A, b = np.empty(7 * N), np.empty(N)
    O.xo_assemble_normal(o.ni, o.nj, o.nk, o.get(FIELD["MATERIAL"]).ravel(), o.get(FIELD["U"]).ravel(),
                         o.get(FIELD["V"]).ravel(), o.get(FIELD["W"]).ravel(), A, b)
    return A.reshape(7, N), b


@needs_ref
@pytest.mark.parametrize("walls", [True, False])
def test_normal_operator_matches_reference(walls):
    O = oracle()
    o = _state(walls=walls)
    r = ref_like(o)
    for f in ("U", "V", "W", "MATERIAL"):
        r.set(FIELD[f], o.get(FIELD[f]))
    ni, nj, nk = o.ni, o.nj, o.nk
    N = ni * nj * nk
    A_r, b_r, p_r = np.empty(7 * N), np.empty(N), np.empty(N)
    assert r.lib.xr_normal_pieces(r.h, A_r, b_r, p_r) == 0
    A_o, b_o = _assemble(o)
    _same(A_o.ravel(), A_r, "matrix")
    _same(b_o, b_r, "rhs")
    # the oracle's PCG with the reference's default solver parameters
    x = np.zeros(N)
    it, res = C.c_int(), C.c_double()
    st = O.xo_pcg(ni, nj, nk, np.ascontiguousarray(A_o.ravel()), b_o, x, 1e-5, 100, 0.97, 0.25, C.byref(it),
                  C.byref(res))
    assert st == 0 and it.value > 3
    _same(x, p_r, "pressure")
    # apply with that pressure: the reference left its result in r's u, v, w
    u, v, w = (np.ascontiguousarray(o.get(FIELD[f]).ravel()) for f in ("U", "V", "W"))
    O.xo_apply_pressure_normal(ni, nj, nk, o.get(FIELD["MATERIAL"]).ravel(), p_r, u, v, w)
    for f, a in (("U", u), ("V", v), ("W", w)):
        _same(a, r.get(FIELD[f]).ravel(), f"apply {f}")


@pytest.mark.parametrize("walls", [True, False])
def test_normal_matrix_closed_form(walls):
    o = _state(walls=walls)
    A, b = _assemble(o)
    m = o.get(FIELD["MATERIAL"]).ravel()
    ni, nj, nk = o.ni, o.nj, o.nk
    N = ni * nj * nk
    off = [-ni * nj, -ni, -1, 0, 1, ni, ni * nj]
    ii = np.arange(N) % ni
    jj = (np.arange(N) // ni) % nj
    kk = np.arange(N) // (ni * nj)
    coords = [kk, jj, ii, None, ii, jj, kk]
    lim = [nk, nj, ni, None, ni, nj, nk]
    nonsolid = np.zeros(N)
    for s in (0, 1, 2, 4, 5, 6):
        lo = s < 3
        inr = coords[s] > 0 if lo else coords[s] < lim[s] - 1
        nb = np.where(inr, m[np.clip(np.arange(N) + off[s], 0, N - 1)], 2)
        fl = (m == 1) & (nb == 1)
        assert np.array_equal(A[s] != 0, fl), f"slot {s} pattern"
        assert np.all(A[s][fl] == -1.0)
        nonsolid += ((m == 1) & (nb != 2))
        # symmetry: A[c, c+off] == A[c+off, c]
        c = np.nonzero(fl)[0]
        assert np.array_equal(A[s][c], A[6 - s][c + off[s]])
    assert np.array_equal(A[3], nonsolid)
    assert np.all(b[m != 1] == 0)


def test_normal_projection_divergence_free():
    """After solve + apply, the reference's own divergence (pressuresolver.cpp:349-420)
    vanishes to the solve's tolerance in every fluid cell."""
    o = _state()
    o.set_params(small_params(pressure_solver=1, pcg_tol=1e-12, pcg_max_iter=500))
    assert o.stage(STAGE["SOLVE"]) == 0
    o.stage(STAGE["APPLY"])
    A, b = _assemble(o)  # b of the projected field = its divergence
    scale = np.abs(o.get(FIELD["U"])).max()
    assert np.abs(b).max() <= 1e-9 * max(scale, 1.0)


def test_normal_golden():
    """The committed fixture of the reference's NormalPressureSolver
    (tests/golden/make_golden.py `normal`): matrix, RHS, pressure and applied
    velocities, reproduced by the oracle bit for bit."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "normal14x11x9.npz")
    g = np.load(path, allow_pickle=False)
    O = oracle()
    ni, nj, nk = (int(x) for x in g["dims"])
    N = ni * nj * nk
    mat = np.ascontiguousarray(g["material"]).ravel()
    A, b = np.empty(7 * N), np.empty(N)
    O.xo_assemble_normal(ni, nj, nk, mat, np.ascontiguousarray(g["u"]).ravel(), np.ascontiguousarray(g["v"]).ravel(),
                         np.ascontiguousarray(g["w"]).ravel(), A, b)
    _same(A, g["A7"], "golden matrix")
    _same(b, g["b"], "golden rhs")
    x = np.zeros(N)
    it, res = C.c_int(), C.c_double()
    assert O.xo_pcg(ni, nj, nk, A, b, x, 1e-5, 100, 0.97, 0.25, C.byref(it), C.byref(res)) == 0
    _same(x, g["pressure"], "golden pressure")
    u, v, w = (np.ascontiguousarray(g[f]).ravel().copy() for f in ("u", "v", "w"))
    O.xo_apply_pressure_normal(ni, nj, nk, mat, np.ascontiguousarray(g["pressure"]), u, v, w)
    _same(u, g["u_out"].ravel(), "golden apply u")
    _same(v, g["v_out"].ravel(), "golden apply v")
    _same(w, g["w_out"].ravel(), "golden apply w")
