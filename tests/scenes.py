"""Scenario builders shared by the parity tests: the same dam break on the
oracle, the reference harness and the GPU library, from one seeded particle set.

Configs follow the dam-break workloads: 1-cell solid walls, a fluid block
addCubicFluid((0,0,0), frac*L), 8 particles per fluid cell (one per 2x2x2
sub-cell, SplitMix64 jitter), gravity -9.81 in y.
"""
import numpy as np

from oracle_lib import FIELD, OracleSim, RefSim, dam_break_scene, default_params

GRAVITY = -9.81


def oracle_dam_break(ni, nj, nk, dx, frac, seed, params=None, mode=1):
    o = OracleSim(ni, nj, nk, dx, params)
    lu, rd = dam_break_scene(ni, nj, nk, dx, *frac)
    o.lib.xo_set_solid_boundary(o.h)
    o.lib.xo_add_cubic_fluid(o.h, np.array(lu), np.array(rd))
    o.lib.xo_seed_particles(o.h, mode, seed)
    o.lib.xo_add_constant_force(o.h, 0.0, GRAVITY, 0.0)
    return o


def ref_like(o):
    """Reference harness with the oracle's scene, particles and parameters."""
    r = RefSim(o.ni, o.nj, o.nk, o.dx, o.params)
    r.set(FIELD["MATERIAL"], o.get(FIELD["MATERIAL"]))
    r.lib.xr_init(r.h)
    r.lib.xr_add_constant_force(r.h, 0.0, GRAVITY, 0.0)
    P, V = o.particles()
    r.set_particles(P, V)
    return r


def gpu_like(o):
    """GPU simulator with the oracle's scene, particles and parameters."""
    import xfluid_amd as xf
    p = xf.default_params()
    for name, _ in p._fields_:
        if name != "reserved":
            setattr(p, name, getattr(o.params, name))
    g = xf.FluidSimulator(o.ni, o.nj, o.nk, p)
    g.setCellWidth(o.dx)
    g.set_field(xf.FIELD["MATERIAL"], o.get(FIELD["MATERIAL"]))
    g.addConstantForce(0.0, GRAVITY, 0.0)
    P, V = o.particles()
    g.set_particles(P, V)
    return g


SYNC_FIELDS = ("U", "V", "W", "USAVED", "VSAVED", "WSAVED", "MATERIAL")


def copy_state(o, g, pressure=False):
    """Copy the oracle's full state into the GPU simulator."""
    for f in SYNC_FIELDS:
        g.set_field(FIELD[f], o.get(FIELD[f]))
    if pressure:
        g.set_field(FIELD["PRESSURE"], o.get(FIELD["PRESSURE"]))
    P, V = o.particles()
    g.set_particles(P, V)


def small_params(**kw):
    base = dict(cfl=5.0, pcg_tol=1e-6, pcg_max_iter=200)
    base.update(kw)
    return default_params(**base)
