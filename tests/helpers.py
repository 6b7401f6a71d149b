"""Shared set-up for the parity tests: one seeded state fed to both the
HIP path and the CPU oracle."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from pycles_amd import cases
from pycles_amd.Grid import Grid
from pycles_amd.ReferenceState import ReferenceState
from pycles_amd.ThermodynamicsSA import ThermodynamicsFactory


def host_case(nx, ny, nz, gw=3, moist=True, dx=50.0, dz=20.0, seed=2, sigma=0.3, order=(5, 5), dt=1.0):
    nl = cases.namelist(nx, ny, nz, gw, dx, dx, dz, moist, order=order, dt=dt, seed=seed, sigma=sigma)
    Gr = Grid(nl)
    RS = ReferenceState(Gr)
    Th = ThermodynamicsFactory(nl)
    fields = (cases.init_moist if moist else cases.init_dry)(Gr, RS, Th, seed=seed, sigma=sigma)
    return nl, Gr, RS, Th, fields


def oracle_sim(Gr, RS, Th, fields, moist, order=(5, 5), dt=1.0, max_iter=0, sgs=None, bomex=None, **kw):
    d = Gr.dims
    dims = O.make_dims(d.n[0], d.n[1], d.n[2], d.gw, *d.dx)
    ref = O.Ref(RS.p0, RS.p0_half, RS.alpha0, RS.alpha0_half, RS.rho0, RS.rho0_half)
    lt = O.Lookup(Th.CC.T, Th.CC.pv) if moist else None
    sim = O.OracleSimulation(dims, ref, moist, lt, scalar_order=order[0], momentum_order=order[1],
                             dt=dt, max_iter=max_iter, sgs=sgs, bomex=bomex, **kw)
    for n in sim.pv_names:
        sim.pv(n)[...] = Gr.from_device_layout(fields[n])
    return sim


def rel_err(a, b):
    """max|a-b| / max|b| (normwise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    if scale == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)


def ulp_perturbed(fields, name="s"):
    """The state with every nonzero value of one field moved up by one FP64
    ulp: the smallest perturbation a rounding difference makes."""
    out = dict(fields)
    f = np.asarray(fields[name])
    out[name] = np.where(f != 0.0, np.nextafter(f, np.inf), 0.0)
    return out


def floor_tolerance(err_floor, tol=1e-12, factor=10.0):
    """Per-field tolerance of a multi-step comparison: tol, or `factor` times
    how far the oracle moves from a one-ulp perturbation of its own start
    over the same steps (where that is larger, rounding differences between
    two correct evaluations are amplified by the dynamics: near-zero
    arguments of the Smagorinsky stability function's square root, secant
    stops, upwind switches)."""
    return {n: max(tol, factor * v) for n, v in err_floor.items()}
