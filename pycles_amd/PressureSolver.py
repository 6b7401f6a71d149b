"""Anelastic pressure projection operator (device).

Mirrors reference ``PressureSolver.pyx``: ``initialize(namelist, Gr, RS, DV,
PM)`` registers ``dynamic_pressure`` and ``divergence``; ``update(Gr, RS, DV,
PV, PM)`` (note the DV-before-PV order of the reference signature) removes
the level mean of w, solves for p and corrects u, v, w.  The FFT/tridiagonal
plan (``PressureFFTParallel.initialize``) is created on the first update, when
the field precision is known.
"""
from __future__ import annotations

import numpy as np

from . import _capi
from . import _device as D


class PressureSolver:
    def __init__(self):
        self.plan = None
        self.write_divergence = True

    def initialize(self, namelist, Gr, RS, DV, PM=None):
        DV.add_variables("dynamic_pressure", "Pa", "p", "dynamic pressure", "sym", PM)
        DV.add_variables("divergence", "1/s", "d", "3d divergence", "sym", PM)
        self.Gr = Gr

    def _plan(self, Gr, RS, precision):
        if self.plan is None:
            rho0 = np.ascontiguousarray(RS.rho0, dtype=np.float64)
            rho0h = np.ascontiguousarray(RS.rho0_half, dtype=np.float64)
            self.plan = _capi.PressurePlan(D.dims_of(Gr), rho0.ctypes.data, rho0h.ctypes.data, precision)
        return self.plan

    def update(self, Gr, RS, DV, PV, PM=None, cfl_out=None, dt=0.0):
        plan = self._plan(Gr, RS, PV.precision)
        ref = D.device_ref(RS)
        n = PV.velocity_names_directional
        plan.update(ref.c, PV.ptr(n[0]), PV.ptr(n[1]), PV.ptr(n[2]), DV.ptr("dynamic_pressure"),
                    DV.ptr("divergence") if self.write_divergence else None, dt,
                    None if cfl_out is None else cfl_out.data_ptr(), D.stream())
