"""SSP-RK3 time stepping (reference ``TimeStepping.pyx``).

``initialize`` reads ``time_stepping`` (ts_type, dt_initial, dt_max, t,
cfl_limit, t_max) like the reference; ``update(Gr, PV, Pa)`` applies the
Shu-Osher stage (``update_third``, :131-151) to every prognostic field,
zeroes the tendencies and fills the z ghost levels; ``adjust_timestep``
computes the CFL maximum on the last stage and sets the next dt
(:94-108, :222-266).  Only ts_type 3 is on the device path.
"""
from __future__ import annotations

import torch

from . import _capi
from . import _device as D


class TimeStepping:
    def initialize(self, namelist, PV, Pa=None):
        ts = namelist["time_stepping"]
        self.ts_type = int(ts.get("ts_type", 3))
        if self.ts_type != 3:
            raise ValueError(f"ts_type {self.ts_type} not supported on the device path (3 only)")
        self.dt = float(ts.get("dt_initial", 1.0))
        self.dt_max = float(ts.get("dt_max", 10.0))
        self.t = float(ts.get("t", 0.0))
        self.cfl_limit = float(ts.get("cfl_limit", 0.7))
        self.t_max = float(ts["t_max"])
        self.rk_step = 0
        self.n_rk_steps = 3
        self.cfl_max = 0.0
        self.PV = PV
        self.cfl_dev = torch.zeros(1, dtype=torch.float64, device=D.require_gpu())
        self._cfl_ready = False

    def copy_ptr(self, PV, name) -> int:
        if getattr(self, "copies", None) is None:
            self.copies = torch.zeros_like(PV.values)
        return self.copies[PV.get_nv(name)].data_ptr()

    def advance(self):
        """The time bookkeeping of update(), for a stage whose RK update ran
        fused into the advection (pycles_advection_rk3)."""
        if self.rk_step == self.n_rk_steps - 1:
            self.t += self.dt

    def update(self, Gr, PV, Pa=None, zero_tendencies=True):
        if getattr(self, "copies", None) is None:
            self.copies = torch.zeros_like(PV.values)
            self.dims = D.dims_of(Gr)
        names = PV.index_name
        _capi.rk3_stage(self.dims, [PV.ptr(n) for n in names], [PV.tptr(n) for n in names],
                        [self.copies[PV.get_nv(n)].data_ptr() for n in names],
                        [PV.bc_type[PV.get_nv(n)] for n in names], self.rk_step, self.dt,
                        PV.precision, D.stream(), 1 if zero_tendencies else 0)
        if self.rk_step == self.n_rk_steps - 1:
            self.t += self.dt

    def cfl_buffer(self):
        """Device slot the pressure correction writes the CFL maximum into."""
        self._cfl_ready = True
        return self.cfl_dev

    def compute_cfl_max(self, Gr, PV, DV=None, Pa=None):
        if not self._cfl_ready:
            n = PV.velocity_names_directional
            _capi.cfl_max(D.dims_of(Gr), PV.ptr(n[0]), PV.ptr(n[1]), PV.ptr(n[2]), self.dt,
                          self.cfl_dev.data_ptr(), PV.precision, D.stream())
        self._cfl_ready = False
        # the library's read raises (PYCLES_ERR_NONFINITE) when a velocity is
        # NaN or infinite, which the device maximum carries, instead of
        # setting a NaN time step; the reference stops on a negative maximum
        self.accept_cfl(_capi.cfl_read(self.cfl_dev.data_ptr(), D.stream()))

    def accept_cfl(self, value):
        """Take the step's CFL maximum (TimeStepping.pyx:255-263)."""
        self.cfl_max = float(value)
        if not self.cfl_max >= 0.0:
            raise RuntimeError(f"CFL_MAX = {self.cfl_max} killing simulation!")

    def restart(self, Re):
        """TimeStepping.restart (TimeStepping.pyx:268-273)."""
        Re.restart_data["TS"] = {"t": self.t, "dt": self.dt}

    def cfl_time_step(self):
        return min(self.dt_max, self.cfl_limit / (self.cfl_max / self.dt))

    def adjust_timestep(self, Gr, PV, DV=None, Pa=None):
        if self.rk_step == self.n_rk_steps - 1:
            self.compute_cfl_max(Gr, PV, DV, Pa)
            self.dt = self.cfl_time_step()
            if self.t + self.dt > self.t_max:
                self.dt = self.t_max - self.t
            if self.dt < 0.0:
                raise RuntimeError(f"dt = {self.dt} killing simulation!")
