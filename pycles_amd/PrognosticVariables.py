"""Prognostic and diagnostic state in HBM.

Mirrors reference ``PrognosticVariables.pyx`` / ``DiagnosticVariables.pyx``:
variables are registered by name with a boundary type ("sym" -> 1.0,
"asym" -> -1.0) and, for PV, a variable type (velocity / scalar); the values
and tendencies live in one flat device buffer with variable v at offset
``v * npg`` (``get_varshift``).  The difference is the per-variable layout:
``[k][j][i]`` over ``nzg`` levels and the interior columns (see ``Grid``),
in the run's precision.

In FP32 runs the scalars s and qt are stored as deviations from an FP64
level profile (``base``), taken as the horizontal mean of the initial field;
``get_variable_array`` and ``set_variable_array`` add and remove it.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _device as D


class _State:
    def _register(self, name, bc_type, Pa=None):
        if bc_type == "sym":
            bc = 1.0
        elif bc_type == "asym":
            bc = -1.0
        else:
            raise ValueError(f"Not a valid bc_type {bc_type!r}")
        self.name_index[name] = self.nv
        self.index_name.append(name)
        self.bc_type.append(bc)
        self.nv = len(self.name_index)

    def get_nv(self, name):
        return self.name_index[name]

    def get_varshift(self, Gr, name):
        return self.name_index[name] * Gr.dims.npdev

    def field(self, name) -> torch.Tensor:
        return self.values[self.name_index[name]]

    def ptr(self, name) -> int:
        return self.field(name).data_ptr()


class PrognosticVariables(_State):
    def __init__(self, Gr):
        self.name_index = {}
        self.index_name = []
        self.units = {}
        self.bc_type = []
        self.var_type = []
        self.nv = 0
        self.nv_scalars = 0
        self.nv_velocities = 0
        self.velocity_directions = [0, 0, 0]
        self.velocity_names_directional = ["", "", ""]
        self.base = {}
        self.precision = D.F64

    def add_variable(self, name, units, nice_name, desc, bc_type, var_type, Pa=None):
        self._register(name, bc_type, Pa)
        self.units[name] = units
        if var_type == "velocity":
            self.var_type.append(0)
            self.nv_velocities += 1
        elif var_type == "scalar":
            self.var_type.append(1)
            self.nv_scalars += 1
        else:
            raise ValueError(f"Not a valid var_type {var_type!r}")

    def set_velocity_direction(self, name, direction, Pa=None):
        self.velocity_directions[direction] = self.get_nv(name)
        self.velocity_names_directional[direction] = name

    def initialize(self, Gr, NS=None, Pa=None, precision=D.F64):
        self.precision = D.precision_code(precision)
        dev = D.require_gpu()
        d = Gr.dims
        shape = (self.nv, d.nzg, d.n[1], d.n[0])
        dt = D.torch_dtype(self.precision)
        self.values = torch.zeros(shape, dtype=dt, device=dev)
        self.tendencies = torch.zeros(shape, dtype=dt, device=dev)
        self.Gr = Gr

    def ptr_next(self, name) -> int:
        """The second value array, written by the fused advection + RK3 stage."""
        if getattr(self, "values_next", None) is None:
            self.values_next = torch.zeros_like(self.values)
        return self.values_next[self.name_index[name]].data_ptr()

    def swap_values(self):
        """Make the array the fused stage wrote the current one."""
        self.values, self.values_next = self.values_next, self.values

    def rk_arrays(self, rk_step):
        """Output and copy value arrays of a fused RK3 stage
        (pycles_advection_rk3).  Three arrays rotate: the stage-0 input stays
        unchanged through the step and is the copy of stages 1 and 2, so
        stage 0 writes no copy, and stage 2 writes its output over it (each
        cell's copy is read before its output is written).  Returns (out,
        copy), copy None at stage 0; make `out` current with set_values."""
        if getattr(self, "_vbufs", None) is None or not any(b is self.values for b in self._vbufs):
            self._vbufs = [self.values, torch.zeros_like(self.values), torch.zeros_like(self.values)]
        if rk_step == 0:
            self._rk_base = self.values
            self._rk_spare = [b for b in self._vbufs if b is not self.values]
            return self._rk_spare[0], None
        if rk_step == 1:
            return self._rk_spare[1], self._rk_base
        return self._rk_base, self._rk_base

    def set_values(self, arr):
        self.values = arr

    def tend(self, name) -> torch.Tensor:
        return self.tendencies[self.name_index[name]]

    def tptr(self, name) -> int:
        return self.tend(name).data_ptr()

    def base_ptr(self, name):
        b = self.base.get(name)
        return None if b is None else b.data_ptr()

    # ---- host <-> device in the reference layout -------------------------
    def set_base_from_device_layout(self, name, dev_layout: np.ndarray):
        """FP32 runs: base(k) = horizontal mean of the given field."""
        if self.precision == D.F32:
            prof = np.asarray(dev_layout, dtype=np.float64).mean(axis=(1, 2))
            self.base[name] = D.to_device(prof)

    def set_device_layout(self, name, arr: np.ndarray):
        """Upload a field given in device layout [k][j][i] (FP64 host array)."""
        t = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(self.values.device)
        if name in self.base:
            t = t - self.base[name][:, None, None]
        self.field(name).copy_(t.to(self.values.dtype))

    def get_device_layout(self, name) -> np.ndarray:
        t = self.field(name).to(torch.float64)
        if name in self.base:
            t = t + self.base[name][:, None, None]
        return t.cpu().numpy()

    def set_variable_array(self, name, Gr, ref_layout: np.ndarray):
        self.set_device_layout(name, Gr.to_device_layout(ref_layout))

    def get_variable_array(self, name, Gr) -> np.ndarray:
        """Reference ghosted [i][j][k] layout, FP64 (reference ``get_variable_array``)."""
        return Gr.from_device_layout(self.get_device_layout(name))

    def get_tendency_array(self, name, Gr) -> np.ndarray:
        return Gr.from_device_layout(self.tend(name).to(torch.float64).cpu().numpy())

    def zero_tendencies(self):
        self.tendencies.zero_()

    # ---- statistics ------------------------------------------------------
    def stats_io(self, Gr, RS, NS, Pa):
        """PrognosticVariables.stats_io (PrognosticVariables.pyx:106-148):
        mean, mean of squares and of cubes, max and min profiles of every
        variable (one device pass each), the translational mean velocities and
        the qt*s product mean."""
        gw = Gr.dims.gw
        for name in self.index_name:
            st = Pa.horizontal_stats(Gr, self.field(name), self.precision, self.base.get(name))
            NS.write_profile(name + "_mean", st[0, gw:-gw], Pa)
            if name == "u":
                NS.write_profile("u_translational_mean", st[0, gw:-gw] + RS.u0, Pa)
            elif name == "v":
                NS.write_profile("v_translational_mean", st[0, gw:-gw] + RS.v0, Pa)
            NS.write_profile(name + "_mean2", st[1, gw:-gw], Pa)
            NS.write_profile(name + "_mean3", st[2, gw:-gw], Pa)
            NS.write_profile(name + "_max", st[3, gw:-gw], Pa)
            NS.write_ts(name + "_max", np.amax(st[3, gw:-gw]), Pa)
            NS.write_profile(name + "_min", st[4, gw:-gw], Pa)
            NS.write_ts(name + "_min", np.amin(st[4, gw:-gw]), Pa)
        if "qt" in self.name_index and "s" in self.name_index:
            st = Pa.horizontal_stats(Gr, self.field("qt"), self.precision, self.base.get("qt"), self.field("s"),
                                     self.base.get("s"))
            NS.write_profile("qt_s_product_mean", st[0, gw:-gw], Pa)

    # ---- restart ---------------------------------------------------------
    def restart(self, Gr, Re):
        """PrognosticVariables.restart (PrognosticVariables.pyx:269-314): the
        registry and the values, here in the device layout [v][k][j][i] and
        the storage precision (FP32: the deviations and their FP64 level
        profiles), so that a resumed run continues bit for bit."""
        Re.restart_data["PV"] = {
            "name_index": dict(self.name_index), "units": dict(self.units),
            "index_name": list(self.index_name), "nv": self.nv, "nv_scalars": self.nv_scalars,
            "nv_velocities": self.nv_velocities, "bc_type": np.array(self.bc_type),
            "var_type": np.array(self.var_type), "velocity_directions": np.array(self.velocity_directions),
            "velocity_names_directional": list(self.velocity_names_directional),
            "precision": "f32" if self.precision == D.F32 else "f64", "layout": "[v][k][j][i]",
            "values": self.values.cpu().numpy(),
            "base": {n: b.cpu().numpy() for n, b in self.base.items()},
        }

    def init_from_restart(self, Gr, Re):
        """PrognosticVariables.init_from_restart (:316-357) into the already
        registered and allocated state, which must match the file's."""
        pv = Re.restart_data["PV"]
        prec = "f32" if self.precision == D.F32 else "f64"
        if list(pv["index_name"]) != list(self.index_name) or pv["precision"] != prec:
            raise ValueError(f"restart file holds {pv['index_name']} in {pv['precision']}, "
                             f"the run {self.index_name} in {prec}")
        vals = np.asarray(pv["values"])
        if tuple(vals.shape) != tuple(self.values.shape):
            raise ValueError(f"restart values of shape {vals.shape}, the run's {tuple(self.values.shape)}")
        self.values.copy_(torch.from_numpy(vals))
        self.base = {n: D.to_device(np.asarray(b)) for n, b in pv["base"].items()}
        self.tendencies.zero_()

    def val_nan(self, PA=None, message=""):
        if not torch.isfinite(self.values).all():
            raise FloatingPointError("Nans found in Prognostic Variables values " + message)

    def tend_nan(self, PA=None, message=""):
        if not torch.isfinite(self.tendencies).all():
            raise FloatingPointError("Nans found in tendencies " + message)


class DiagnosticVariables(_State):
    def __init__(self):
        self.name_index = {}
        self.index_name = []
        self.units = {}
        self.bc_type = []
        self.nv = 0

    def add_variables(self, name, units, nice_name, desc, bc_type, Pa=None):
        self._register(name, bc_type, Pa)
        self.units[name] = units

    def initialize(self, Gr, NS=None, Pa=None, precision=D.F64):
        self.precision = D.precision_code(precision)
        d = Gr.dims
        self.values = torch.zeros((self.nv, d.nzg, d.n[1], d.n[0]),
                                  dtype=D.torch_dtype(self.precision), device=D.require_gpu())

    def get_device_layout(self, name) -> np.ndarray:
        return self.field(name).to(torch.float64).cpu().numpy()

    def get_variable_array(self, name, Gr) -> np.ndarray:
        return Gr.from_device_layout(self.get_device_layout(name))

    def stats_io(self, Gr, NS, Pa):
        """DiagnosticVariables.stats_io (DiagnosticVariables.pyx:155-185)."""
        gw = Gr.dims.gw
        for name in self.index_name:
            st = Pa.horizontal_stats(Gr, self.field(name), self.precision)
            NS.write_profile(name + "_mean", st[0, gw:-gw], Pa)
            NS.write_profile(name + "_mean2", st[1, gw:-gw], Pa)
            NS.write_profile(name + "_mean3", st[2, gw:-gw], Pa)
            NS.write_profile(name + "_max", st[3, gw:-gw], Pa)
            NS.write_ts(name + "_max", np.amax(st[3, gw:-gw]), Pa)
            NS.write_profile(name + "_min", st[4, gw:-gw], Pa)
            NS.write_ts(name + "_min", np.amin(st[4, gw:-gw]), Pa)
