"""Saturation-adjustment and dry thermodynamics operators (device).

``ThermodynamicsSA`` mirrors reference ``ThermodynamicsSA.pyx``: the
constructor builds the Clausius-Clapeyron table (``ClausiusClapeyron``),
``initialize`` registers s, qt and the diagnostic fields, ``update(Gr, RS,
PV, DV)`` runs ``pycles_thermo_sa`` (qt clipping, saturation adjustment,
buoyancy into the w tendency, N^2, theta_rho, theta_l).  ``entropy``,
``alpha``, ``eos``, ``get_pv_star`` and ``get_lh`` are the reference's scalar
helpers used while building the reference state and initial conditions.
``ThermodynamicsDry`` mirrors ``ThermodynamicsDry.pyx``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _capi
from . import _device as D
from . import thermo_host as H
from .Thermodynamics import ClausiusClapeyron, LatentHeat


class ThermodynamicsSA:
    def __init__(self, namelist, LH: LatentHeat | None = None, Par=None):
        self.LH = LH or LatentHeat()
        self.CC = ClausiusClapeyron().initialize(namelist, self.LH, Par)
        th = (namelist or {}).get("thermodynamics", {})
        self.do_qt_clipping = th.get("do_qt_clipping", True)
        self.max_iter = int(th.get("max_iter", 100))
        self.mlp = None
        self.newton_steps = 1

    def initialize(self, Gr, PV, DV, NS=None, Pa=None):
        PV.add_variable("s", "J kg^-1 K^-1", "s", "specific entropy", "sym", "scalar", Pa)
        PV.add_variable("qt", "kg/kg", "q_t", "total water mass fraction", "sym", "scalar", Pa)
        for name in ("buoyancy", "alpha", "temperature", "buoyancy_frequency", "qv", "ql", "qi",
                     "theta_rho", "thetali"):
            DV.add_variables(name, "", name, name, "sym", Pa)

    # scalar helpers (host) ------------------------------------------------
    def entropy(self, p0, T, qt, ql, qi):
        return float(H.entropy_sa(p0, T, qt, ql, qi))

    def alpha(self, p0, T, qt, qv):
        return float(H.alpha_c(p0, T, qt, qv))

    def eos(self, p0, s, qt):
        return H.eos_sa(self.CC.table, p0, s, qt)

    def get_pv_star(self, t):
        return float(self.CC.table(t))

    def get_lh(self, t):
        return float(H.latent_heat_variable(t))

    # device ---------------------------------------------------------------
    def _prepare(self, Gr):
        if getattr(self, "_lt", None) is None:
            self._tab = D.to_device(self.CC.pv)
            t = self.CC.table
            self._lt = _capi.Lookup(self._tab.data_ptr(), len(self.CC.pv), t.x_min, t.dx, t.dxi)
            self.dims = D.dims_of(Gr)
            self.counters = torch.zeros(4, dtype=torch.int64, device=D.require_gpu())

    def options(self, wt_overwrite=False):
        return _capi.SAOptions(self.do_qt_clipping, self.max_iter, _capi.LH_VARIABLE,
                               None if self.mlp is None else self.mlp.handle_ptr(),
                               self.newton_steps, self.counters.data_ptr(), wt_overwrite)

    def update(self, Gr, RS, PV, DV, wt_overwrite=False):
        """wt_overwrite: the w tendency is zero (or stale and to be replaced),
        so the buoyancy term is written instead of added (same bits)."""
        self._prepare(Gr)
        ref = D.device_ref(RS)
        base = None
        if PV.precision == D.F32:
            base = _capi.ScalarBase(PV.base_ptr("s"), PV.base_ptr("qt"))
        _capi.thermo_sa(self.dims, ref.c, self._lt, PV.ptr("s"), PV.ptr("qt"), PV.tptr("w"),
                        DV.ptr("temperature"), DV.ptr("qv"), DV.ptr("ql"), DV.ptr("alpha"),
                        DV.ptr("buoyancy"), DV.ptr("theta_rho"), DV.ptr("buoyancy_frequency"),
                        DV.ptr("thetali"), base, self.options(wt_overwrite), PV.precision, D.stream())


class ThermodynamicsDry:
    def __init__(self, namelist, LH: LatentHeat | None = None, Pa=None):
        self.LH = LH or LatentHeat()
        self.CC = ClausiusClapeyron().initialize(namelist, self.LH, Pa)

    def initialize(self, Gr, PV, DV, NS=None, Pa=None):
        PV.add_variable("s", "J kg^-1 K^-1", "s", "specific entropy", "sym", "scalar", Pa)
        for name in ("buoyancy", "alpha", "temperature", "buoyancy_frequency", "theta"):
            DV.add_variables(name, "", name, name, "sym", Pa)

    def entropy(self, p0, T, qt, ql, qi):
        return float(H.entropy_dry(p0, T))

    def eos(self, p0, s, qt):
        return float(H.eos_dry(p0, s)), 0.0, 0.0

    def alpha(self, p0, T, qt, qv):
        return float(H.alpha_c(p0, T, 0.0, 0.0))

    def update(self, Gr, RS, PV, DV, wt_overwrite=False):
        if getattr(self, "dims", None) is None:
            self.dims = D.dims_of(Gr)
        ref = D.device_ref(RS)
        base = _capi.ScalarBase(PV.base_ptr("s"), None) if PV.precision == D.F32 else None
        _capi.thermo_dry(self.dims, ref.c, PV.ptr("s"), PV.tptr("w"), DV.ptr("temperature"),
                         DV.ptr("alpha"), DV.ptr("buoyancy"), DV.ptr("theta"),
                         DV.ptr("buoyancy_frequency"), base, PV.precision, D.stream(),
                         1 if wt_overwrite else 0)


def ThermodynamicsFactory(namelist, Micro=None, LH=None, Par=None):
    """Reference ``Thermodynamics.pyx:123-128``: None_Dry -> dry, None_SA -> SA."""
    scheme = (namelist or {}).get("microphysics", {}).get("scheme", "None_Dry")
    if scheme == "None_SA":
        return ThermodynamicsSA(namelist, LH, Par)
    if scheme == "None_Dry":
        return ThermodynamicsDry(namelist, LatentHeat(model=1), Par)
    raise ValueError(f"microphysics scheme {scheme!r} is outside the device path")
