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
        # Newton corrections after the learned guess: the second takes the
        # FP32 result from the residual check's 3e-4 K to FP32 rounding (the
        # 1-step T error against FP64 falls from 9.5e-7 to 2.1e-7 relative)
        # and leaves no cell to the fallback, for ~1% of the step
        self.newton_steps = int(th.get("newton_steps", 2))
        # the FP64 secant's stopping test |dT| < secant_tol; 0 keeps the
        # reference's 1e-3 K (a smaller value is for measuring how far that
        # stopping point moves a run)
        self.secant_tol = float(th.get("secant_tol", 0.0))

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

    def stats_io(self, Gr, RS, PV, DV, NS, Pa):
        """ThermodynamicsSA.stats_io (ThermodynamicsSA.pyx:256-389): profiles
        of the entropy potential temperature, the dry potential temperature
        and the relative humidity, then liquid_stats (:395-492)."""
        self._prepare(Gr)
        gw = Gr.dims.gw
        ref = D.device_ref(RS)
        base = _capi.ScalarBase(PV.base_ptr("s"), PV.base_ptr("qt")) if PV.precision == D.F32 else None
        tmp = torch.empty((3,) + tuple(PV.field("s").shape), dtype=PV.values.dtype, device=PV.values.device)
        _capi.sa_stat_fields(self.dims, ref.c, self._lt, PV.ptr("s"), PV.ptr("qt"), DV.ptr("temperature"),
                             DV.ptr("qv"), base, PV.precision, tmp[0].data_ptr(), tmp[1].data_ptr(),
                             tmp[2].data_ptr(), D.stream())
        for q, name, moments in ((0, "thetas", True), (1, "theta", True), (2, "rh", False)):
            st = Pa.horizontal_stats(Gr, tmp[q], PV.precision)
            NS.write_profile(name + "_mean", st[0, gw:-gw], Pa)
            if moments:
                NS.write_profile(name + "_mean2", st[1, gw:-gw], Pa)
                NS.write_profile(name + "_mean3", st[2, gw:-gw], Pa)
            NS.write_profile(name + "_max", st[3, gw:-gw], Pa)
            NS.write_profile(name + "_min", st[4, gw:-gw], Pa)
            if moments:
                NS.write_ts(name + "_max", np.amax(st[3, gw:-gw]), Pa)
                NS.write_ts(name + "_min", np.amin(st[4, gw:-gw]), Pa)
        self.liquid_stats(Gr, RS, PV, DV, NS, Pa)

    def liquid_stats(self, Gr, RS, PV, DV, NS, Pa):
        """ThermodynamicsSA.liquid_stats (ThermodynamicsSA.pyx:395-492)."""
        cf, cover, cb, ct, lwp = Pa.cloud_stats(Gr, RS, DV.field("ql"), DV.precision)
        NS.write_profile("cloud_fraction", cf, Pa)
        NS.write_ts("cloud_fraction", cover, Pa)
        NS.write_ts("cloud_base", cb, Pa)
        NS.write_ts("cloud_top", ct, Pa)
        NS.write_ts("lwp", lwp, Pa)

    def restart(self, Re):
        """The learned first guess, when one is in use (no reference
        counterpart: the reference has no learned part)."""
        if self.mlp is not None:
            Re.restart_data["Th"] = {"learned_guess": self.mlp.state_dict()}

    def init_from_restart(self, Re):
        st = Re.restart_data.get("Th", {}).get("learned_guess")
        if st is not None:
            from .LearnedGuess import SaturationGuess
            self.mlp = SaturationGuess(self.CC, n_pairs=1 << 7, batch=128, epochs=1).load_state_dict(st)

    def options(self, wt_overwrite=False):
        return _capi.SAOptions(self.do_qt_clipping, self.max_iter, _capi.LH_VARIABLE,
                               None if self.mlp is None else self.mlp.handle_ptr(),
                               self.newton_steps, self.counters.data_ptr(), wt_overwrite, self.secant_tol)

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
