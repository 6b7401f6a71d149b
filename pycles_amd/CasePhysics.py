"""Surface, large-scale forcing and damping of the BOMEX case (device).

Mirror reference ``Surface.pyx`` (``SurfaceFactory``, ``SurfaceBomex``,
``SurfaceNone``), ``Forcing.pyx`` (``Forcing`` with ``ForcingBomex``, a no-op
otherwise) and ``Damping.pyx`` (``Damping`` with ``Rayleigh`` or ``Dummy``)
with their ``initialize``/``update`` signatures.  The profiles (geostrophic
wind, radiative cooling, drying, subsidence; the damping coefficients) are
built once on the host with the reference's expressions, copied to HBM, and
every per-cell update runs in ``pycles_surface_bomex``,
``pycles_forcing_bomex`` and ``pycles_rayleigh_damping``.  Horizontal means
come from ``pycles_level_mean``.  Other cases' surface and forcing schemes
are not on the device path.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _capi
from . import _device as D
from . import thermo_host as H


def _means(Gr, PV, names, dev):
    """Level means of the stored fields (FP64 [len(names)][nzg])."""
    out = torch.empty((len(names), Gr.dims.nzg), dtype=torch.float64, device=dev)
    dims = D.dims_of(Gr)
    for q, n in enumerate(names):
        _capi.level_mean(dims, PV.ptr(n), out[q].data_ptr(), PV.precision, D.stream())
    return out


def _qt_base(PV):
    return _capi.ScalarBase(PV.base_ptr("s"), PV.base_ptr("qt")) if PV.precision == D.F32 else None


class SurfaceNone:
    def initialize(self, Gr, Ref, NS=None, Pa=None):
        return None

    def update(self, Gr, Ref, PV, DV, Pa=None, TS=None):
        return None

    def stats_io(self, Gr, NS, Pa=None):
        return None


class SurfaceBomex:
    """SurfaceBomex (Surface.pyx:297-378)."""

    theta_flux = 8.0e-3
    ustar_ = 0.28
    theta_surface = 299.1
    qt_surface = 22.45e-3
    qt_flux_value = 5.2e-5

    def __init__(self, LH=None):
        self.dry_case = False
        self.flux2d = None

    def initialize(self, Gr, Ref, NS=None, Pa=None):
        d = Gr.dims
        self.T_surface = Ref.Tg
        self.flux2d = torch.zeros((8, d.n[1], d.n[0]), dtype=torch.float64, device=D.require_gpu())
        eps_vi = H.P.eps_vi
        self.buoyancy_flux = H.P.g * ((self.theta_flux + (eps_vi - 1.0) * (self.theta_surface * self.qt_flux_value
                                                                          + self.qt_surface * self.theta_flux))
                                      / (self.theta_surface * (1.0 + (eps_vi - 1) * self.qt_surface)))
        for name in ("uw_surface_mean", "vw_surface_mean", "s_flux_surface_mean", "shf_surface_mean",
                     "lhf_surface_mean", "obukhov_length_mean", "friction_velocity_mean",
                     "buoyancy_flux_surface_mean"):
            if NS is not None:
                NS.add_ts(name, Gr, Pa)

    def update(self, Gr, Ref, PV, DV, Pa=None, TS=None):
        dev = PV.values.device
        t_mean = torch.empty(Gr.dims.nzg, dtype=torch.float64, device=dev)
        _capi.level_mean(D.dims_of(Gr), DV.ptr("temperature"), t_mean.data_ptr(), DV.precision, D.stream())
        self._t_mean = t_mean
        _capi.surface_bomex(D.dims_of(Gr), D.device_ref(Ref).c, PV.ptr("u"), PV.ptr("v"), PV.ptr("qt"),
                            DV.ptr("temperature"), DV.ptr("qv"), DV.ptr("ql"), _qt_base(PV),
                            [PV.tptr("u"), PV.tptr("v"), PV.tptr("s"), PV.tptr("qt")], Ref.u0, Ref.v0,
                            t_mean.data_ptr(), self.flux2d.data_ptr(), PV.precision, D.stream())

    def stats_io(self, Gr, NS, Pa=None):
        """SurfaceBase.stats_io (Surface.pyx:187-214): horizontal means of the
        column fluxes and diagnostics."""
        m = self.flux2d.cpu().numpy().mean(axis=(1, 2))
        for q, name in enumerate(("uw_surface_mean", "vw_surface_mean", "s_flux_surface_mean", None,
                                  "shf_surface_mean", "lhf_surface_mean", "buoyancy_flux_surface_mean",
                                  "obukhov_length_mean")):
            if name:
                NS.write_ts(name, float(m[q]), Pa)
        NS.write_ts("friction_velocity_mean", self.ustar_, Pa)

    def restart(self, Re):
        Re.restart_data["surf"] = {"T_surf": self.T_surface}

    def init_from_restart(self, Re):
        self.T_surface = Re.restart_data["surf"]["T_surf"]


def SurfaceFactory(namelist, LH=None, Par=None):
    """Surface.pyx:44-71: by namelist['meta']['casename']."""
    case = (namelist or {}).get("meta", {}).get("casename")
    if case == "Bomex":
        return SurfaceBomex(LH)
    return SurfaceNone()


class ForcingNone:
    def initialize(self, Gr, RS, Th=None, NS=None, Pa=None):
        return None

    def update(self, Gr, RS, PV, DV, Pa=None):
        return None


class ForcingBomex:
    """ForcingBomex (Forcing.pyx:102-203)."""

    coriolis_param = 0.376e-4

    def initialize(self, Gr, RS, Th=None, NS=None, Pa=None):
        n = Gr.dims.nlg[2]
        z = Gr.zl_half
        ug = np.zeros(n)
        vg = np.zeros(n)
        dtdt = np.zeros(n)
        dqtdt = np.zeros(n)
        sub = np.zeros(n)
        for k in range(n):
            ug[k] = -10.0 + (1.8e-3) * z[k]
            ex = (RS.p0_half[k] / H.P.p_tilde) ** H.P.kappa
            if z[k] <= 1500.0:
                dtdt[k] = (-2.0 / (3600 * 24.0)) * ex
            if z[k] > 1500.0:
                dtdt[k] = (-2.0 / (3600 * 24.0) + (z[k] - 1500.0) * (0.0 - -2.0 / (3600 * 24.0))
                           / (3000.0 - 1500.0)) * ex
            if z[k] <= 300.0:
                dqtdt[k] = -1.2e-8
            if 300.0 < z[k] <= 500.0:
                dqtdt[k] = -1.2e-8 + (z[k] - 300.0) * (0.0 - -1.2e-8) / (500.0 - 300.0)
            if z[k] <= 1500.0:
                sub[k] = 0.0 + z[k] * (-0.65 / 100.0 - 0.0) / (1500.0 - 0.0)
            if 1500.0 < z[k] <= 2100.0:
                sub[k] = -0.65 / 100 + (z[k] - 1500.0) * (0.0 - -0.65 / 100.0) / (2100.0 - 1500.0)
        self.profiles_host = np.stack([ug, vg, dtdt, dqtdt, sub])
        self.ug, self.vg, self.dtdt, self.dqtdt, self.subsidence = self.profiles_host
        self.profiles = D.to_device(self.profiles_host)

    def update(self, Gr, RS, PV, DV, Pa=None):
        uv = _means(Gr, PV, ("u", "v"), PV.values.device)
        _capi.forcing_bomex(D.dims_of(Gr), D.device_ref(RS).c, self.profiles.data_ptr(), uv.data_ptr(),
                            PV.ptr("u"), PV.ptr("v"), PV.ptr("s"), PV.ptr("qt"), DV.ptr("temperature"),
                            DV.ptr("ql"), _qt_base(PV), [PV.tptr("u"), PV.tptr("v"), PV.tptr("s"), PV.tptr("qt")],
                            self.coriolis_param, RS.u0, RS.v0, PV.precision, D.stream())


class Forcing:
    """Forcing.pyx:29-86: by namelist['meta']['casename']."""

    def __init__(self, namelist, LH=None, Pa=None):
        case = (namelist or {}).get("meta", {}).get("casename")
        self.scheme = ForcingBomex() if case == "Bomex" else ForcingNone()

    def initialize(self, Gr, RS, Th=None, NS=None, Pa=None):
        self.scheme.initialize(Gr, RS, Th, NS, Pa)

    def update(self, Gr, RS, PV, DV, Pa=None):
        self.scheme.update(Gr, RS, PV, DV, Pa)


class Dummy:
    def initialize(self, Gr, RS):
        return None

    def update(self, Gr, RS, PV, DV, Pa=None):
        return None


class Rayleigh:
    """Rayleigh (Damping.pyx:54-133)."""

    def __init__(self, namelist, Pa=None):
        r = namelist["damping"]["Rayleigh"]
        self.z_d = float(r["z_d"])
        self.gamma_r = float(r["gamma_r"])

    def initialize(self, Gr, RS):
        n = Gr.dims.nlg[2]
        z_top = Gr.dims.dx[2] * Gr.dims.n[2]
        gz = np.zeros(n)
        gzh = np.zeros(n)
        for k in range(n):
            if Gr.zl_half[k] >= z_top - self.z_d:
                gzh[k] = self.gamma_r * math.sin((math.pi / 2.0) * (1.0 - (z_top - Gr.zl_half[k]) / self.z_d)) ** 2.0
            if Gr.zl[k] >= z_top - self.z_d:
                gz[k] = self.gamma_r * math.sin((math.pi / 2.0) * (1.0 - (z_top - Gr.zl[k]) / self.z_d)) ** 2.0
        self.gamma_z_host, self.gamma_zhalf_host = gz, gzh
        self.gamma_z = D.to_device(gz)
        self.gamma_zhalf = D.to_device(gzh)

    def update(self, Gr, RS, PV, DV, Pa=None):
        names = list(PV.index_name)
        means = _means(Gr, PV, names, PV.values.device)
        gam = [(self.gamma_zhalf if n == "w" else self.gamma_z).data_ptr() for n in names]
        _capi.rayleigh_damping(D.dims_of(Gr), [PV.ptr(n) for n in names], [PV.tptr(n) for n in names], gam,
                               means.data_ptr(), PV.precision, D.stream())


class Damping:
    """Damping.pyx:23-44: namelist['damping']['scheme'] None or Rayleigh."""

    def __init__(self, namelist, Pa=None):
        scheme = (namelist or {}).get("damping", {}).get("scheme", "None")
        self.scheme = Rayleigh(namelist, Pa) if scheme == "Rayleigh" else Dummy()

    def initialize(self, Gr, RS):
        self.scheme.initialize(Gr, RS)

    def update(self, Gr, RS, PV, DV, Pa=None):
        self.scheme.update(Gr, RS, PV, DV, Pa)
