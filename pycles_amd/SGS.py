"""Kinematics, subgrid-scale closure and diffusion operators (device).

Mirror reference ``Kinematics.pyx``, ``SGS.pyx``, ``ScalarDiffusion.pyx`` and
``MomentumDiffusion.pyx`` with their ``update`` signatures:

- ``Kinematics.update(Gr, PV)`` (Kinematics.pyx:41-58).  The reference
  stores nine velocity gradients and nine strain-rate components; here the
  strain rate is formed in registers where it is used, so ``update`` only
  marks the velocities of the stage, and the magnitude |S| is written by
  the SGS kernel (it needs no more than the velocities).
- ``SGS(namelist)`` with ``sgs.scheme`` ``Smagorinsky`` (cs, prt; SGS.pyx:
  102-152: ``pycles_sgs_smagorinsky`` computes |S|, nu and D in one pass) or
  ``UniformViscosity`` (SGS.pyx:60-90: constant nu and D).  The wall and
  iles variants and the TKE scheme are not on the device path.
- ``ScalarDiffusion.update(Gr, RS, PV, DV)`` and
  ``MomentumDiffusion.update(Gr, Rs, PV, DV, Ke)``: ``pycles_diffusion``
  with the scalar, momentum and entropy-source parts.  ``diffuse`` does both
  operators' work in one pass (same bits).
"""
from __future__ import annotations

import torch

from . import _capi
from . import _device as D


class Kinematics:
    def __init__(self):
        self.strain_rate_mag = None

    def initialize(self, Gr, NS=None, Pa=None, precision=D.F64):
        d = Gr.dims
        self.strain_rate_mag = torch.zeros((d.nzg, d.n[1], d.n[0]), dtype=D.torch_dtype(D.precision_code(precision)),
                                           device=D.require_gpu())

    def update(self, Gr, PV):
        """The strain-rate magnitude of the stage's velocities is computed by
        the SGS update, which runs before anything changes them."""
        return None


class UniformViscosity:
    """SGS.pyx:60-90: constant viscosity and diffusivity, set once."""

    def __init__(self, namelist):
        uv = namelist.get("sgs", {}).get("UniformViscosity", {})
        self.const_diffusivity = float(uv.get("diffusivity", 0.0))
        self.const_viscosity = float(uv.get("viscosity", 0.0))
        self.is_init = False

    def initialize(self, Gr, PV, NS=None, Pa=None):
        return None

    def update(self, Gr, DV, PV, Ke, Sur=None, Pa=None):
        if not self.is_init:
            DV.field("diffusivity").fill_(self.const_diffusivity)
            DV.field("viscosity").fill_(self.const_viscosity)
            self.is_init = True


class Smagorinsky:
    """SGS.pyx:92-152 (smagorinsky_update, sgs.h:4-19)."""

    def __init__(self, namelist):
        sm = namelist.get("sgs", {}).get("Smagorinsky", {})
        self.cs = float(sm.get("cs", 0.17))
        self.prt = float(sm.get("prt", 1.0 / 3.0))
        if sm.get("wall", False) or sm.get("iles", False):
            raise ValueError("Smagorinsky wall / iles variants are not on the device path")

    def initialize(self, Gr, PV, NS=None, Pa=None):
        return None

    def update(self, Gr, DV, PV, Ke, Sur=None, Pa=None):
        n = PV.velocity_names_directional
        _capi.sgs_smagorinsky(D.dims_of(Gr), [PV.ptr(n[0]), PV.ptr(n[1]), PV.ptr(n[2])],
                              DV.ptr("buoyancy_frequency"), self.cs, self.prt, Ke.strain_rate_mag.data_ptr(),
                              DV.ptr("viscosity"), DV.ptr("diffusivity"), PV.precision, D.stream())


class SGS:
    """SGS.pyx:17-57: the scheme named by namelist['sgs']['scheme']."""

    def __init__(self, namelist):
        scheme = namelist["sgs"]["scheme"]
        if scheme == "UniformViscosity":
            self.scheme = UniformViscosity(namelist)
        elif scheme == "Smagorinsky":
            self.scheme = Smagorinsky(namelist)
        else:
            raise ValueError(f"sgs scheme {scheme!r} is outside the device path")

    def initialize(self, Gr, PV, NS=None, Pa=None):
        self.scheme.initialize(Gr, PV, NS, Pa)

    def update(self, Gr, DV, PV, Ke, Sur=None, Pa=None):
        self.scheme.update(Gr, DV, PV, Ke, Sur, Pa)


def _scalars(PV):
    names = [n for n in ("s", "qt") if n in PV.name_index]
    base = None
    if PV.precision == D.F32 and names:
        base = _capi.ScalarBase(*[PV.base_ptr(n) for n in names])
    return names, base


def _diffuse(Gr, RS, PV, DV, Ke, flags):
    names, base = _scalars(PV)
    n = PV.velocity_names_directional
    ref = D.device_ref(RS)
    mom = flags & (_capi.DIFF_MOMENTUM | _capi.DIFF_ENTROPY_SOURCE)
    _capi.diffusion(D.dims_of(Gr), ref.c, [PV.ptr(x) for x in n],
                    [PV.tptr(x) for x in n] if flags & _capi.DIFF_MOMENTUM else None,
                    [PV.ptr(x) for x in names], [PV.tptr(x) for x in names], base,
                    DV.ptr("viscosity") if mom else None, DV.ptr("diffusivity"),
                    Ke.strain_rate_mag.data_ptr() if Ke is not None and mom else None,
                    DV.ptr("temperature") if flags & _capi.DIFF_ENTROPY_SOURCE else None, flags, PV.precision,
                    D.stream())


class ScalarDiffusion:
    """ScalarDiffusion.pyx:23-119 (second order, no qt entropy source)."""

    def __init__(self, namelist, LH=None, DV=None, Pa=None):
        if DV is not None:
            DV.add_variables("diffusivity", "m^2s^{-1}", "D_t", "eddy diffusivity", "sym", Pa)
        if (namelist or {}).get("diffusion", {}).get("qt_entropy_source", False):
            raise ValueError("the qt entropy source of scalar diffusion is not on the device path")

    def initialize(self, Gr, PV, DV, NS=None, Pa=None):
        return None

    def update(self, Gr, RS, PV, DV):
        _diffuse(Gr, RS, PV, DV, None, _capi.DIFF_SCALARS)


class MomentumDiffusion:
    """MomentumDiffusion.pyx:22-66."""

    def __init__(self, DV=None, Pa=None):
        if DV is not None:
            DV.add_variables("viscosity", "m^2s^{-1}", r"\nu_t", "eddy viscosity", "sym", Pa)

    def initialize(self, Gr, PV, DV, NS=None, Pa=None):
        return None

    def update(self, Gr, Rs, PV, DV, Ke):
        _diffuse(Gr, Rs, PV, DV, Ke, _capi.DIFF_MOMENTUM | _capi.DIFF_ENTROPY_SOURCE)


def diffuse(SD, MD, Gr, RS, PV, DV, Ke):
    """SD.update then MD.update in one pass of the device kernel."""
    flags = 0
    if SD is not None:
        flags |= _capi.DIFF_SCALARS
    if MD is not None:
        flags |= _capi.DIFF_MOMENTUM | _capi.DIFF_ENTROPY_SOURCE
    if flags:
        _diffuse(Gr, RS, PV, DV, Ke, flags)
