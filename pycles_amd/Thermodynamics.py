"""Latent heat selection and the Clausius-Clapeyron pv*(T) table.

Mirrors reference ``Thermodynamics.pyx``: ``LatentHeat`` carries the latent
heat model that the microphysics choice installs (``Microphysics.pyx:31-53``:
constant for ``None_Dry``, variable for ``None_SA``); ``ClausiusClapeyron``
integrates d ln pv*/dT = L(T)/(Rv T^2) from the triple point (Tt, 611.7 Pa)
up and down over ``linspace(Tmin, Tmax, n_lookup)`` with ``scipy.integrate.odeint``
(``Thermodynamics.pyx:45-116``).  Where the reference hands the kernels C
function pointers for L(T) and lambda(T), the HIP kernels take an enum
(``PYCLES_LH_VARIABLE`` / ``PYCLES_LH_CONSTANT``).
"""
from __future__ import annotations

import numpy as np
from scipy.integrate import odeint

from . import parameters as P
from . import thermo_host as H

LH_VARIABLE = 0
LH_CONSTANT = 1


class LatentHeat:
    def __init__(self, namelist=None, Par=None, model=LH_VARIABLE):
        self.model = model

    def L(self, T, Lambda=1.0):
        if self.model == LH_CONSTANT:
            return 2.501e6
        return H.latent_heat_variable(T)

    def Lambda(self, T):
        return 1.0


class ClausiusClapeyron:
    def initialize(self, namelist, LH: LatentHeat, Par=None):
        cc = (namelist or {}).get("ClausiusClapeyron", {})
        Tmin = cc.get("temperature_min", 100.15)
        Tmax = cc.get("temperature_max", 380.0)
        n_lookup = cc.get("n_lookup", 512)
        T = np.linspace(Tmin, Tmax, n_lookup)
        tp_close_index = np.max(np.where(T <= P.Tt))
        if T[tp_close_index] == P.Tt:
            raise ValueError("ClausiusClapeyron table temperatures contain Tt; pick other bounds")
        T_above = np.append([P.Tt], T[tp_close_index + 1:])
        T_below = np.append(T[:tp_close_index + 1], [P.Tt])[::-1]

        def rhs(z, T_):
            return LH.L(T_, LH.Lambda(T_)) / (P.Rv * T_ * T_)

        pv0 = np.log(P.pv_star_t)
        pv_above = np.exp(odeint(rhs, pv0, T_above, hmax=0.1)[1:])
        pv_below = np.exp(odeint(rhs, pv0, T_below, hmax=0.1)[1:])[::-1]
        pv = np.append(pv_below, pv_above)
        pv[pv < 1e-11] = 1e-11
        self.T = T
        self.pv = np.ascontiguousarray(pv, dtype=np.float64)
        self.table = H.PvStarTable(T, self.pv)
        return self
