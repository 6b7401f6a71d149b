"""Scalar (host) thermodynamic functions used while setting up a run.

These back the scalar helpers that the reference exposes on its
thermodynamics classes (``ThermodynamicsSA.entropy/alpha/eos``,
``ThermodynamicsSA.pyx:138-192``; ``ThermodynamicsDry.entropy/eos/alpha``,
``ThermodynamicsDry.pyx:93-103``), which ``ReferenceState.initialize`` and the
case initializers call once per level or per cell.  They are initialization
code, not the per-stage path: the per-stage path runs on the GPU only.
Vectorised over numpy arrays where the initializers need whole fields.
"""
from __future__ import annotations

import numpy as np

from . import parameters as P


def latent_heat_variable(T):
    """Reference ``Microphysics.pxd:71-75``."""
    TC = T - 273.15
    return (2500.8 - 2.36 * TC + 0.0016 * TC * TC - 0.00006 * TC * TC * TC) * 1000.0


def latent_heat_constant(T):
    """Reference ``Microphysics.pxd:68-69``."""
    return 2.501e6 + 0.0 * T


def exner(p0):
    return np.power(p0 / P.p_tilde, P.kappa)


def pd_c(p0, qt, qv):
    return p0 * (1.0 - qt) / (1.0 - qt + P.eps_vi * qv)


def pv_c(p0, qt, qv):
    return p0 * P.eps_vi * qv / (1.0 - qt + P.eps_vi * qv)


def sd_c(pd, T):
    return P.sd_tilde + P.cpd * np.log(T / P.T_tilde) - P.Rd * np.log(pd / P.p_tilde)


def sv_c(pv, T):
    return P.sv_tilde + P.cpv * np.log(T / P.T_tilde) - P.Rv * np.log(pv / P.p_tilde)


def qv_star_c(p0, qt, pv):
    return P.eps_v * (1.0 - qt) * pv / (p0 - pv)


def alpha_c(p0, T, qt, qv):
    return (P.Rd * T) / p0 * (1.0 - qt + P.eps_vi * qv)


def entropy_sa(p0, T, qt, ql, qi):
    """Moist entropy, Pressel et al. 2015 eq. (40)."""
    qv = qt - ql - qi
    pd = pd_c(p0, qt, qv)
    pv = pv_c(p0, qt, qv)
    L = latent_heat_variable(T)
    return sd_c(pd, T) * (1.0 - qt) + sv_c(pv, T) * qt + (-L / T) * (ql + qi)


def entropy_dry(pd, T):
    return sd_c(pd, T)


def eos_dry(pd, s):
    return P.T_tilde * np.exp((s - P.sd_tilde + P.Rd * np.log(pd / P.p_tilde)) / P.cpd)


class PvStarTable:
    """Linear interpolation in a pv*(T) table, as reference ``lookup.h:48-54``."""

    def __init__(self, T, pv):
        self.x_min = float(np.min(T))
        self.dx = float(T[1] - T[0])
        self.dxi = 1.0 / self.dx
        self.y = np.asarray(pv, dtype=np.float64)

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        indx = np.floor((x - self.x_min) * self.dxi).astype(np.int64)
        y1 = self.y[indx]
        y2 = self.y[indx + 1]
        x1 = self.x_min + self.dx * indx
        return y1 + (x - x1) * (y2 - y1) * self.dxi


def eos_sa(table: PvStarTable, p0: float, s: float, qt: float, max_iter: int = 200):
    """Saturation adjustment for one state (secant, stops at |dT| < 1e-3 K
    with sigma >= 0, as reference ``thermodynamics_sa.h:17-66``).
    Returns (T, ql, qi)."""
    pv_1 = pv_c(p0, qt, qt)
    pd_1 = p0 - pv_1
    T_1 = P.T_tilde * np.exp((s - (1.0 - qt) * (P.sd_tilde - P.Rd * np.log(pd_1 / P.p_tilde))
                              - qt * (P.sv_tilde - P.Rv * np.log(pv_1 / P.p_tilde)))
                             / ((1.0 - qt) * P.cpd + qt * P.cpv))
    qv_star_1 = qv_star_c(p0, qt, table(T_1))
    if qt <= qv_star_1:
        return float(T_1), 0.0, 0.0
    sigma_1 = qt - qv_star_1
    L_1 = latent_heat_variable(T_1)
    s_1 = sd_c(pd_1, T_1) * (1.0 - qt) + sv_c(pv_1, T_1) * qt + (-L_1 / T_1) * sigma_1
    f_1 = s - s_1
    T_2 = T_1 + sigma_1 * L_1 / ((1.0 - qt) * P.cpd + qv_star_1 * P.cpv)
    for _ in range(max_iter):
        qv_star_2 = qv_star_c(p0, qt, table(T_2))
        pv_2 = pv_c(p0, qt, qv_star_2)
        pd_2 = p0 - pv_2
        sigma_2 = qt - qv_star_2
        L_2 = latent_heat_variable(T_2)
        s_2 = sd_c(pd_2, T_2) * (1.0 - qt) + sv_c(pv_2, T_2) * qt + (-L_2 / T_2) * sigma_2
        f_2 = s - s_2
        T_n = T_2 - f_2 * (T_2 - T_1) / (f_2 - f_1)
        T_1, T_2, f_1 = T_2, T_n, f_2
        if abs(T_2 - T_1) < 1.0e-3 and sigma_2 >= 0.0:
            break
    return float(T_2), float(sigma_2), 0.0
