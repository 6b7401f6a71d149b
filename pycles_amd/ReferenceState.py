"""Hydrostatic anelastic reference state (reference ``ReferenceState.pyx:39-200``).

Integrates d ln p/dz = -g / (Rd T (1 - qtg + eps_vi (qtg - ql - qi))) with
T from the thermodynamics' ``eos`` at constant surface entropy, on full levels
z[gw-1 .. nlg-gw] and on half levels [0, z_half interior], mirrors the
ghost levels, and evaluates alpha0/rho0 per level.  The profiles are inputs
to every kernel; they are computed once on the host and copied to HBM.
"""
from __future__ import annotations

import numpy as np
from scipy.integrate import odeint

from . import parameters as P


class ReferenceState:
    def __init__(self, Gr):
        n = Gr.dims.nlg[2]
        for name in ("p0", "p0_half", "alpha0", "alpha0_half", "rho0", "rho0_half"):
            setattr(self, name, np.zeros(n))
        self.Pg = 1.0e5
        self.Tg = 300.0
        self.qtg = 0.0
        self.u0 = 0.0
        self.v0 = 0.0

    def initialize(self, Gr, Thermodynamics, NS=None, Pa=None):
        d = Gr.dims
        gw = d.gw
        self.sg = Thermodynamics.entropy(self.Pg, self.Tg, self.qtg, 0.0, 0.0)

        def rhs(p, z):
            T, ql, qi = Thermodynamics.eos(float(np.exp(p[0])), self.sg, self.qtg)
            return -P.g / (P.Rd * T * (1.0 - self.qtg + P.eps_vi * (self.qtg - ql - qi)))

        z = np.array(Gr.z[gw - 1:-gw + 1])
        z_half = np.append([0.0], np.array(Gr.z_half[gw:-gw]))
        p0 = np.log(self.Pg)
        p = np.zeros(d.ng[2])
        p_half = np.zeros(d.ng[2])
        p[gw - 1:-gw + 1] = odeint(rhs, p0, z, hmax=1.0)[:, 0]
        p_half[gw:-gw] = odeint(rhs, p0, z_half, hmax=1.0)[1:, 0]
        p[:gw - 1] = p[2 * gw - 2:gw - 1:-1]
        p[-gw + 1:] = p[-gw - 1:-2 * gw:-1]
        p_half[:gw] = p_half[2 * gw - 1:gw - 1:-1]
        p_half[-gw:] = p_half[-gw - 1:-2 * gw - 1:-1]
        p = np.exp(p)
        p_half = np.exp(p_half)
        alpha = np.zeros(d.ng[2])
        alpha_half = np.zeros(d.ng[2])
        for k in range(d.ng[2]):
            T, ql, qi = Thermodynamics.eos(p[k], self.sg, self.qtg)
            alpha[k] = Thermodynamics.alpha(p[k], T, self.qtg, self.qtg - (ql + qi))
            Th, qlh, qih = Thermodynamics.eos(p_half[k], self.sg, self.qtg)
            alpha_half[k] = Thermodynamics.alpha(p_half[k], Th, self.qtg, self.qtg - (qlh + qih))
            s = Thermodynamics.entropy(p_half[k], Th, self.qtg, qlh, qih)
            if abs(s - self.sg) / self.sg > 0.01:
                raise RuntimeError("reference profile entropy not constant: saturation adjustment error")
        self.p0 = p
        self.p0_half = p_half
        self.alpha0 = alpha
        self.alpha0_half = alpha_half
        self.rho0 = 1.0 / alpha
        self.rho0_half = 1.0 / alpha_half
        return self

    def restart(self, Gr, Re):
        """ReferenceState.restart (ReferenceState.pyx:202-223)."""
        Re.restart_data["Ref"] = {n: np.array(getattr(self, n)) for n in ("p0", "p0_half", "alpha0", "alpha0_half")}
        Re.restart_data["Ref"].update({n: float(getattr(self, n, 0.0)) for n in ("Tg", "Pg", "sg", "qtg", "u0", "v0")})

    def init_from_restart(self, Gr, Re):
        """ReferenceState.init_from_restart (:225-245): rho0 = 1/alpha0."""
        r = Re.restart_data["Ref"]
        for n in ("Tg", "Pg", "sg", "qtg", "u0", "v0"):
            setattr(self, n, r[n])
        for n in ("p0", "p0_half", "alpha0", "alpha0_half"):
            setattr(self, n, np.array(r[n], dtype=np.float64))
        self.rho0 = 1.0 / self.alpha0
        self.rho0_half = 1.0 / self.alpha0_half
        self._device_ref = None
        return self
