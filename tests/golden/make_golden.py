"""Writes tests/golden/thermo_profiles.npz: the Clausius-Clapeyron pv*(T)
table and the anelastic reference profiles of the moist (BOMEX-like, 20 m,
128 levels) and dry (40 m, 64 levels) configurations, as this build's host
code computes them with scipy's odeint.  They are the inputs the oracle and
the HIP path share; the fixture fixes them so that a change in the host-side
integration shows up as a test failure.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pycles_amd import cases  # noqa: E402
from pycles_amd.Grid import Grid  # noqa: E402
from pycles_amd.ReferenceState import ReferenceState  # noqa: E402
from pycles_amd.ThermodynamicsSA import ThermodynamicsFactory  # noqa: E402

GRIDS = {"moist": dict(nx=8, ny=8, nz=128, gw=3, dx=50.0, dy=50.0, dz=20.0, moist=True),
         "dry": dict(nx=8, ny=8, nz=64, gw=3, dx=100.0, dy=100.0, dz=40.0, moist=False)}


def compute():
    out = {}
    for name, g in GRIDS.items():
        nl = cases.namelist(**g)
        Gr = Grid(nl)
        RS = ReferenceState(Gr)
        Th = ThermodynamicsFactory(nl)
        (cases.init_moist if g["moist"] else cases.init_dry)(Gr, RS, Th)
        for f in ("p0", "p0_half", "alpha0", "alpha0_half"):
            out[f"{name}_{f}"] = np.asarray(getattr(RS, f), dtype=np.float64)
        if g["moist"]:
            out["cc_T"] = np.asarray(Th.CC.T, dtype=np.float64)
            out["cc_pv"] = np.asarray(Th.CC.pv, dtype=np.float64)
    return out


if __name__ == "__main__":
    np.savez(os.path.join(HERE, "thermo_profiles.npz"), **compute())
    print("wrote", os.path.join(HERE, "thermo_profiles.npz"))
