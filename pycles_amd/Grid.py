"""Grid description (reference ``Grid.pyx``, ``Csrc/grid.h``).

``Grid.dims`` carries the reference ``DimStruct`` fields for a single GPU
owning the whole domain (``nl == n``, ``indx_lo == 0``).  The device layout
used by the HIP kernels differs from the reference's flat ``[i][j][k]``
ghosted array: each field is ``[k][j][i]`` (x fastest) over ``nzg = nz+2*gw``
levels and the ``ny*nx`` interior columns; x and y are periodic by index
wrap, so only z carries ghost levels.  ``to_device_layout`` /
``from_device_layout`` convert between the two at ingress and egress.
"""
from __future__ import annotations

import numpy as np


class DimStruct:
    pass


class Grid:
    def __init__(self, namelist, Parallel=None):
        g = namelist["grid"]
        d = DimStruct()
        d.dims = int(g.get("dims", 3))
        d.dx = [float(g["dx"]), float(g["dy"]), float(g["dz"])]
        d.dxi = [1.0 / v for v in d.dx]
        d.gw = int(g["gw"])
        d.n = [int(g["nx"]), int(g["ny"]), int(g["nz"])]
        d.ng = [n + 2 * d.gw for n in d.n]
        d.nl = list(d.n)
        d.nlg = list(d.ng)
        d.indx_lo = [0, 0, 0]
        d.indx_lo_g = [0, 0, 0]
        d.npd = d.n[0] * d.n[1] * d.n[2]
        d.npl = d.npd
        d.npg = d.nlg[0] * d.nlg[1] * d.nlg[2]
        # device layout
        d.nzg = d.n[2] + 2 * d.gw
        d.plane = d.n[0] * d.n[1]
        d.npdev = d.plane * d.nzg
        self.dims = d
        self.compute_coordinates()

    def compute_coordinates(self):
        """Cell edges and centres (``Grid.pyx:125-173``)."""
        d = self.dims
        idx = [np.arange(-d.gw, n + d.gw, dtype=np.float64) for n in d.n]
        self.x, self.y, self.z = [(i + 1.0) * dx for i, dx in zip(idx, d.dx)]
        self.x_half, self.y_half, self.z_half = [(i + 0.5) * dx for i, dx in zip(idx, d.dx)]
        self.zl = self.z
        self.zl_half = self.z_half

    def restart(self, Re):
        """Grid.restart (Grid.pyx:184-203)."""
        d = self.dims
        Re.restart_data["Gr"] = {"dims": d.dims, "n": np.array(d.n), "ng": np.array(d.ng), "nl": np.array(d.nl),
                                 "nlg": np.array(d.nlg), "indx_lo_g": np.array(d.indx_lo_g),
                                 "indx_lo": np.array(d.indx_lo), "gw": d.gw, "dx": np.array(d.dx)}

    # ---- layout conversion (host numpy) ---------------------------------
    def to_device_layout(self, ref_field: np.ndarray) -> np.ndarray:
        """Reference ghosted [i][j][k] -> device [k][j][i] (x/y interior)."""
        d = self.dims
        gw = d.gw
        f = np.asarray(ref_field).reshape(d.nlg)
        inner = f[gw:gw + d.n[0], gw:gw + d.n[1], :]
        return np.ascontiguousarray(inner.transpose(2, 1, 0))

    def from_device_layout(self, dev_field: np.ndarray) -> np.ndarray:
        """Device [k][j][i] -> reference ghosted [i][j][k]; x/y ghosts are
        the periodic images (what update_all_bcs leaves there)."""
        d = self.dims
        gw = d.gw
        f = np.asarray(dev_field).reshape(d.nzg, d.n[1], d.n[0]).transpose(2, 1, 0)
        f = np.pad(f, ((gw, gw), (gw, gw), (0, 0)), mode="wrap")
        return np.ascontiguousarray(f)
