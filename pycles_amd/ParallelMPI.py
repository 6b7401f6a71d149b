"""Single-GPU counterpart of the reference's ParallelMPI object and its
statistics writer.

The reference threads a ``ParallelMPI`` instance through every operator for
logging (``root_print``), fatal errors (``kill`` -> MPI_Abort) and the
horizontal reductions of its statistics (``ParallelMPI.pyx:17-540``).  One
GPU owns the whole domain here, so the reductions are device kernels of the
library (``pycles_horizontal_stats``, one pass giving the mean, the means of
squares and cubes, the maximum, the minimum and, with a mask, the
conditional means of a level), and ``kill`` raises instead of aborting.

``NetCDFIO_Stats`` keeps the reference's writer interface (``add_profile``,
``add_ts``, ``write_profile``, ``write_ts``, ``open_files``/``close_files``,
``write_simulation_time``).  netCDF4 is not available, so the profiles and
time series are kept per output time in memory and written, on
``close_files``, to an ``.npz`` file (plain arrays, readable with
``numpy.load``) under the reference's ``stats_io`` output path.
"""
from __future__ import annotations

import os

import numpy as np


class ParallelMPI:
    rank = 0
    size = 1

    def __init__(self, namelist=None, verbose=False):
        self.verbose = verbose
        self._ws = {}

    def root_print(self, msg):
        if self.verbose:
            print(msg, flush=True)

    def kill(self, msg="killed"):
        raise RuntimeError(msg)

    def barrier(self):
        return None

    # ---- horizontal reductions (device) --------------------------------
    def _workspace(self, Gr, kind):
        import torch
        from . import _capi
        from . import _device as D
        key = (kind, tuple(Gr.dims.n), Gr.dims.gw)
        ws = self._ws.get(key)
        if ws is None:
            dims = D.dims_of(Gr)
            nbytes = (_capi.horizontal_stats_workspace_bytes(dims) if kind == "stats"
                      else _capi.cloud_stats_workspace_bytes(dims))
            ws = torch.empty(max(int(nbytes), 8), dtype=torch.uint8, device=D.require_gpu())
            self._ws[key] = ws
        return ws

    def horizontal_stats(self, Gr, a, precision, base_a=None, b=None, base_b=None, mask=None):
        """Per-level statistics of a device field (optionally times a second
        one, optionally masked): an FP64 numpy array [6][nzg] of the mean,
        mean of squares, mean of cubes, maximum, minimum and mask sum."""
        import torch
        from . import _capi
        from . import _device as D
        out = torch.empty((6, Gr.dims.nzg), dtype=torch.float64, device=D.require_gpu())
        p = lambda t: None if t is None else t.data_ptr()  # noqa: E731
        _capi.horizontal_stats(D.dims_of(Gr), p(a), p(base_a), p(b), p(base_b), p(mask), precision,
                               out.data_ptr(), self._workspace(Gr, "stats").data_ptr(), D.stream())
        return out.cpu().numpy()

    def HorizontalMean(self, Gr, a, precision, base=None):
        return self.horizontal_stats(Gr, a, precision, base)[0]

    def HorizontalMeanofSquares(self, Gr, a, b, precision, base_a=None, base_b=None):
        return self.horizontal_stats(Gr, a, precision, base_a, b, base_b)[0]

    def HorizontalMaximum(self, Gr, a, precision, base=None):
        return self.horizontal_stats(Gr, a, precision, base)[3]

    def HorizontalMinimum(self, Gr, a, precision, base=None):
        return self.horizontal_stats(Gr, a, precision, base)[4]

    def HorizontalMeanConditional(self, Gr, a, mask, precision, base=None):
        return self.horizontal_stats(Gr, a, precision, base, mask=mask)[0]

    def cloud_stats(self, Gr, RS, ql, precision):
        """ThermodynamicsSA.liquid_stats on the device: (cloud fraction
        profile [nz], cloud cover, cloud base, cloud top, liquid water path)."""
        import torch
        from . import _capi
        from . import _device as D
        nz = Gr.dims.n[2]
        out = torch.empty(nz + 4, dtype=torch.float64, device=D.require_gpu())
        r = D.device_ref(RS)
        _capi.cloud_stats(D.dims_of(Gr), r.t["rho0_half"].data_ptr(), ql.data_ptr(), precision, out.data_ptr(),
                          self._workspace(Gr, "cloud").data_ptr(), D.stream())
        o = out.cpu().numpy()
        return o[:nz], float(o[nz]), float(o[nz + 1]), float(o[nz + 2]), float(o[nz + 3])


class NetCDFIO_Stats:
    """Statistics writer with the reference's interface (NetCDFIO.pyx:19-260);
    netCDF4 being absent, each output time is one .npz file of profiles
    (`profile/<name>`) and time-series values (`ts/<name>`)."""

    def __init__(self):
        self.frequency = 1e30
        self.last_output_time = 0.0
        self.profiles = {}
        self.timeseries = {}
        self.path = None
        self._current = None
        self.written = []

    def initialize(self, namelist, Gr=None, Pa=None):
        """Output path and frequency from namelist['stats_io'] (as the
        reference, NetCDFIO.pyx:27-62); without that section nothing is written."""
        sio = (namelist or {}).get("stats_io")
        if not sio:
            return self
        self.frequency = float(sio.get("frequency", 1e30))
        meta = namelist.get("meta", {})
        root = str(namelist.get("output", {}).get("output_root", "./"))
        uuid = str(meta.get("uuid", "00000"))
        self.path = os.path.join(root + "Output." + str(meta.get("simname", "sim")) + "." + uuid[-5:],
                                 str(sio.get("stats_dir", "stats")))
        os.makedirs(self.path, exist_ok=True)
        return self

    def add_profile(self, name, Gr=None, Pa=None, *a, **k):
        self.profiles.setdefault(name, [])

    def add_reference_profile(self, name, Gr=None, Pa=None, *a, **k):
        self.profiles.setdefault("reference/" + name, [])

    def add_ts(self, name, Gr=None, Pa=None, *a, **k):
        self.timeseries.setdefault(name, [])

    def open_files(self, Pa=None):
        self._current = {}

    def write_simulation_time(self, t, Pa=None):
        if self._current is None:
            self._current = {}
        self._current["t"] = np.array(t, dtype=np.float64)

    def write_profile(self, name, data, Pa=None):
        if self._current is None:
            self._current = {}
        self._current["profile/" + name] = np.asarray(data, dtype=np.float64).copy()

    def write_reference_profile(self, name, data, Pa=None):
        self.write_profile("reference/" + name, data, Pa)

    def write_ts(self, name, data, Pa=None):
        if self._current is None:
            self._current = {}
        self._current["ts/" + name] = np.array(data, dtype=np.float64)

    def close_files(self, Pa=None):
        """Finish one output time: keep it and, with an output path, write it."""
        cur = self._current or {}
        self._current = None
        for k, v in cur.items():
            if k.startswith("profile/"):
                self.profiles.setdefault(k[8:], []).append(v)
            elif k.startswith("ts/"):
                self.timeseries.setdefault(k[3:], []).append(float(v))
        if self.path is not None:
            t = float(cur.get("t", len(self.written)))
            fn = os.path.join(self.path, f"stats_{t:.6f}.npz")
            np.savez(fn, **cur)
            self.written.append(fn)
        return cur
