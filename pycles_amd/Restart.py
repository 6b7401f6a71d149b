"""Checkpoint and restart of the device state (reference ``Restart.pyx``).

The reference pickles a per-rank dictionary (``Restart.pyx:92-152``) that the
operators fill in ``restart(Gr, Re)`` and read back in
``init_from_restart(Gr, Re)`` (prognostic values, ``PrognosticVariables.pyx:
269-357``; time stepping and reference state alike).  Here the same
dictionary is filled by the same operator methods, and written as one binary
file: an 8-byte magic, the length of a JSON header, the JSON header (names,
shapes, dtypes and byte offsets of the arrays, and every scalar), then the
arrays' raw bytes.  Reading it executes nothing from the file (no pickle).

The prognostic values are saved in the device layout and storage precision
(FP32 runs: the FP32 deviations and their FP64 level profiles), so a run that
stops after step n and restarts from the file continues bit for bit as the
uninterrupted run does.

    Re = Restart(namelist, Pa); Re.initialize()
    PV.restart(Gr, Re); TS.restart(Re); RS.restart(Gr, Re); Re.write(Pa)
    Re.read(Pa); PV.init_from_restart(Gr, Re); ...
"""
from __future__ import annotations

import glob
import json
import os
import shutil

import numpy as np

MAGIC = b"PYCLESR1"


def _flatten(data, prefix, arrays):
    """Split a nested dict into JSON-able scalars and named numpy arrays."""
    out = {}
    for k, v in data.items():
        key = f"{prefix}/{k}" if prefix else str(k)
        if isinstance(v, dict):
            out[k] = {"__dict__": _flatten(v, key, arrays)}
        elif isinstance(v, np.ndarray):
            if v.dtype.hasobject:
                raise TypeError(f"restart entry {key}: object arrays are not supported")
            arrays[key] = np.ascontiguousarray(v)
            out[k] = {"__array__": key}
        elif isinstance(v, (np.floating, np.integer, np.bool_)):
            out[k] = v.item()
        elif isinstance(v, (list, tuple)) and any(isinstance(x, np.ndarray) for x in v):
            raise TypeError(f"restart entry {key}: lists of arrays are not supported")
        else:
            json.dumps(v)  # raises for anything that is not plain data
            out[k] = v
    return out


def _unflatten(tree, arrays):
    out = {}
    for k, v in tree.items():
        if isinstance(v, dict) and "__dict__" in v:
            out[k] = _unflatten(v["__dict__"], arrays)
        elif isinstance(v, dict) and "__array__" in v:
            out[k] = arrays[v["__array__"]]
        else:
            out[k] = v
    return out


def write_file(path, data: dict):
    """Write a restart dictionary (nested dicts of JSON scalars and numpy
    arrays) to `path`."""
    arrays = {}
    tree = _flatten(data, "", arrays)
    table = []
    off = 0
    for name, a in arrays.items():
        table.append({"name": name, "dtype": a.dtype.str, "shape": list(a.shape), "offset": off,
                      "nbytes": int(a.nbytes)})
        off += (a.nbytes + 63) // 64 * 64
    header = json.dumps({"format": 1, "data": tree, "arrays": table}).encode()
    pad = (-(len(MAGIC) + 8 + len(header))) % 64
    header += b" " * pad
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(MAGIC)
        f.write(np.uint64(len(header)).tobytes())
        f.write(header)
        for a in arrays.values():
            f.write(a.tobytes(order="C"))
            f.write(b"\0" * ((a.nbytes + 63) // 64 * 64 - a.nbytes))
    os.replace(tmp, path)


def read_file(path) -> dict:
    """Read a file written by write_file; raises ValueError if it is not one."""
    with open(path, "rb") as f:
        if f.read(len(MAGIC)) != MAGIC:
            raise ValueError(f"{path}: not a pycles_amd restart file")
        hlen = int(np.frombuffer(f.read(8), dtype=np.uint64)[0])
        hdr = json.loads(f.read(hlen).decode())
        if hdr.get("format") != 1:
            raise ValueError(f"{path}: unsupported restart format {hdr.get('format')!r}")
        start = len(MAGIC) + 8 + hlen
        arrays = {}
        for ent in hdr["arrays"]:
            f.seek(start + ent["offset"])
            buf = f.read(ent["nbytes"])
            if len(buf) != ent["nbytes"]:
                raise ValueError(f"{path}: array {ent['name']} truncated")
            arrays[ent["name"]] = np.frombuffer(buf, dtype=np.dtype(ent["dtype"])).reshape(ent["shape"]).copy()
    return _unflatten(hdr["data"], arrays)


class Restart:
    """Restart bookkeeping with the reference's namelist keys and paths
    (``Restart.pyx:12-70``): ``restart.frequency`` (s, default 30),
    ``restart.delete_old``, ``restart.times_retained``, ``restart.init_from``
    with ``restart.input_path``; files under
    ``<output_root>Output.<simname>.<uuid[-5:]>/Restart/<int(t)>/``."""

    FILE = "0.pcr"  # one GPU = rank 0

    def __init__(self, namelist, Pa=None):
        nl = namelist or {}
        meta = nl.get("meta", {})
        self.uuid = str(meta.get("uuid", "00000"))
        try:
            outpath = os.path.join(str(nl["output"]["output_root"]) + "Output." + str(meta["simname"]) + "."
                                   + self.uuid[-5:])
            self.restart_path = os.path.join(outpath, "Restart")
        except KeyError:
            outpath = None
            self.restart_path = "./restart." + self.uuid[-5:]
        # restart files are written only when the namelist has a restart
        # section (the reference writes them in every run)
        self.enabled = "restart" in nl
        rs = nl.get("restart", {})
        self.frequency = float(rs.get("frequency", 30.0))
        self.delete_old = bool(rs.get("delete_old", False))
        tr = rs.get("times_retained", [])
        self.times_retained = [str(t) for t in (tr if isinstance(tr, list) else [tr])]
        self.is_restart_run = bool(rs.get("init_from", False))
        self.input_path = str(rs["input_path"]) if self.is_restart_run else None
        self._outpath = outpath

    def initialize(self):
        self.restart_data = {}
        self.last_restart_time = 0.0
        return self

    def write(self, Pa=None):
        """Write restart_data to <restart_path>/<int(last_restart_time)>/ and
        empty it; an existing directory is not overwritten (as the reference)."""
        self.restart_data["last_restart_time"] = self.last_restart_time
        os.makedirs(self.restart_path, exist_ok=True)
        path = os.path.join(self.restart_path, str(int(self.last_restart_time)))
        if self.delete_old:
            for d in glob.glob(os.path.join(self.restart_path, "*_recent")):
                os.rename(d, d[:-7] + "_old")
            for d in glob.glob(os.path.join(self.restart_path, "*_new")):
                os.rename(d, d[:-4] + "_recent")
            path += "_new"
        if os.path.exists(path):
            if Pa is not None:
                Pa.root_print("Restart path exists; for safety not overwriting.")
            self.free_memory()
            return None
        os.mkdir(path)
        write_file(os.path.join(path, self.FILE), self.restart_data)
        self.free_memory()
        if self.delete_old:
            for d in glob.glob(os.path.join(self.restart_path, "*_old")):
                stem = os.path.basename(d)[:-4]
                if stem in self.times_retained:
                    os.rename(d, os.path.join(self.restart_path, stem))
                else:
                    shutil.rmtree(d)
        return path

    def read(self, Pa=None):
        self.restart_data = read_file(os.path.join(self.input_path, self.FILE))
        if self.delete_old:
            os.rename(self.input_path, self.input_path + "_original")
        return self

    def free_memory(self):
        self.restart_data = {}

    def cleanup(self):
        """Strip the suffixes that delete_old and a restarted read left on the
        directories (reference ``Restart.cleanup``, Restart.pyx:166-186)."""
        for suffix in ("_original", "_recent", "_new"):
            for d in glob.glob(os.path.join(self.restart_path, "*" + suffix)):
                os.rename(d, d[:-len(suffix)])
