"""Benchmark: Mcell-updates/s per SSP-RK3 step of the dynamical core.

Workload: the moist 512x512x256 configuration (dx=dy=50 m, dz=20 m, WENO5
scalars and momentum, saturation adjustment, FFT + tridiagonal pressure,
dt=1 s), seeded synthetic state.  A step is the full SSP-RK3 step (three
stages of thermodynamics, advection, RK update with boundaries, pressure
projection; CFL once).  Inputs are resident in HBM before timing starts.

The same initial state is run in plain FP64 and in the FP32 form; ``value``
is the FP32-form throughput, ``fp64`` the FP64 throughput, ``speedup`` their
ratio, and ``fp32_vs_fp64`` the max / RMS relative difference per field after
warmup+steps steps.  ``roofline`` is for the dominant operator of the FP32
form, timed with HIP events on the working stream during the timed steps;
``cpu_baseline`` times the CPU oracle (FP64 C restatement, OpenMP) on a
bounded sample on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_state(cfg_name, nx=None, ny=None, nz=None):
    from pycles_amd import cases
    from pycles_amd.Grid import Grid
    from pycles_amd.ReferenceState import ReferenceState
    from pycles_amd.ThermodynamicsSA import ThermodynamicsFactory
    c = dict(cases.CONFIGS[cfg_name])
    if nx:
        c.update(nx=nx, ny=ny, nz=nz)
    nl = cases.namelist(**c)
    Gr = Grid(nl)
    RS = ReferenceState(Gr)
    Th = ThermodynamicsFactory(nl)
    if c["moist"]:
        fields = cases.init_moist(Gr, RS, Th, seed=c["seed"], sigma=c["sigma"], cloud=c.get("cloud"))
    else:
        fields = cases.init_dry(Gr, RS, Th, seed=c["seed"], sigma=c["sigma"])
    return c, nl, Gr, RS, Th, fields


def run_form(nl, Gr, RS, Th, fields, precision, steps, warmup, learned=None):
    import torch
    from pycles_amd.Simulation3d import Simulation3d
    nl = json.loads(json.dumps(nl))
    nl["mi355x"]["precision"] = precision
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)
    if learned is not None:
        sim.Th.mlp = learned
    sim.begin_run()
    for _ in range(warmup):
        sim.step()
    torch.cuda.synchronize()
    sim.timers = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = sim.phase_times()
    sim.timers = None
    return sim, (t1 - t0) * 1e3 / steps, phases


def field_errors(sim32, sim64):
    import torch
    out = {}
    pairs = [(n, "PV") for n in ("u", "v", "w", "s", "qt")] + [(n, "DV") for n in ("temperature", "ql")]
    for n, kind in pairs:
        if kind == "PV":
            a = sim32.PV.get_device_layout(n)
            b = sim64.PV.get_device_layout(n)
        else:
            a = sim32.DV.get_device_layout(n)
            b = sim64.DV.get_device_layout(n)
        gw = sim64.Gr.dims.gw
        a, b = a[gw:-gw], b[gw:-gw]
        scale = np.max(np.abs(b))
        d = np.abs(a - b)
        if n in ("temperature", "qt", "s"):
            rel = d / np.abs(b)
            out[n] = {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel * rel)))}
        else:
            out[n] = {"max_abs": float(d.max()), "rms_abs": float(np.sqrt(np.mean(d * d))),
                      "max_rel_to_max": float(d.max() / scale) if scale else 0.0}
    del torch
    return out


def cpu_baseline():
    """FP64 CPU oracle, one SSP-RK3 step of the moist configuration on a
    128x128x64 sample of the same spacing and profiles, on the host cores."""
    from oracle import oracle as O
    from tests.helpers import oracle_sim
    threads = min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    c, nl, Gr, RS, Th, fields = build_state("moist512", 128, 128, 64)
    sim = oracle_sim(Gr, RS, Th, fields, True, dt=c["dt"])
    sim.begin_run()
    t0 = time.perf_counter()
    sim.step()
    dt = time.perf_counter() - t0
    n = 128 * 128 * 64
    return {"value": round(n / dt / 1e6, 3), "unit": "Mcell-updates/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": "port",
            "sample": "one SSP-RK3 step of the moist configuration on a 128x128x64 subdomain (1.05 M cells), "
                      f"FP64 CPU oracle with OpenMP, {dt:.2f} s"}


# Algorithmic HBM bytes per cell for each operator in one stage (field passes
# x storage bytes); see DESIGN.md "roofline accounting".
PASSES = {"thermo": 2 + 1 + 1 + 8, "advection": 3 * 15, "rk3": 5 * 4, "pressure": 22}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="moist512")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU "
              "decomposition is a design only).", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench needs a GPU"
    log(f"[bench] building {args.config} state")
    c, nl, Gr, RS, Th, fields = build_state(args.config)
    ncell = c["nx"] * c["ny"] * c["nz"]
    log("[bench] FP32 form")
    sim32, ms32, ph32 = run_form(nl, Gr, RS, Th, fields, "f32", args.steps, args.warmup)
    log(f"[bench] FP32 {ms32:.3f} ms/step {ph32}")
    res64 = None
    if not args.no_fp64:
        log("[bench] FP64 form")
        sim64, ms64, ph64 = run_form(nl, Gr, RS, Th, fields, "f64", args.steps, args.warmup)
        log(f"[bench] FP64 {ms64:.3f} ms/step {ph64}")
        errs = field_errors(sim32, sim64)
        res64 = {"ms_per_step": round(ms64, 4), "value": round(ncell / (ms64 * 1e-3) / 1e6, 2),
                 "phases_ms_per_step": {k: round(v / args.steps, 4) for k, v in ph64.items()}}
    value = ncell / (ms32 * 1e-3) / 1e6
    # dominant operator of the FP32 form
    dom = max(ph32, key=ph32.get)
    per_launch_ms = ph32[dom] / (3 * args.steps)
    bytes_per_launch = PASSES[dom] * ncell * 4
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "operator": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "algorithmic_bytes_per_launch": bytes_per_launch, "ms_per_launch": round(per_launch_ms, 4)}
    out = {"metric": "Mcell-updates/s per RK3 step: FP32+learned vs plain FP64 at <=1e-5 rel err",
           "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms32, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded BOMEX-like moist state)",
           "config": {"workload": f"{args.config}: {c['nx']}x{c['ny']}x{c['nz']} moist SA, WENO{c['order'][0]}/"
                                  f"{c['order'][1]}, FFT+tridiagonal pressure, dt={c['dt']} s",
                      "cells": ncell},
           "phases_ms_per_step": {k: round(v / args.steps, 4) for k, v in ph32.items()},
           "roofline": roof}
    if res64:
        out["fp64"] = res64
        out["speedup_vs_fp64"] = round(res64["ms_per_step"] / ms32, 3)
        out["fp32_vs_fp64"] = errs
    if not args.no_cpu:
        log("[bench] CPU baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
