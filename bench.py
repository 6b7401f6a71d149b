"""Benchmark: Mcell-updates/s per SSP-RK3 step of the dynamical core.

Workload: the moist 512x512x256 configuration (dx=dy=50 m, dz=20 m, WENO5
scalars and momentum, saturation adjustment with a saturated cloud layer,
FFT + tridiagonal pressure, dt=1 s), seeded synthetic state.  A step is the
full SSP-RK3 step (three stages of thermodynamics, scalar and momentum
advection, RK update with boundaries, pressure projection; CFL once).  Inputs
are resident in HBM before timing starts.

Forms, all from the same initial state:
  value      FP32 + learned guess (the network is trained here first, on the
             GPU, outside the timed region)
  fp32_plain FP32 with the reference's secant
  fp64       plain FP64 (the reference's arithmetic)
Errors of the FP32 + learned form against FP64 are measured after 1 and after
`--error-steps` steps.  Every single-kernel operator is timed with HIP events
on the working stream during the timed steps; `roofline` is for the one with
the largest average launch time.  `cpu_baseline` times the CPU oracle (FP64 C
restatement, OpenMP) for one step of the same workload on the host cores.

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
METRIC = "Mcell-updates/s per RK3 step: FP32+learned vs plain FP64 at <=1e-5 rel err"

# Algorithmic HBM bytes per interior cell of one launch, FP32 storage (4 B):
#   thermo  reads s, qt; writes T, qv, ql, alpha, buoyancy, theta_rho, N^2,
#           thetali and the w tendency                          11 x 4
#   adv_x   reads u, v, w, s, qt; writes the s, qt, u, v tendencies; reads
#           and writes the w tendency                            11 x 4
#   adv_y   reads the five fields; reads and writes five tendencies  15 x 4
#   adv_z   as adv_y                                             15 x 4
#   rk3     per variable: value and tendency in, value out, and the stage
#           copy (written in stage 0, read in stages 1, 2); five variables  20 x 4
#           (only when the RK3 update is not fused into the z pass)
#   adv_z_rk3  the z pass with the RK3 stage fused: reads the five fields and
#           their tendencies, writes the five new values (stage 0: 15 x 4);
#           stages 1, 2 also read the stage copies (20 x 4); the launch time
#           is the mean over the three stages, so are the bytes   (15 + 20 + 20) / 3 x 4
#   adv_yz, adv_yz_rk3  the y and z passes as one kernel (FP32, WENO5): the
#           bytes of the z pass alone (the y pass's own 15 x 4 disappear)
KERNEL_BYTES = {"thermo": 44, "adv_x": 44, "adv_y": 60, "adv_z": 60, "rk3": 80, "adv_z_rk3": 220 / 3,
                "adv_yz": 60, "adv_yz_rk3": 220 / 3}
# Algorithmic flops per interior cell of the advection passes: WENO5 scalar and
# momentum advection is ~2265 flop per cell per stage over the three directions
# (2 scalar and 9 momentum face fluxes, interpolation and divergence), split
# evenly over the directional passes; the fused RK3 stage adds 4 flop per
# variable.  VALU peak: FP32 vector with packed FMA (the kernels issue no
# packed FP32, so half of it is their ceiling).
KERNEL_FLOPS = {"adv_x": 755, "adv_y": 755, "adv_z": 755, "adv_z_rk3": 775, "adv_yz": 1510, "adv_yz_rk3": 1530}
VALU_PEAK_TFLOPS = 157.3
CUS, CLOCK_GHZ = 256, 2.4
# Field passes per moist SSP-RK3 step (thermo 6, scalar advection 7,
# momentum advection 7, RK update 20, projection 22 per stage; x3)
STEP_PASSES = 186


def kernel_names(so, mo):
    """Kernel instantiations timed as each phase (scalar order so, momentum order mo)."""
    adv = lambda d: f"adv::adv_{d}<PolicyF32,{so},{mo},2,true>"  # noqa: E731
    return {"thermo": "thermo_sa_f32_kernel<true,true>", "adv_x": adv("x"), "adv_y": adv("y"), "adv_z": adv("z"),
            "rk3": "rk3_kernel<float,...>", "adv_z_rk3": adv("z") + " (RK3 fused)", "adv_yz": adv("yz"),
            "adv_yz_rk3": adv("yz") + " (RK3 fused)"}


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


def make_sim(nl, Gr, RS, Th, fields, precision, learned=None):
    from pycles_amd.Simulation3d import Simulation3d
    nl = json.loads(json.dumps(nl))
    nl["mi355x"]["precision"] = precision
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=fields)
    sim.Th.mlp = learned
    return sim


def run_form(nl, Gr, RS, Th, fields, precision, steps, warmup, learned=None):
    """Warm up, then time exactly `steps` steps between two synchronisations."""
    import torch
    sim = make_sim(nl, Gr, RS, Th, fields, precision, learned)
    sim.begin_run()
    for _ in range(warmup):
        sim.step()
    torch.cuda.synchronize()
    counters = None
    if hasattr(sim.Th, "counters"):
        sim.Th.counters.zero_()
    sim.timers = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = sim.phase_times()
    sim.timers = None
    if hasattr(sim.Th, "counters"):
        counters = sim.Th.counters.cpu().numpy().tolist()
    sim.Th.mlp = None
    return sim, (t1 - t0) * 1e3 / steps, phases, counters


def field_errors(sim32, sim64):
    out = {}
    pairs = [(n, "PV") for n in ("u", "v", "w", "s", "qt")] + [(n, "DV") for n in ("temperature", "ql")]
    gw = sim64.Gr.dims.gw
    for n, kind in pairs:
        src32 = sim32.PV if kind == "PV" else sim32.DV
        src64 = sim64.PV if kind == "PV" else sim64.DV
        a = src32.get_device_layout(n)[gw:-gw]
        b = src64.get_device_layout(n)[gw:-gw]
        d = np.abs(a - b)
        if n in ("temperature", "qt", "s"):
            rel = d / np.abs(b)
            # where the error sits: the 99.99th percentile and the share of
            # cells beyond the 1e-5 gate, beside the maximum and RMS
            out[n] = {"max_rel": float(rel.max()), "rms_rel": float(np.sqrt(np.mean(rel * rel))),
                      "p9999_rel": float(np.quantile(rel, 0.9999)),
                      "frac_above_1e-5": float(np.count_nonzero(rel > 1e-5)) / rel.size}
        else:
            scale = float(np.max(np.abs(b)))
            out[n] = {"max_abs": float(d.max()), "rms_abs": float(np.sqrt(np.mean(d * d))),
                      "max_rel_to_max": float(d.max() / scale) if scale else 0.0}
    return out


def perturbed_fields(fields, kind):
    """The initial state with a perturbation whose growth sets the floor of any
    comparison with the FP64 run:
      ulp64  every nonzero prognostic value moved up by one FP64 ulp;
      fp32   the state exactly as the FP32 form stores it (u, v, w rounded to
             FP32; s, qt as level mean + FP32-rounded deviation), kept in FP64."""
    out = {}
    for n, f in fields.items():
        if kind == "ulp64":
            out[n] = np.where(f != 0.0, np.nextafter(f, np.inf), 0.0)
        elif n in ("s", "qt"):
            base = f.mean(axis=(1, 2), keepdims=True)
            out[n] = base + (f - base).astype(np.float32).astype(np.float64)
        else:
            out[n] = f.astype(np.float32).astype(np.float64)
    return out


class FP32Storage:
    """After every step, round an FP64 run's prognostic state to what the FP32
    form stores: u, v, w to FP32, s and qt to their level profile (the
    initial horizontal mean) plus an FP32-rounded deviation.  The run then
    carries FP32 storage's rounding, step after step, with FP64 arithmetic:
    how far it moves from the plain FP64 run is the floor of any comparison
    of an FP32-storage form with FP64."""

    def __init__(self, sim, fields):
        import torch
        self.sim = sim
        self.base = {n: torch.from_numpy(np.asarray(fields[n], dtype=np.float64).mean(axis=(1, 2))).to(
            sim.PV.values.device)[:, None, None] for n in ("s", "qt")}

    def round(self):
        PV = self.sim.PV
        for n in PV.index_name:
            f = PV.field(n)
            if n in self.base:
                b = self.base[n]
                f.copy_(b + (f - b).float().double())
            else:
                f.copy_(f.float().double())


def error_study(nl, Gr, RS, Th, fields, learned, nsteps, floors=True,
                floor_kinds=("ulp64", "fp32", "storage", "secant")):
    """FP32 + learned against FP64 from the same state after 1 and nsteps
    steps; with `floors`, also FP64 runs from the perturbed states of
    `perturbed_fields`, an FP64 run rounded to FP32 storage after every
    step (FP32Storage), and an FP64 run whose saturation-adjustment secant
    stops at |dT| < 1e-7 K instead of the reference's 1e-3 K (how far the
    reference's own stopping point moves the run), each against the same
    FP64 run (their noise floors)."""
    import torch
    runs = {"fp32_learned": make_sim(nl, Gr, RS, Th, fields, "f32", learned),
            "fp64": make_sim(nl, Gr, RS, Th, fields, "f64")}
    storage = None
    if floors:
        for kind in ("ulp64", "fp32"):
            if kind in floor_kinds:
                runs["fp64_from_" + kind] = make_sim(nl, Gr, RS, Th, perturbed_fields(fields, kind), "f64")
        if "storage" in floor_kinds:
            runs["fp64_fp32_storage"] = make_sim(nl, Gr, RS, Th, perturbed_fields(fields, "fp32"), "f64")
            storage = FP32Storage(runs["fp64_fp32_storage"], fields)
        if "secant" in floor_kinds:
            runs["fp64_secant_1e-7"] = make_sim(nl, Gr, RS, Th, fields, "f64")
    out = {"errors_vs_fp64": {}, "noise_floor_fp64": {}}
    for name, sim in runs.items():
        sim.Th.mlp = learned if name == "fp32_learned" else None
        sim.Th.secant_tol = 1e-7 if name == "fp64_secant_1e-7" else 0.0
        sim.begin_run()
    # the runs share Th: its network and secant tolerance are set per run
    for n in range(1, nsteps + 1):
        for name, sim in runs.items():
            sim.Th.mlp = learned if name == "fp32_learned" else None
            sim.Th.secant_tol = 1e-7 if name == "fp64_secant_1e-7" else 0.0
            sim.step()
        if storage is not None:
            storage.round()
        if n == 1 or n == nsteps:
            torch.cuda.synchronize()
            key = f"after_{n}_steps"
            out["errors_vs_fp64"][key] = field_errors(runs["fp32_learned"], runs["fp64"])
            for name in runs:
                if name.startswith("fp64_") and name != "fp64":
                    out["noise_floor_fp64"].setdefault(name, {})[key] = field_errors(runs[name], runs["fp64"])
            e = out["errors_vs_fp64"][key]
            log(f"[bench] errors after {n} steps: T {e['temperature']} qt {e['qt']}")
            for name, fl in out["noise_floor_fp64"].items():
                log(f"[bench] floor {name} after {n} steps: T {fl[key]['temperature']} qt {fl[key]['qt']}")
    runs["fp32_learned"].Th.mlp = None
    runs["fp32_learned"].Th.secant_tol = 0.0
    return out


def error_gate(errors, tol=1e-5, floor=None, secant_floor=None):
    """The metric's gate: T and qt within tol (max relative) of FP64.  With
    `floor` (the FP32-storage run's errors), also the gate against what FP32
    storage alone gives: within max(tol, 2 x floor).  With `secant_floor`
    (the errors of an FP64 run whose saturation adjustment converges to
    1e-7 K instead of stopping at the reference's 1e-3 K), the gate against
    the reference method's own resolution: within max(tol, 2 x that)."""
    out = {"tolerance": tol, "fields": ["temperature", "qt"], "failing": []}
    gates = {}
    if floor:
        gates["floor_gate"] = ("max_rel <= max(tol, 2 x FP64 rounded to FP32 storage every step)", floor)
    if secant_floor:
        gates["reference_stopping_gate"] = (
            "max_rel <= max(tol, 2 x FP64 with its secant converged to 1e-7 K instead of the reference's 1e-3 K)",
            secant_floor)
    for g, (rule, _) in gates.items():
        out[g] = {"rule": rule, "failing": []}
    for key, e in errors.items():
        for f in out["fields"]:
            v = e[f]["max_rel"]
            if not v <= tol:
                out["failing"].append(f"{f} {key}: {v:.3g}")
            for g, (_, fl) in gates.items():
                if key in fl:
                    lim = max(tol, 2.0 * fl[key][f]["max_rel"])
                    if not v <= lim:
                        out[g]["failing"].append(f"{f} {key}: {v:.3g} > {lim:.3g}")
    out["pass"] = not out["failing"]
    for g in gates:
        out[g]["pass"] = not out[g]["failing"]
    return out


def train_guess(Th, pairs_log2=22, epochs=40):
    import torch
    from pycles_amd.LearnedGuess import SaturationGuess
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = SaturationGuess(Th.CC, n_pairs=1 << pairs_log2, epochs=epochs).fit()
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    held = g.pairs(1 << 16, 424242)
    T = held[:, 3].cpu().numpy()
    y = g.predict(held).cpu().numpy().astype(np.float64)
    err = y - T
    info = {"pairs": g.n, "epochs": g.epochs, "batch": g.batch, "train_s": round(t_train, 3),
            "loss_first_epoch": g.loss_history[0], "loss_last_epoch": g.loss_history[-1],
            "heldout": "65536 saturated pairs, packed bf16 inference",
            "heldout_rms_K": float(np.sqrt(np.mean(err ** 2))), "heldout_max_K": float(np.abs(err).max()),
            "inference": "bf16 MFMA 32x32x16, FP32 accumulate, fused into the thermodynamics kernel"}
    return g, info


def load_guess(Th, path):
    from pycles_amd.LearnedGuess import SaturationGuess
    with np.load(path, allow_pickle=False) as z:
        st = {"params": z["params"], "norm": z["norm"]}
    g = SaturationGuess(Th.CC)
    g.load_state_dict(st)
    return g, {"loaded_from": os.path.basename(path),
               "inference": "bf16 MFMA 32x32x16, FP32 accumulate, fused into the thermodynamics kernel"}


def per_kernel(phases, steps):
    out = {}
    for k, nbytes in KERNEL_BYTES.items():
        if k not in phases:
            continue
        ms = phases[k] / (3 * steps)
        out[k] = {"ms_per_launch": round(ms, 4), "bytes_per_cell": round(nbytes, 2)}
    return out


def hbm_ceiling(nbytes=1 << 30, reps=5):
    """The HBM rate the device reaches with the library's streaming kernels
    (16-byte loads and stores, grid-stride, plain and non-temporal) on
    `nbytes` buffers: copy, read-only and write-only, over a range of grid
    sizes; the best average launch time of `reps` per variant, timed with
    HIP events on the launching stream."""
    import torch
    from pycles_amd import _capi
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    st = torch.cuda.current_stream()
    out = {"bytes": nbytes, "variants": {}}
    best = {}
    for mode, name, moved in ((0, "copy", 2), (1, "read", 1), (2, "write", 1)):
        for nt in (0, 4):
            for blocks in (1024, 4096, 16384, 65536):
                _capi.stream_kernel(mode + nt, a.data_ptr(), b.data_ptr(), nbytes, blocks, st.cuda_stream)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(st)
                for _ in range(reps):
                    _capi.stream_kernel(mode + nt, a.data_ptr(), b.data_ptr(), nbytes, blocks, st.cuda_stream)
                e1.record(st)
                e1.synchronize()
                ms = e0.elapsed_time(e1) / reps
                gbs = moved * nbytes / (ms * 1e-3) / 1e9
                key = f"{name}{'_nt' if nt else ''}_g{blocks}"
                out["variants"][key] = round(gbs, 1)
                if gbs > best.get(name, (0.0, ""))[0]:
                    best[name] = (gbs, key)
    out["best_GBs"] = {k: round(v[0], 1) for k, v in best.items()}
    out["best_variant"] = {k: v[1] for k, v in best.items()}
    del a, b
    return out


def cpu_baseline(c, Gr, RS, Th, fields):
    """FP64 CPU oracle: one SSP-RK3 step of the same workload on the host cores."""
    from tests.helpers import oracle_sim
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sim = oracle_sim(Gr, RS, Th, fields, True, dt=c["dt"])
    sim.begin_run()
    t0 = time.perf_counter()
    sim.step()
    dt = time.perf_counter() - t0
    n = c["nx"] * c["ny"] * c["nz"]
    return {"value": round(n / dt / 1e6, 3), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"one SSP-RK3 step of the full {c['nx']}x{c['ny']}x{c['nz']} workload, FP64 CPU oracle "
                      f"(C, OpenMP, {threads} threads): {dt:.2f} s"}


def measured_traffic():
    """HBM bytes and VALU wavefront instructions per launch of each kernel,
    from committed rocprofv3 --pmc runs."""
    p = os.path.join(ROOT, "profiles", "traffic_moist512.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def cpu_reference_c1(repeats=5):
    """The dry 64^3 configuration's SSP-RK3 step on one CPU core with the
    reference's own compiled kernels where they build (momentum advection and
    the boundary fill, oracle/_ref) and the FP64 oracle for the rest (scalar
    advection, dry thermodynamics, RK update, pressure projection): the
    reference's CPU path as one MPI rank would run it.  Median of `repeats`."""
    os.environ["OMP_NUM_THREADS"] = "1"
    import ctypes as C
    from oracle import oracle as O
    from oracle import reflib
    from tests.helpers import oracle_sim
    c, nl, Gr, RS, Th, fields = build_state("dry64")
    d = Gr.dims
    rd = reflib.make_dims(d.n[0], d.n[1], d.n[2], d.gw, *d.dx)
    R = reflib.lib()
    times = []
    for _ in range(repeats):
        sim = oracle_sim(Gr, RS, Th, fields, False, dt=c["dt"])
        r = sim.ref.arrays  # p0, p0_half, alpha0, alpha0_half, rho0, rho0_half

        def momentum():
            for a in range(3):
                for b in range(3):
                    R.compute_advective_tendencies_m(C.byref(rd), reflib.ptr(r[4]), reflib.ptr(r[5]),
                                                     reflib.ptr(r[2]), reflib.ptr(r[3]),
                                                     reflib.ptr(sim.pv("uvw"[a])), reflib.ptr(sim.pv("uvw"[b])),
                                                     reflib.ptr(sim.pt("uvw"[a])), a, b, c["order"][1])

        def bcs():
            for n in sim.pv_names:
                reflib.update_all_bcs(rd, sim.pv(n), O.BC[n])

        sim.momentum_advection_update = momentum
        sim.update_all_bcs = bcs
        sim.begin_run()
        t0 = time.perf_counter()
        sim.step()
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    n = c["nx"] * c["ny"] * c["nz"]
    return {"metric": "Mcell-updates/s per RK3 step (reference CPU path)", "value": round(n / t / 1e6, 4),
            "unit": "Mcell-updates/s", "cores": 1, "kind": "reference",
            "sample": f"dry 64^3, WENO5, one SSP-RK3 step, dt = 1 s, median of {repeats}: {t:.3f} s; "
                      "momentum advection and boundary fill from the reference's compiled C (oracle/_ref), "
                      "scalar advection, dry thermodynamics, RK update and pressure from the oracle (C, 1 thread)",
            "seconds_per_step": t, "cpu": _cpu_name()}


def _cpu_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="moist512")
    ap.add_argument("--error-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--guess-pairs-log2", type=int, default=22)
    ap.add_argument("--guess-epochs", type=int, default=40)
    ap.add_argument("--newton-steps", type=int, default=None,
                    help="Newton corrections after the learned guess (default: the thermodynamics' setting)")
    ap.add_argument("--no-floor", action="store_true", help="skip the FP64 noise-floor runs of the error study")
    ap.add_argument("--save-guess", default=None, help="write the trained network (npz) here")
    ap.add_argument("--load-guess", default=None, help="use this trained network (npz) instead of training")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device copy (the reference kernel for HBM counter scaling)")
    ap.add_argument("--cpu-reference-c1", action="store_true",
                    help="only time the reference's CPU path on the dry 64^3 configuration, one core (no GPU)")
    args = ap.parse_args()
    if args.cpu_reference_c1:
        print(json.dumps(cpu_reference_c1()), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU "
              "decomposition is a design only).", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench needs a GPU"
    log(f"[bench] building {args.config} state")
    c, nl, Gr, RS, Th, fields = build_state(args.config)
    ncell = c["nx"] * c["ny"] * c["nz"]
    mcell = lambda ms: ncell / (ms * 1e-3) / 1e6  # noqa: E731

    if args.copy_calibration:
        from pycles_amd import _capi
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        _capi.stream_kernel(0, a.data_ptr(), b.data_ptr(), 1 << 30, 16384, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 1 GiB read + 1 GiB written (pycles_stream copy, 16384 blocks)")
        del a, b
    ceiling = hbm_ceiling()
    log(f"[bench] HBM ceiling of the streaming kernels: {ceiling['best_GBs']} GB/s")
    if args.newton_steps is not None:
        Th.newton_steps = args.newton_steps
    if args.load_guess:
        guess, ginfo = load_guess(Th, args.load_guess)
        log(f"[bench] loaded the saturation guess from {args.load_guess}")
    else:
        log("[bench] training the saturation guess")
        guess, ginfo = train_guess(Th, args.guess_pairs_log2, args.guess_epochs)
        log(f"[bench] trained in {ginfo['train_s']} s, held-out rms {ginfo['heldout_rms_K']:.4f} K")
    if args.save_guess:
        st = guess.state_dict()
        np.savez(args.save_guess, params=st["params"], norm=st["norm"])

    log("[bench] FP32 + learned form")
    simL, msL, phL, cntL = run_form(nl, Gr, RS, Th, fields, "f32", args.steps, args.warmup, learned=guess)
    log(f"[bench] FP32+learned {msL:.3f} ms/step {phL}")
    del simL
    out_plain = None
    if not args.no_plain:
        simP, msP, phP, cntP = run_form(nl, Gr, RS, Th, fields, "f32", args.steps, args.warmup)
        log(f"[bench] FP32 plain {msP:.3f} ms/step")
        del simP
        out_plain = {"ms_per_step": round(msP, 4), "value": round(mcell(msP), 2),
                     "phases_ms_per_step": {k: round(v / args.steps, 4) for k, v in phP.items()},
                     "sa_iterations_per_saturated_cell": round(cntP[1] / max(cntP[0], 1), 3)}
    out64 = None
    if not args.no_fp64:
        sim64, ms64, ph64, _ = run_form(nl, Gr, RS, Th, fields, "f64", args.steps, args.warmup)
        log(f"[bench] FP64 {ms64:.3f} ms/step")
        del sim64
        out64 = {"ms_per_step": round(ms64, 4), "value": round(mcell(ms64), 2),
                 "phases_ms_per_step": {k: round(v / args.steps, 4) for k, v in ph64.items()}}
    torch.cuda.empty_cache()

    kernels = per_kernel(phL, args.steps)
    traffic = measured_traffic() if args.config == "moist512" else {}
    for k, v in kernels.items():
        ach = v["bytes_per_cell"] * ncell / (v["ms_per_launch"] * 1e-3) / 1e9
        v["achieved_GBs"] = round(ach, 1)
        v["hbm_frac"] = round(ach / HBM_PEAK_GBS, 4)
        if k in KERNEL_FLOPS and tuple(c["order"]) == (5, 5):  # flop counts are WENO5's
            tf = KERNEL_FLOPS[k] * ncell / (v["ms_per_launch"] * 1e-3) / 1e12
            v["flop_per_cell"] = KERNEL_FLOPS[k]
            v["achieved_TFLOPs"] = round(tf, 2)
            v["valu_frac"] = round(tf / VALU_PEAK_TFLOPS, 4)
        if k in traffic and "bytes_per_launch" in traffic[k]:
            v["traffic_bytes_per_launch"] = traffic[k]["bytes_per_launch"]
        if k in traffic and "valu_wave_instr_per_launch" in traffic[k]:
            # one wavefront VALU instruction per cycle per CU (4 SIMDs, each
            # issuing a wave64 instruction every 4 cycles) at 2.4 GHz
            issue_ms = traffic[k]["valu_wave_instr_per_launch"] / (CUS * CLOCK_GHZ * 1e9) * 1e3
            v["valu_issue_ms"] = round(issue_ms, 4)
            v["valu_issue_frac"] = round(issue_ms / v["ms_per_launch"], 4)
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_launch"])
    kd = kernels[dom]
    tr = traffic.get(dom)
    roof = {"bound": "hbm", "kernel": kernel_names(*c["order"])[dom], "achieved": kd["achieved_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": kd["hbm_frac"],
            "frac_of_copy_ceiling": round(kd["achieved_GBs"] / ceiling["best_GBs"]["copy"], 4),
            "traffic": tr.get("bytes_per_launch") if tr else None,
            "valu_issue_frac": kd.get("valu_issue_frac"),
            "algorithmic_bytes_per_launch": kd["bytes_per_cell"] * ncell, "ms_per_launch": kd["ms_per_launch"]}
    # the whole step against its algorithmic bytes: 186 field passes per
    # moist SSP-RK3 step (62 per stage), FP32 storage
    step_bytes = STEP_PASSES * 4 * ncell
    step_roof = {"passes_per_step": STEP_PASSES, "bytes_per_step": step_bytes,
                 "achieved_GBs": round(step_bytes / (msL * 1e-3) / 1e9, 1),
                 "frac": round(step_bytes / (msL * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    ginfo["newton_steps"] = Th.newton_steps
    ginfo["saturated_fraction"] = round(cntL[0] / (ncell * 3 * args.steps), 4)
    ginfo["newton_iterations_per_saturated_cell"] = round(cntL[1] / max(cntL[0], 1), 3)
    ginfo["fallback_fraction"] = cntL[3] / max(cntL[0], 1)

    out = {"metric": METRIC, "value": round(mcell(msL), 2), "unit": "Mcell-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(msL, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded BOMEX-like moist state with a saturated layer)",
           "config": {"workload": f"{args.config}: {c['nx']}x{c['ny']}x{c['nz']} moist SA, WENO{c['order'][0]}/"
                                  f"{c['order'][1]}, FFT+tridiagonal pressure, dt={c['dt']} s, "
                                  "FP32 + learned saturation guess", "cells": ncell},
           "phases_ms_per_step": {k: round(v / args.steps, 4) for k, v in phL.items()},
           "roofline": roof, "step_roofline": step_roof, "hbm_ceiling": ceiling, "kernels": kernels,
           "learned_guess": ginfo}
    if out64:
        out["fp64"] = out64
        out["speedup_vs_fp64"] = round(out64["ms_per_step"] / msL, 3)
    if out_plain:
        out["fp32_plain"] = out_plain
    if not args.no_fp64 and args.error_steps > 0:
        log(f"[bench] error study: {args.error_steps} steps")
        es = error_study(nl, Gr, RS, Th, fields, guess, args.error_steps, floors=not args.no_floor)
        out["errors_vs_fp64"] = es["errors_vs_fp64"]
        out["error_gate"] = error_gate(es["errors_vs_fp64"],
                                       floor=es["noise_floor_fp64"].get("fp64_fp32_storage"),
                                       secant_floor=es["noise_floor_fp64"].get("fp64_secant_1e-7"))
        if es["noise_floor_fp64"]:
            out["noise_floor_fp64"] = es["noise_floor_fp64"]
        torch.cuda.empty_cache()
    if not args.no_cpu:
        log("[bench] CPU baseline")
        out["cpu_baseline"] = cpu_baseline(c, Gr, RS, Th, fields)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
