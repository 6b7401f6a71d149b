"""Seeded initial states for the benchmark configurations.

The reference initializes each case on the host (e.g. ``InitStableBubble``,
``Initialization.pyx:64-105``; ``InitBomex``, ``:260-368``) with numpy's
global RNG.  Here the noise comes from a counter-based generator,
SplitMix64(seed, field, cell index) -> U[0,1), with Box-Muller for normals, so
a state is reproducible independent of layout and thread count.  Fields are
built in the device layout ``[k][j][i]`` as FP64 numpy arrays, the z ghost
levels are filled with the boundary rule of each variable, and the arrays are
uploaded once.  White-noise velocities are strongly divergent, so the seeded
velocities are projected onto the divergence-free space once before the run
(``project_initial``); advecting entropy with the raw noise would change s by
O(100) J/kg/K in the first stage.

Configurations (sizes, spacings, profiles) follow the benchmark definitions:
  dry64     64^3, dx=dy=100 m, dz=40 m, dry, theta=300 K + U(+-0.1 K) below
            1 km, u=5+N(0,0.5), v,w=N(0,0.5)
  moist     BOMEX-like moist profiles (qt 17 -> 16.3 g/kg at 520 m -> 4.3 g/kg
            at 2 km, held above; theta_l 298.7 K to 520 m then +3.85 K/km;
            qt +-0.025 g/kg, theta_l +-0.1 K uniform noise below 1.6 km;
            u=-8.75+N(0,sigma), v,w=N(0,sigma)) on any grid.  These profiles
            peak at 94% relative humidity (510 m), so the configurations that
            need saturated cells (the 512^2 and 1024^2 ones: ~20% of cells)
            add a saturated layer, qt = max(qt, 1.02 qv*), for 500-1500 m.
Profiles are evaluated at cell centres z_half (the reference's InitBomex
evaluates them at ``Gr.zl_half``, which Grid.pyx:166-167 sets to full levels).
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft

from . import parameters as P
from . import thermo_host as H

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _splitmix64(x):
    x = x + _G
    z = (x ^ (x >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


_CHUNK = 1 << 21


def _workers():
    return max(1, min(16, os.cpu_count() or 1))


def _chunked(n, fn):
    """Fill an array of n values chunk by chunk on a thread pool (numpy's
    element-wise operations release the GIL); the values do not depend on
    the chunking."""
    out = np.empty(n)
    bounds = [(a, min(a + _CHUNK, n)) for a in range(0, n, _CHUNK)]

    def run(b):
        out[b[0]:b[1]] = fn(b[0], b[1])

    if len(bounds) <= 1:
        for b in bounds:
            run(b)
    else:
        with ThreadPoolExecutor(max_workers=_workers()) as ex:
            list(ex.map(run, bounds))
    return out


def _uniform_range(key, a, b):
    with np.errstate(over="ignore"):
        z = _splitmix64(np.arange(a, b, dtype=np.uint64) ^ key)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _key(seed, field):
    with np.errstate(over="ignore"):
        return _splitmix64(np.uint64(seed) * _G ^ (np.uint64(field) << np.uint64(56)))


def uniform(seed: int, field: int, shape):
    """U[0,1) per cell of an interior block of the given (nz, ny, nx) shape."""
    key = _key(seed, field)
    return _chunked(int(np.prod(shape)), lambda a, b: _uniform_range(key, a, b)).reshape(shape)


def normal(seed: int, field: int, shape):
    k1, k2 = _key(seed, 2 * field), _key(seed, 2 * field + 1)

    def box_muller(a, b):
        u1 = 1.0 - _uniform_range(k1, a, b)
        u2 = _uniform_range(k2, a, b)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)

    return _chunked(int(np.prod(shape)), box_muller).reshape(shape)


def fill_z_ghosts(f: np.ndarray, gw: int, bc: float):
    """set_bcs in z on a device-layout array (prognostic_variables.h:205-270)."""
    nzg = f.shape[0]
    if bc == 1.0:
        for m in range(gw):
            f[gw - 1 - m] = f[gw + m]
            f[nzg - gw + m] = f[nzg - gw - 1 - m]
    else:
        f[gw - 1] = 0.0
        f[nzg - gw] = 0.0
        for m in range(1, gw):
            f[gw - 1 - m] = bc * f[gw - 1 + m]
            f[nzg - gw + m] = bc * f[nzg - gw - m]
    return f


CONFIGS = {
    "dry64": dict(nx=64, ny=64, nz=64, gw=3, dx=100.0, dy=100.0, dz=40.0, moist=False,
                  order=(5, 5), dt=1.0, steps=1, seed=1, sigma=0.5),
    "moist256": dict(nx=256, ny=256, nz=128, gw=3, dx=50.0, dy=50.0, dz=20.0, moist=True,
                     order=(5, 5), dt=2.0, steps=10, seed=2, sigma=0.3),
    "moist512": dict(nx=512, ny=512, nz=256, gw=3, dx=50.0, dy=50.0, dz=20.0, moist=True,
                     order=(5, 5), dt=1.0, steps=100, seed=3, sigma=0.3, cloud=(500.0, 1500.0)),
    "moist1024": dict(nx=1024, ny=1024, nz=192, gw=5, dx=25.0, dy=25.0, dz=10.0, moist=True,
                      order=(9, 5), dt=0.5, steps=20, seed=5, sigma=1.0, cloud=(500.0, 880.0)),
}


def namelist(nx, ny, nz, gw, dx, dy, dz, moist, order=(5, 5), dt=1.0, t_max=1e9, precision="f64",
             cfl_limit=1e6, **extra):
    nl = {
        "grid": {"dims": 3, "nx": nx, "ny": ny, "nz": nz, "gw": gw, "dx": dx, "dy": dy, "dz": dz},
        "scalar_transport": {"order": order[0]},
        "momentum_transport": {"order": order[1]},
        "time_stepping": {"ts_type": 3, "dt_initial": dt, "dt_max": dt, "t_max": t_max,
                          "cfl_limit": cfl_limit},
        "thermodynamics": {"do_qt_clipping": True},
        "microphysics": {"scheme": "None_SA" if moist else "None_Dry"},
        "mi355x": {"precision": precision},
        "initialization": {"seed": extra.get("seed", 1), "sigma": extra.get("sigma", 0.5),
                           "case": "moist" if moist else "dry", "cloud": extra.get("cloud")},
    }
    return nl


def config_namelist(name, precision="f64", **over):
    c = dict(CONFIGS[name])
    c.update(over)
    return namelist(precision=precision, **c)


def _velocities(Gr, seed, sigma, u_mean):
    d = Gr.dims
    gw = d.gw
    shape = (d.n[2], d.n[1], d.n[0])
    out = {}
    for f, (name, mean) in enumerate((("u", u_mean), ("v", 0.0), ("w", 0.0))):
        a = np.zeros((d.nzg, d.n[1], d.n[0]))
        a[gw:gw + d.n[2]] = mean + sigma * normal(seed, 10 + f, shape)
        if name == "w":
            a[gw + d.n[2] - 1] = 0.0  # top interior face w
        out[name] = fill_z_ghosts(a, gw, -1.0 if name == "w" else 1.0)
    return out


def project_initial(Gr, RS, u, v, w):
    """Make seeded velocity noise divergence-free before the run starts.

    Uses the discrete operators of the pressure projection (second-order
    divergence of rho0*u, PressureSolver.pyx:138-175; modified wavenumbers and
    tridiagonal rows, PressureFFTParallel.pyx:64-148; (0,0) mode zero) with
    numpy FFTs.  Initialization only: the per-stage projection runs on the GPU.
    """
    d = Gr.dims
    gw, nz = d.gw, d.n[2]
    dx, dy, dz = d.dx
    sl = slice(gw, gw + nz)
    ah = RS.alpha0_half[sl][:, None, None]
    a0 = RS.alpha0[sl][:, None, None]
    a0m = RS.alpha0[gw - 1:gw - 1 + nz][:, None, None]
    w = w - w[:, :, :].mean(axis=(1, 2), keepdims=True)
    ui, vi = u[sl], v[sl]
    div = (ui / ah - np.roll(ui, 1, axis=2) / ah) / dx
    div = div + (vi / ah - np.roll(vi, 1, axis=1) / ah) / dy
    div = div + (w[sl] / a0 - w[gw - 1:gw - 1 + nz] / a0m) / dz
    spec = scipy.fft.fft2(div, axes=(1, 2), workers=_workers())
    kx = np.arange(d.n[0]); kx = np.where(kx <= d.n[0] // 2, kx, kx - d.n[0])
    ky = np.arange(d.n[1]); ky = np.where(ky <= d.n[1] // 2, ky, ky - d.n[1])
    kx2 = (2.0 * np.cos(2.0 * np.pi / d.n[0] * kx) - 2.0) / dx / dx
    ky2 = (2.0 * np.cos(2.0 * np.pi / d.n[1] * ky) - 2.0) / dy / dy
    kk = ky2[:, None] + kx2[None, :]
    rho0, rho0h = RS.rho0, RS.rho0_half
    dzi2 = 1.0 / dz / dz
    a = np.array([0.0] + [dzi2 * rho0[k + gw - 1] for k in range(1, nz)])
    c = np.array([dzi2 * rho0[k + gw] for k in range(nz - 1)] + [0.0])
    dd = np.array([(rho0[k + gw] + (rho0[k + gw - 1] if k > 0 else 0.0)) if k < nz - 1
                   else rho0[k + gw - 1] for k in range(nz)]) * dzi2
    dd[0] = rho0[gw] * dzi2
    kk0 = kk.copy()
    kk0[0, 0] = 1.0  # placeholder; the (0,0) mode is zeroed below
    x = np.empty_like(spec)

    def tdma_rows(j0, j1):
        # the tridiagonal solve of the columns of wavenumber rows [j0, j1)
        kr = kk0[j0:j1]
        scr = np.empty((nz,) + kr.shape)
        b = rho0h[gw] * kr - dd[0]
        scr[0] = c[0] / b
        x[0, j0:j1] = spec[0, j0:j1] / b
        for k in range(1, nz):
            b = rho0h[k + gw] * kr - dd[k]
            m = 1.0 / (b - a[k] * scr[k - 1])
            scr[k] = c[k] * m
            x[k, j0:j1] = (spec[k, j0:j1] - a[k] * x[k - 1, j0:j1]) * m
        for k in range(nz - 2, -1, -1):
            x[k, j0:j1] = x[k, j0:j1] - scr[k] * x[k + 1, j0:j1]

    step = max(1, d.n[1] // _workers())
    with ThreadPoolExecutor(max_workers=_workers()) as ex:
        list(ex.map(lambda j: tdma_rows(j, min(j + step, d.n[1])), range(0, d.n[1], step)))
    x[:, 0, 0] = 0.0
    p = np.zeros_like(u)
    p[sl] = np.real(scipy.fft.ifft2(x, axes=(1, 2), workers=_workers()))
    fill_z_ghosts(p, gw, 1.0)
    u = u - (np.roll(p, -1, axis=2) - p) / dx
    v = v - (np.roll(p, -1, axis=1) - p) / dy
    w = w.copy()
    w[:-1] = w[:-1] - (p[1:] - p[:-1]) / dz
    return fill_z_ghosts(u, gw, 1.0), fill_z_ghosts(v, gw, 1.0), fill_z_ghosts(w, gw, -1.0)


def init_dry(Gr, RS, Th, seed=1, sigma=0.5):
    """Dry state of the 64^3 configuration (isentropic reference, theta_s=300 K, p_s=1000 hPa)."""
    RS.Pg = 1.0e5
    RS.Tg = 300.0
    RS.qtg = 0.0
    RS.u0 = 0.0
    RS.v0 = 0.0
    RS.initialize(Gr, Th)
    d = Gr.dims
    gw = d.gw
    shape = (d.n[2], d.n[1], d.n[0])
    zh = Gr.z_half[gw:gw + d.n[2]][:, None, None]
    p0h = RS.p0_half[gw:gw + d.n[2]][:, None, None]
    theta = 300.0 + np.where(zh < 1000.0, -0.1 + 0.2 * uniform(seed, 20, shape), 0.0)
    T = theta * H.exner(p0h)
    s = np.zeros((d.nzg, d.n[1], d.n[0]))
    s[gw:gw + d.n[2]] = H.entropy_dry(p0h, T)
    fields = _velocities(Gr, seed, sigma, 5.0)
    fields["u"], fields["v"], fields["w"] = project_initial(Gr, RS, fields["u"], fields["v"], fields["w"])
    fields["s"] = fill_z_ghosts(s, gw, 1.0)
    return fields


def moist_profiles(z):
    thetal = np.where(z <= 520.0, 298.7, 298.7 + (z - 520.0) * 3.85e-3)
    qt = np.where(z <= 520.0, 17.0 + z * (16.3 - 17.0) / 520.0,
                  np.where(z <= 2000.0, 16.3 + (z - 520.0) * (4.3 - 16.3) / (2000.0 - 520.0), 4.3))
    return thetal, qt / 1000.0


def init_moist(Gr, RS, Th, seed=2, sigma=0.3, cloud=None):
    RS.Pg = 1.015e5
    RS.Tg = 300.4
    RS.qtg = 0.02245
    RS.u0 = 0.0
    RS.v0 = 0.0
    RS.initialize(Gr, Th)
    d = Gr.dims
    gw = d.gw
    shape = (d.n[2], d.n[1], d.n[0])
    zh = Gr.z_half[gw:gw + d.n[2]][:, None, None]
    p0h = RS.p0_half[gw:gw + d.n[2]][:, None, None]
    thetal, qt = moist_profiles(zh)
    if cloud is not None:
        # saturated layer: qt raised to 1.02 qv*(T, p0) between cloud[0] and cloud[1]
        T0 = thetal * H.exner(p0h)
        qvs = H.qv_star_c(p0h, qt, Th.CC.table(T0))
        lay = (zh >= cloud[0]) & (zh <= cloud[1])
        qt = np.where(lay, np.maximum(qt, 1.02 * qvs), qt)
    pert = zh <= 1600.0
    thetal = thetal + np.where(pert, -0.1 + 0.2 * uniform(seed, 30, shape), 0.0)
    qt = qt + np.where(pert, (-0.025 + 0.05 * uniform(seed, 31, shape)) / 1000.0, 0.0)
    T = thetal * H.exner(p0h)
    s = np.zeros((d.nzg, d.n[1], d.n[0]))
    q = np.zeros_like(s)
    s[gw:gw + d.n[2]] = H.entropy_sa(p0h, T, qt, 0.0, 0.0)
    q[gw:gw + d.n[2]] = qt
    fields = _velocities(Gr, seed, sigma, -8.75)
    fields["u"], fields["v"], fields["w"] = project_initial(Gr, RS, fields["u"], fields["v"], fields["w"])
    fields["s"] = fill_z_ghosts(s, gw, 1.0)
    fields["qt"] = fill_z_ghosts(q, gw, 1.0)
    return fields


def bomex_profiles(z):
    """InitBomex's liquid-water potential temperature, total water (kg/kg)
    and u profiles (Initialization.pyx:296-329)."""
    z = np.asarray(z, dtype=np.float64)
    thetal = np.empty_like(z)
    qt = np.empty_like(z)
    u = np.empty_like(z)
    for k, zk in enumerate(z):
        if zk <= 520.0:
            thetal[k] = 298.7
        if 520.0 < zk <= 1480.0:
            thetal[k] = 298.7 + (zk - 520) * (302.4 - 298.7) / (1480.0 - 520.0)
        if 1480.0 < zk <= 2000:
            thetal[k] = 302.4 + (zk - 1480.0) * (308.2 - 302.4) / (2000.0 - 1480.0)
        if zk > 2000.0:
            thetal[k] = 308.2 + (zk - 2000.0) * (311.85 - 308.2) / (3000.0 - 2000.0)
        if zk <= 520:
            qt[k] = 17.0 + zk * (16.3 - 17.0) / 520.0
        if 520.0 < zk <= 1480.0:
            qt[k] = 16.3 + (zk - 520.0) * (10.7 - 16.3) / (1480.0 - 520.0)
        if 1480.0 < zk <= 2000.0:
            qt[k] = 10.7 + (zk - 1480.0) * (4.2 - 10.7) / (2000.0 - 1480.0)
        if zk > 2000.0:
            qt[k] = 4.2 + (zk - 2000.0) * (3.0 - 4.2) / (3000.0 - 2000.0)
        qt[k] /= 1000.0
        if zk <= 700.0:
            u[k] = -8.75
        if zk > 700.0:
            u[k] = -8.75 + (zk - 700.0) * (-4.61 - -8.75) / (3000.0 - 700.0)
    return thetal, qt, u


def init_bomex(Gr, RS, Th, seed=6):
    """InitBomex (Initialization.pyx:260-368): reference state Pg = 1.015e5,
    Tg = 300.4, qtg = 0.02245; the Galilean translation u0 = (max u + min u)/2;
    u - u0, v = w = 0; theta_l + U(-0.05, 0.05) K and qt + U(-0.0125, 0.0125)
    g/kg up to 1600 m; s = entropy(p0_half, theta_l exner, qt).  The noise
    comes from the counter-based generator instead of numpy's global one."""
    RS.Pg = 1.015e5
    RS.Tg = 300.4
    RS.qtg = 0.02245
    RS.initialize(Gr, Th)
    d = Gr.dims
    shape = (d.nzg, d.n[1], d.n[0])
    zh = Gr.zl_half[:, None, None]
    p0h = RS.p0_half[:, None, None]
    thetal, qt, u = bomex_profiles(Gr.zl_half)
    RS.v0 = 0.0
    RS.u0 = 0.5 * (np.amax(u) + np.amin(u))
    pert = zh <= 1600.0
    theta_pert = (uniform(seed, 30, shape) - 0.5) * 0.1
    qt_pert = (uniform(seed, 31, shape) - 0.5) * 0.025 / 1000.0
    temp = np.where(pert, (thetal[:, None, None] + theta_pert), thetal[:, None, None]) * H.exner(p0h)
    q = np.where(pert, qt[:, None, None] + qt_pert, qt[:, None, None] + 0.0 * qt_pert)
    fields = {"u": np.broadcast_to((u - RS.u0)[:, None, None], shape).copy(),
              "v": np.full(shape, 0.0 - RS.v0), "w": np.zeros(shape),
              "s": H.entropy_sa(p0h, temp, q, 0.0, 0.0), "qt": q}
    for n, bc in (("u", 1.0), ("v", 1.0), ("w", -1.0), ("s", 1.0), ("qt", 1.0)):
        fill_z_ghosts(fields[n], d.gw, bc)
    return fields


def bomex_namelist(nx=64, ny=64, nz=75, gw=3, dx=100.0, dz=40.0, dt=10.0, precision="f64", seed=6,
                   cfl_limit=0.7, t_max=21600.0):
    """The reference's Bomex namelist (generate_namelist.py:311-393) on the
    device path: None_SA thermodynamics, Smagorinsky SGS, Rayleigh damping
    (gamma_r = 0.2 /s over the top 600 m), WENO5 transport, CFL-limited
    dt up to dt_max."""
    nl = namelist(nx, ny, nz, gw, dx, dx, dz, True, order=(5, 5), dt=dt, t_max=t_max, precision=precision,
                  cfl_limit=cfl_limit, seed=seed)
    nl["time_stepping"]["dt_max"] = dt
    nl["meta"] = {"casename": "Bomex", "simname": "Bomex", "uuid": "bomex-0000-00000"}
    nl["sgs"] = {"scheme": "Smagorinsky"}
    nl["diffusion"] = {"qt_entropy_source": False}
    nl["damping"] = {"scheme": "Rayleigh", "Rayleigh": {"gamma_r": 0.2, "z_d": 600.0}}
    nl["initialization"]["case"] = "bomex"
    return nl


def initialize(namelist, Gr, PV, RS, Th):
    """Build the initial state and upload it (the InitializationFactory step)."""
    ini = namelist.get("initialization", {})
    moist = namelist["microphysics"]["scheme"] == "None_SA"
    seed = int(ini.get("seed", 1))
    sigma = float(ini.get("sigma", 0.5))
    if ini.get("case") == "bomex":
        fields = init_bomex(Gr, RS, Th, seed=seed)
    elif moist:
        fields = init_moist(Gr, RS, Th, seed=seed, sigma=sigma, cloud=ini.get("cloud"))
    else:
        fields = init_dry(Gr, RS, Th, seed=seed, sigma=sigma)
    for name, arr in fields.items():
        if name in ("s", "qt"):
            PV.set_base_from_device_layout(name, arr)
        PV.set_device_layout(name, arr)
    return fields
