"""CPU oracle for the PyCLES dynamical core (ctypes face of liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module.  Product code under ``pycles_amd/`` never does.

Arrays are numpy float64 in the reference layout: one variable is
``[nlg0, nlg1, nlg2]`` C-order, z fastest (reference
``PrognosticVariables.pxd:45-46``).  ``OracleSimulation.stage`` follows the
per-stage call order of ``Simulation3d.run`` (``Simulation3d.pyx:158-180``)
restricted to the hot path: Th.update, SA.update, MA.update, TS.update,
PV.Update_all_bcs, Pr.update, TS.adjust_timestep.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


class OrcDims(C.Structure):
    _fields_ = [("n", C.c_long * 3), ("nlg", C.c_long * 3), ("gw", C.c_long),
                ("npg", C.c_long), ("dx", C.c_double * 3), ("dxi", C.c_double * 3)]


class OrcLookup(C.Structure):
    _fields_ = [("n", C.c_long), ("x_min", C.c_double), ("dx", C.c_double),
                ("dxi", C.c_double), ("y", _dp)]


class OrcRef(C.Structure):
    _fields_ = [(name, _dp) for name in ("p0", "p0_half", "alpha0", "alpha0_half", "rho0", "rho0_half")]


class OrcSAStats(C.Structure):
    _fields_ = [("saturated", C.c_long), ("iterations", C.c_long),
                ("max_iterations", C.c_long), ("capped", C.c_long)]


def build():
    """Compile liboracle.so from oracle.c with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        d = C.c_double
        L.orc_interp_weno5.restype = d
        L.orc_interp_weno7.restype = d
        L.orc_interp_weno9.restype = d
        L.orc_interp_4.restype = d
        L.orc_interp_6.restype = d
        L.orc_interp_8.restype = d
        for f in ("orc_interp_weno5", "orc_interp_weno7", "orc_interp_weno9",
                  "orc_interp_4", "orc_interp_6", "orc_interp_8"):
            getattr(L, f).argtypes = [_dp]
        L.orc_latent_heat_variable.restype = d
        L.orc_latent_heat_variable.argtypes = [d]
        L.orc_lookup_eval.restype = d
        L.orc_lookup_eval.argtypes = [C.POINTER(OrcLookup), d]
        L.orc_eos_sa.restype = C.c_int
        L.orc_eos_sa.argtypes = [C.POINTER(OrcLookup), d, d, d, _dp, _dp, _dp, _dp, C.c_int]
        L.orc_entropy_sa.restype = d
        L.orc_entropy_sa.argtypes = [d, d, d, d, d]
        L.orc_alpha.restype = d
        L.orc_alpha.argtypes = [d, d, d, d]
        L.orc_exner.restype = d
        L.orc_exner.argtypes = [d]
        L.orc_eos_dry.restype = d
        L.orc_eos_dry.argtypes = [d, d]
        L.orc_entropy_dry.restype = d
        L.orc_entropy_dry.argtypes = [d, d]
        L.orc_cfl_max.restype = d
        L.orc_rk3_update.argtypes = [C.c_long, C.c_int, d, _dp, _dp, _dp]
        L.orc_tdma_solve.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp]
        L.orc_fft.argtypes = [C.c_long, _dp, _dp, C.c_int]
        L.orc_update_bcs.argtypes = [C.POINTER(OrcDims), _dp, d]
        L.orc_surface_bomex.argtypes = [C.POINTER(OrcDims), C.POINTER(OrcRef)] + [_dp] * 10 + [d, d, d, C.c_int, _dp]
        L.orc_forcing_bomex_profiles.argtypes = [C.POINTER(OrcDims), _dp, _dp, _dp]
        L.orc_forcing_bomex.argtypes = [C.POINTER(OrcDims), C.POINTER(OrcRef)] + [_dp] * 11 + [d, d, d]
        L.orc_rayleigh_profiles.argtypes = [C.POINTER(OrcDims), _dp, _dp, d, d, _dp, _dp]
        L.orc_rayleigh_update.argtypes = [C.POINTER(OrcDims), _dp, _dp, _dp]
        L.orc_entropyflux_from_thetaflux_qtflux.restype = d
        L.orc_entropyflux_from_thetaflux_qtflux.argtypes = [d] * 6
        L.orc_kinematics_update.argtypes = [C.POINTER(OrcDims), C.POINTER(_dp), _dp, _dp, _dp]
        L.orc_smagorinsky_update.argtypes = [C.POINTER(OrcDims), _dp, _dp, _dp, _dp, d, d]
        L.orc_scalar_diffusion_update.argtypes = [C.POINTER(OrcDims), C.POINTER(OrcRef), _dp, C.c_int,
                                                  C.POINTER(_dp), C.POINTER(_dp), _dp]
        L.orc_momentum_diffusion_update.argtypes = [C.POINTER(OrcDims), C.POINTER(OrcRef), _dp, _dp,
                                                    C.POINTER(_dp), _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def make_dims(nx, ny, nz, gw, dx, dy, dz) -> OrcDims:
    d = OrcDims()
    lib().orc_dims_init(C.byref(d), C.c_long(nx), C.c_long(ny), C.c_long(nz), C.c_long(gw),
                        C.c_double(dx), C.c_double(dy), C.c_double(dz))
    return d


class Lookup:
    """Holds an ``OrcLookup`` over a pv*(T) table (reference lookup.h:16-41)."""

    def __init__(self, T: np.ndarray, pv: np.ndarray):
        self.y = np.ascontiguousarray(pv, dtype=np.float64)
        self.c = OrcLookup()
        self.c.n = len(self.y)
        self.c.x_min = float(np.min(T))
        self.c.dx = float(T[1] - T[0])
        self.c.dxi = 1.0 / self.c.dx
        self.c.y = ptr(self.y)

    def __call__(self, x: float) -> float:
        return lib().orc_lookup_eval(C.byref(self.c), C.c_double(x))


class Ref:
    """Reference-state profiles (length nlg[2]) bundled as ``OrcRef``."""

    def __init__(self, p0, p0_half, alpha0, alpha0_half, rho0, rho0_half):
        self.arrays = [np.ascontiguousarray(a, dtype=np.float64)
                       for a in (p0, p0_half, alpha0, alpha0_half, rho0, rho0_half)]
        self.c = OrcRef(*[ptr(a) for a in self.arrays])


def weno(order: int, q) -> float:
    q = np.ascontiguousarray(q, dtype=np.float64)
    f = {5: lib().orc_interp_weno5, 7: lib().orc_interp_weno7, 9: lib().orc_interp_weno9}[order]
    return f(ptr(q))


def centred(order: int, q) -> float:
    q = np.ascontiguousarray(q, dtype=np.float64)
    f = {5: lib().orc_interp_4, 7: lib().orc_interp_6, 9: lib().orc_interp_8}[order]
    return f(ptr(q))


def eos_sa(lt: Lookup, p0, s, qt, max_iter=0):
    T, qv, ql, qi = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    it = lib().orc_eos_sa(C.byref(lt.c), p0, s, qt, C.byref(T), C.byref(qv), C.byref(ql),
                          C.byref(qi), max_iter)
    return T.value, qv.value, ql.value, qi.value, it


def tdma(x, a, b, c):
    x = np.array(x, dtype=np.float64)
    scr = np.zeros_like(x)
    lib().orc_tdma_solve(len(x), ptr(x), ptr(np.asarray(a, np.float64)), ptr(np.asarray(b, np.float64)),
                         ptr(np.asarray(c, np.float64)), ptr(scr))
    return x


def fft(x, inverse=False):
    x = np.asarray(x, dtype=np.complex128)
    re = np.ascontiguousarray(x.real.copy())
    im = np.ascontiguousarray(x.imag.copy())
    lib().orc_fft(len(re), ptr(re), ptr(im), 1 if inverse else 0)
    return re + 1j * im


MOIST_PV = ("u", "v", "w", "s", "qt")
DRY_PV = ("u", "v", "w", "s")
MOIST_DV = ("temperature", "qv", "ql", "qi", "alpha", "buoyancy", "theta_rho",
            "buoyancy_frequency", "thetali", "dynamic_pressure", "divergence")
DRY_DV = ("temperature", "alpha", "buoyancy", "theta", "buoyancy_frequency",
          "dynamic_pressure", "divergence")
BC = {"u": 1.0, "v": 1.0, "w": -1.0, "s": 1.0, "qt": 1.0}


class OracleSimulation:
    """Reference-layout FP64 state plus the hot-path stage sequence."""

    def __init__(self, dims: OrcDims, ref: Ref, moist: bool, lookup: Lookup | None,
                 scalar_order=5, momentum_order=5, dt=1.0, dt_max=None, cfl_limit=1e6,
                 t_max=1e30, max_iter=0, sgs=None, bomex=None):
        self.d = dims
        self.ref = ref
        self.moist = moist
        self.lt = lookup
        self.so = scalar_order
        self.mo = momentum_order
        self.shape = tuple(dims.nlg)
        self.npg = dims.npg
        self.pv_names = MOIST_PV if moist else DRY_PV
        self.dv_names = MOIST_DV if moist else DRY_DV
        nv = len(self.pv_names)
        self.values = np.zeros((nv,) + self.shape)
        self.tend = np.zeros_like(self.values)
        self.copy = np.zeros_like(self.values)
        self.dv = {n: np.zeros(self.shape) for n in self.dv_names}
        self.flux = np.zeros((3,) + self.shape)
        self.dt = dt
        self.dt_max = dt if dt_max is None else dt_max
        self.cfl_limit = cfl_limit
        self.t = 0.0
        self.t_max = t_max
        self.max_iter = max_iter
        self.cfl_max = 0.0
        self.sa_stats = OrcSAStats()
        # BOMEX surface, forcing and Rayleigh damping: {"zl", "zl_half", "u0",
        # "v0", "gamma_r", "z_d"} (the grid's full and half level heights)
        self.bomex = bomex
        if bomex is not None:
            n2 = dims.nlg[2]
            self.forcing_prof = np.zeros(5 * n2)
            lib().orc_forcing_bomex_profiles(C.byref(dims), ptr(np.ascontiguousarray(bomex["zl_half"])),
                                             ptr(ref.arrays[1]), ptr(self.forcing_prof))
            self.gamma_z = np.zeros(n2)
            self.gamma_zhalf = np.zeros(n2)
            lib().orc_rayleigh_profiles(C.byref(dims), ptr(np.ascontiguousarray(bomex["zl"])),
                                        ptr(np.ascontiguousarray(bomex["zl_half"])), bomex["gamma_r"],
                                        bomex["z_d"], ptr(self.gamma_z), ptr(self.gamma_zhalf))
            self.surface_flux2d = np.zeros(8 * dims.nlg[0] * dims.nlg[1])
        # Smagorinsky SGS with scalar and momentum diffusion: {"cs", "prt"}
        self.sgs = sgs
        if sgs is not None:
            self.dv["viscosity"] = np.zeros(self.shape)
            self.dv["diffusivity"] = np.zeros(self.shape)
            self.vgrad = np.zeros((9,) + self.shape)
            self.strain = np.zeros((9,) + self.shape)
            self.smag = np.zeros(self.shape)

    # -- accessors ---------------------------------------------------------
    def pv(self, name):
        return self.values[self.pv_names.index(name)]

    def pt(self, name):
        return self.tend[self.pv_names.index(name)]

    # -- operators (each mirrors one reference update) ---------------------
    def update_all_bcs(self):
        for n in self.pv_names:
            lib().orc_update_bcs(C.byref(self.d), ptr(self.pv(n)), BC[n])

    def thermo_update(self):
        L = lib()
        r = C.byref(self.ref.c)
        if self.moist:
            dv = self.dv
            L.orc_thermo_sa_update(C.byref(self.d), r, C.byref(self.lt.c), ptr(self.pv("s")),
                                   ptr(self.pv("qt")), ptr(self.pt("w")), ptr(dv["temperature"]),
                                   ptr(dv["qv"]), ptr(dv["ql"]), ptr(dv["qi"]), ptr(dv["alpha"]),
                                   ptr(dv["buoyancy"]), ptr(dv["theta_rho"]),
                                   ptr(dv["buoyancy_frequency"]), ptr(dv["thetali"]), 1,
                                   self.max_iter, C.byref(self.sa_stats))
        else:
            dv = self.dv
            L.orc_thermo_dry_update(C.byref(self.d), r, ptr(self.pv("s")), ptr(self.pt("w")),
                                    ptr(dv["temperature"]), ptr(dv["alpha"]), ptr(dv["buoyancy"]),
                                    ptr(dv["theta"]), ptr(dv["buoyancy_frequency"]))

    def scalar_advection_update(self):
        names = [n for n in self.pv_names if n in ("s", "qt")]
        vel = (_dp * 3)(*[ptr(self.pv(n)) for n in ("u", "v", "w")])
        phi = (_dp * len(names))(*[ptr(self.pv(n)) for n in names])
        tend = (_dp * len(names))(*[ptr(self.pt(n)) for n in names])
        lib().orc_scalar_advection_update(C.byref(self.d), C.byref(self.ref.c), vel, len(names), phi,
                                          tend, ptr(self.flux), self.so)

    def momentum_advection_update(self):
        vel = (_dp * 3)(*[ptr(self.pv(n)) for n in ("u", "v", "w")])
        tend = (_dp * 3)(*[ptr(self.pt(n)) for n in ("u", "v", "w")])
        lib().orc_momentum_advection_update(C.byref(self.d), C.byref(self.ref.c), vel, tend, self.mo,
                                            ptr(self.flux))

    def kinematics_update(self):
        """Kinematics.update (Kinematics.pyx:41-58)."""
        vel = (_dp * 3)(*[ptr(self.pv(n)) for n in ("u", "v", "w")])
        lib().orc_kinematics_update(C.byref(self.d), vel, ptr(self.vgrad), ptr(self.strain), ptr(self.smag))

    def sgs_update(self):
        """Smagorinsky.update (SGS.pyx:146-152), no wall or iles adjustment."""
        lib().orc_smagorinsky_update(C.byref(self.d), ptr(self.dv["viscosity"]), ptr(self.dv["diffusivity"]),
                                     ptr(self.dv["buoyancy_frequency"]), ptr(self.smag), self.sgs["cs"],
                                     self.sgs["prt"])

    def scalar_diffusion_update(self):
        names = [n for n in self.pv_names if n in ("s", "qt")]
        phi = (_dp * len(names))(*[ptr(self.pv(n)) for n in names])
        tend = (_dp * len(names))(*[ptr(self.pt(n)) for n in names])
        lib().orc_scalar_diffusion_update(C.byref(self.d), C.byref(self.ref.c), ptr(self.dv["diffusivity"]),
                                          len(names), phi, tend, ptr(self.flux))

    def momentum_diffusion_update(self):
        tend = (_dp * 3)(*[ptr(self.pt(n)) for n in ("u", "v", "w")])
        lib().orc_momentum_diffusion_update(C.byref(self.d), C.byref(self.ref.c), ptr(self.strain),
                                            ptr(self.dv["viscosity"]), tend, ptr(self.flux), ptr(self.smag),
                                            ptr(self.dv["temperature"]), ptr(self.pt("s")))

    def surface_update(self):
        """SurfaceBomex.update (Surface.pyx:326-373)."""
        tmean = np.zeros(self.shape[2])
        lib().orc_horizontal_mean(C.byref(self.d), ptr(self.dv["temperature"]), ptr(tmean))
        dv = self.dv
        lib().orc_surface_bomex(C.byref(self.d), C.byref(self.ref.c), ptr(self.pv("u")), ptr(self.pv("v")),
                                ptr(self.pv("qt")), ptr(dv["temperature"]), ptr(dv["qv"]), ptr(dv["ql"]),
                                ptr(self.pt("u")), ptr(self.pt("v")), ptr(self.pt("s")), ptr(self.pt("qt")),
                                self.bomex["u0"], self.bomex["v0"], tmean[self.d.gw], 0, ptr(self.surface_flux2d))

    def damping_update(self):
        """Rayleigh.update (Damping.pyx:95-133): w with the half-level profile."""
        for n in self.pv_names:
            g = self.gamma_zhalf if n == "w" else self.gamma_z
            lib().orc_rayleigh_update(C.byref(self.d), ptr(g), ptr(self.pv(n)), ptr(self.pt(n)))

    def forcing_update(self):
        """ForcingBomex.update (Forcing.pyx:147-203)."""
        dv = self.dv
        lib().orc_forcing_bomex(C.byref(self.d), C.byref(self.ref.c), ptr(self.forcing_prof), ptr(self.pv("u")),
                                ptr(self.pv("v")), ptr(self.pv("s")), ptr(self.pv("qt")), ptr(dv["temperature"]),
                                ptr(dv["ql"]), ptr(self.pt("u")), ptr(self.pt("v")), ptr(self.pt("s")),
                                ptr(self.pt("qt")), 0.376e-4, self.bomex["u0"], self.bomex["v0"])

    def rk_update(self, rk_step):
        lib().orc_rk3_update(self.values.size, rk_step, self.dt, ptr(self.values), ptr(self.tend),
                             ptr(self.copy))
        if rk_step == 2:
            self.t += self.dt

    def pressure_update(self):
        lib().orc_pressure_update(C.byref(self.d), C.byref(self.ref.c), ptr(self.pv("u")),
                                  ptr(self.pv("v")), ptr(self.pv("w")),
                                  ptr(self.dv["dynamic_pressure"]), ptr(self.dv["divergence"]))

    def adjust_timestep(self):
        """TimeStepping.adjust_timestep on the last stage (TimeStepping.pyx:94-108)."""
        L = lib()
        L.orc_cfl_max.argtypes = [C.POINTER(OrcDims), _dp, _dp, _dp, C.c_double]
        self.cfl_max = L.orc_cfl_max(C.byref(self.d), ptr(self.pv("u")), ptr(self.pv("v")),
                                     ptr(self.pv("w")), self.dt)
        self.dt = min(self.dt_max, self.cfl_limit / (self.cfl_max / self.dt))
        if self.t + self.dt > self.t_max:
            self.dt = self.t_max - self.t

    def stage(self, rk_step):
        """Simulation3d.run's stage (Simulation3d.pyx:158-177) restricted to
        the operators restated here."""
        if self.sgs is not None:
            self.kinematics_update()
        self.thermo_update()
        self.scalar_advection_update()
        self.momentum_advection_update()
        if self.bomex is not None:
            self.surface_update()
        if self.sgs is not None:
            self.sgs_update()
        if self.bomex is not None:
            self.damping_update()
        if self.sgs is not None:
            self.scalar_diffusion_update()
            self.momentum_diffusion_update()
        if self.bomex is not None:
            self.forcing_update()
        self.rk_update(rk_step)
        self.update_all_bcs()
        self.pressure_update()
        if rk_step == 2:
            self.adjust_timestep()

    def begin_run(self):
        """The two calls ahead of the loop in Simulation3d.run (:141-148): the
        boundary fill and a thermodynamics update.  The latter adds buoyancy
        to the (not yet zeroed) w tendency, which the first stage then keeps."""
        self.update_all_bcs()
        self.thermo_update()

    def step(self):
        for rk in range(3):
            self.stage(rk)

    def divergence(self):
        div = np.zeros(self.shape)
        lib().orc_divergence(C.byref(self.d), C.byref(self.ref.c), ptr(self.pv("u")), ptr(self.pv("v")),
                             ptr(self.pv("w")), ptr(div))
        return div


# --------------------------------------------------------------------------
# Statistics (numpy restatement; reference layout [i][j][k], ghosts on every
# side, reductions over the interior columns and every level)

def horizontal_stats(d: OrcDims, a, b=None, mask=None):
    """ParallelMPI.HorizontalMean / MeanofSquares / MeanofCubes /
    Maximum / Minimum / MeanConditional (ParallelMPI.pyx:196-540) of v = a
    (times b): rows mean, mean of squares, mean of cubes, max, min, mask sum.
    Sums in the reference's serial order (i, then j) per level."""
    gw = d.gw
    sl = (slice(gw, d.nlg[0] - gw), slice(gw, d.nlg[1] - gw))
    v = np.asarray(a, dtype=np.float64)[sl]
    if b is not None:
        v = v * np.asarray(b, dtype=np.float64)[sl]
    m = np.ones_like(v) if mask is None else np.asarray(mask, dtype=np.float64)[sl]
    n2 = d.nlg[2]
    out = np.zeros((6, n2))
    flat_v = v.reshape(-1, n2)
    flat_m = m.reshape(-1, n2)
    for k in range(n2):
        s1 = s2 = s3 = sm = 0.0
        for x, w in zip(flat_v[:, k], flat_m[:, k]):
            s1 += x * w
            s2 += x * x * w
            s3 += x * x * x * w
            sm += w
        out[0, k], out[1, k], out[2, k], out[5, k] = s1, s2, s3, sm
        sel = flat_v[:, k][flat_m[:, k] != 0.0]
        out[3, k] = max(-9e12, sel.max()) if sel.size else -9e12
        out[4, k] = min(9e12, sel.min()) if sel.size else 9e12
    if mask is None:
        out[:3] *= 1.0 / float(d.n[0] * d.n[1])
    else:
        out[:3] /= np.maximum(out[5], 1.0)
    return out


def sa_stat_fields(p0_half, lt: Lookup, s, qt, T, qv):
    """The fields of ThermodynamicsSA.stats_io (ThermodynamicsSA.pyx:256-389):
    thetas_c(s, qt), theta_c(p0_half, T), pv_c(p0_half, qt, qv)/pv*(T)
    (thermodynamic_functions.h:9-12, 23-25, 36-41)."""
    from pycles_amd import parameters as P  # constants only
    p0 = np.asarray(p0_half, dtype=np.float64)[None, None, :]
    cpm = (1.0 - qt) * P.cpd + qt * P.cpv
    thetas = P.T_tilde * np.exp((s - (1.0 - qt) * P.sd_tilde - qt * P.sv_tilde) / cpm)
    theta = T / np.power(p0 / P.p_tilde, P.kappa)
    pv_star = np.vectorize(lt.__call__)(T)
    pv = p0 * P.eps_vi * qv / (1.0 - qt + P.eps_vi * qv)
    return thetas, theta, pv / pv_star


def liquid_stats(d: OrcDims, rho0_half, ql):
    """ThermodynamicsSA.liquid_stats (ThermodynamicsSA.pyx:395-492) over the
    interior: cloud fraction profile, cloud cover, cloud base and top (z_half
    = (k + 1/2) dz), liquid water path with interior level k weighted by
    rho0_half[k] as the reference indexes it (:478-480)."""
    gw = d.gw
    q = np.asarray(ql, dtype=np.float64)[gw:d.nlg[0] - gw, gw:d.nlg[1] - gw, gw:d.nlg[2] - gw]
    nz = d.n[2]
    div = float(d.n[0] * d.n[1])
    cf = np.zeros(nz)
    for k in range(nz):
        for _ in range(np.count_nonzero(q[:, :, k] > 0.0)):
            cf[k] += 1.0 / div          # one addition per cloudy cell (:447-451)
    cover = float(np.count_nonzero((q > 0.0).any(axis=2))) / div
    dz = d.dx[2]
    cb, ct = 99999.9, -99999.9
    for k in range(nz):
        if np.any(q[:, :, k] > 0.0):
            z = (k + 0.5) * dz
            cb, ct = min(cb, z), max(ct, z)
    lwp = 0.0
    cols = q.reshape(-1, nz)
    for col in cols:
        c = 0.0
        for k in range(nz):
            c += rho0_half[k] * col[k] * dz
        lwp += c
    return cf, cover, cb, ct, lwp / div
