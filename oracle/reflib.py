"""ctypes face of oracle/_ref/libpycles_ref.so: the reference's own C kernels.

TEST INFRASTRUCTURE ONLY.  ``make -C oracle ref`` compiles, straight from the
reference tree, the headers that build without generated code:
``Csrc/momentum_advection.h`` (with the interpolation and flux-divergence
headers it includes), ``Csrc/prognostic_variables.h``, ``Csrc/kinematics.h``
and ``Csrc/momentum_diffusion.h``.  Tests use it to
check the oracle bit for bit; it is never shipped and never measured.  It
exists only where the reference tree is present, so tests that use it skip
elsewhere.

Arrays follow the reference layout: one variable is ``[nlg0, nlg1, nlg2]``
C-order, z fastest.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_ref", "libpycles_ref.so")
_KIN_PATH = os.path.join(_HERE, "_ref", "libpycles_ref_kin.so")
REF_SRC = "/root/reference/Csrc"
_lib = None
_kin = None
_dp = C.POINTER(C.c_double)


class DimStruct(C.Structure):
    """``struct DimStruct`` of Csrc/grid.h:2-25."""
    _fields_ = [("dims", C.c_long), ("n", C.c_long * 3), ("ng", C.c_long * 3),
                ("nl", C.c_long * 3), ("nlg", C.c_long * 3), ("npd", C.c_long),
                ("npl", C.c_long), ("npg", C.c_long), ("gw", C.c_long),
                ("indx_lo", C.c_long * 3), ("indx_lo_g", C.c_long * 3),
                ("nbuffer", C.c_long * 3), ("ghosted_stride", C.c_long * 3),
                ("dx", C.c_double * 3), ("dxi", C.c_double * 3)]


def available() -> bool:
    """True when the library is built or the reference sources are here to build it."""
    return os.path.exists(_LIB_PATH) or os.path.isdir(REF_SRC)


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        d = C.c_double
        for name, nargs in (("interp_4", 4), ("interp_6", 6), ("interp_8", 8),
                            ("interp_weno5", 5), ("interp_weno7", 7), ("interp_weno9", 9)):
            f = getattr(L, name)
            f.restype = d
            f.argtypes = [d] * nargs
        L.compute_advective_tendencies_m.argtypes = [C.POINTER(DimStruct), _dp, _dp, _dp, _dp, _dp, _dp,
                                                      _dp, C.c_ssize_t, C.c_ssize_t, C.c_int]
        L.scalar_flux_divergence.argtypes = [C.POINTER(DimStruct), _dp, _dp, _dp, _dp, d, C.c_ssize_t]
        L.build_buffer.argtypes = [C.c_long, C.c_long, C.c_long, C.POINTER(DimStruct), _dp, _dp]
        L.buffer_to_values.argtypes = [C.c_ssize_t, C.c_ssize_t, C.POINTER(DimStruct), _dp, _dp]
        L.set_bcs.argtypes = [C.c_ssize_t, C.c_ssize_t, d, C.POINTER(DimStruct), _dp]
        L.compute_diffusive_flux_m.argtypes = [C.POINTER(DimStruct), _dp, _dp, _dp, _dp, _dp, C.c_ssize_t,
                                               C.c_ssize_t]
        L.momentum_flux_divergence.argtypes = [C.POINTER(DimStruct), _dp, _dp, _dp, _dp, C.c_ssize_t, C.c_ssize_t]
        L.compute_entropy_source.argtypes = [C.POINTER(DimStruct), _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def kin():
    global _kin
    if _kin is None:
        if not os.path.exists(_KIN_PATH):
            build()
        K = C.CDLL(_KIN_PATH)
        K.compute_velocity_gradient.argtypes = [C.POINTER(DimStruct), _dp, _dp, C.c_ssize_t]
        K.compute_strain_rate.argtypes = [C.POINTER(DimStruct), _dp, _dp]
        K.compute_strain_rate_mag.argtypes = [C.POINTER(DimStruct), _dp, _dp]
        _kin = K
    return _kin


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def make_dims(nx, ny, nz, gw, dx, dy, dz) -> DimStruct:
    """One-rank DimStruct filled as Grid.pyx:25-121 fills it."""
    s = DimStruct()
    s.dims = 3
    n = (nx, ny, nz)
    for i in range(3):
        s.n[i] = n[i]
        s.ng[i] = n[i] + 2 * gw
        s.nl[i] = n[i]
        s.nlg[i] = n[i] + 2 * gw
        s.indx_lo[i] = 0
        s.indx_lo_g[i] = 0
    s.gw = gw
    s.npd = nx * ny * nz
    s.npl = nx * ny * nz
    s.npg = s.nlg[0] * s.nlg[1] * s.nlg[2]
    g = s.nlg
    s.nbuffer[0] = gw * max(g[1] * g[2], g[1], g[2])
    s.nbuffer[1] = gw * max(g[0] * g[2], g[0], g[2])
    s.nbuffer[2] = gw * max(g[0] * g[1], g[0], g[1])
    s.ghosted_stride[0] = max(g[1] * g[2], g[1], g[2])
    s.ghosted_stride[1] = max(1, g[0])
    s.ghosted_stride[2] = 1
    for i, h in enumerate((dx, dy, dz)):
        s.dx[i] = h
        s.dxi[i] = 1.0 / h
    return s


def interp(name: str, *q) -> float:
    return getattr(lib(), name)(*[float(x) for x in q])


def momentum_tendencies(dims: DimStruct, rho0, rho0_half, alpha0, alpha0_half, vel, order):
    """MomentumAdvection.update (MomentumAdvection.pyx:54-68): every
    (advected, advecting) pair through compute_advective_tendencies_m,
    accumulating into zeroed tendencies in the reference's loop order."""
    L = lib()
    tend = [np.zeros_like(v) for v in vel]
    for a in range(3):
        for b in range(3):
            L.compute_advective_tendencies_m(C.byref(dims), ptr(rho0), ptr(rho0_half), ptr(alpha0),
                                             ptr(alpha0_half), ptr(vel[a]), ptr(vel[b]), ptr(tend[a]),
                                             a, b, order)
    return tend


def scalar_flux_divergence(dims: DimStruct, alpha0, alpha0_half, flux, tend, d):
    lib().scalar_flux_divergence(C.byref(dims), ptr(alpha0), ptr(alpha0_half), ptr(flux), ptr(tend),
                                 dims.dx[d], d)


def update_all_bcs(dims: DimStruct, values: np.ndarray, bc_factor: float):
    """PrognosticVariables.update_all_bcs (PrognosticVariables.pyx:183-233) for
    one variable on one rank: x and y are periodic, so the Sendrecv returns the
    rank's own send buffer; z has no neighbour, so set_bcs fills it."""
    L = lib()
    buf = np.zeros(max(dims.nbuffer[0], dims.nbuffer[1], dims.nbuffer[2]))
    for d in range(3):
        for s in (-1, 1):
            L.build_buffer(0, d, s, C.byref(dims), ptr(values), ptr(buf))
            if d < 2:
                L.buffer_to_values(d, s, C.byref(dims), ptr(values), ptr(buf))
            else:
                L.set_bcs(d, s, bc_factor, C.byref(dims), ptr(values))


def kinematics(dims: DimStruct, vel):
    """Kinematics.update (Kinematics.pyx:41-58) with the reference's
    compute_velocity_gradient / compute_strain_rate / compute_strain_rate_mag:
    returns (vgrad[9 npg], strain[9 npg], strain_mag[npg])."""
    L = kin()
    npg = dims.npg
    vgrad = np.zeros(9 * npg)
    # compute_strain_rate_mag reads up to one x-plane before the strain array
    # (i = 1, the (0,0) pair at ijk - 2 istride): give it zeros to read
    pad = dims.nlg[1] * dims.nlg[2]
    sbuf = np.zeros(pad + 9 * npg)
    strain = sbuf[pad:]
    smag = np.zeros(npg)
    count = 0
    for vi in range(3):
        for d in range(3):
            g = vgrad[count * npg:(count + 1) * npg]
            L.compute_velocity_gradient(C.byref(dims), ptr(vel[vi].reshape(-1)), ptr(g), d)
            count += 1
    L.compute_strain_rate(C.byref(dims), ptr(vgrad), ptr(strain))
    L.compute_strain_rate_mag(C.byref(dims), ptr(strain), ptr(smag))
    return vgrad, strain, smag


def momentum_diffusion(dims: DimStruct, rho0, rho0_half, alpha0, alpha0_half, strain, visc, tend, smag, T,
                       s_tend):
    """MomentumDiffusion.update (MomentumDiffusion.pyx:44-66): the nine
    stress fluxes, their divergence into tend[3] and the entropy source into
    s_tend, with the reference's own functions."""
    L = lib()
    npg = dims.npg
    flux = np.zeros(npg)
    count = 0
    for i1 in range(3):
        for i2 in range(3):
            L.compute_diffusive_flux_m(C.byref(dims), ptr(strain[count * npg:(count + 1) * npg]), ptr(visc),
                                       ptr(flux), ptr(rho0), ptr(rho0_half), i1, i2)
            L.momentum_flux_divergence(C.byref(dims), ptr(alpha0), ptr(alpha0_half), ptr(flux), ptr(tend[i1]),
                                       i1, i2)
            count += 1
    L.compute_entropy_source(C.byref(dims), ptr(visc), ptr(smag), ptr(T), ptr(s_tend))
