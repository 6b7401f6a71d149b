"""Scalar and momentum advection operators (device).

``ScalarAdvection`` mirrors reference ``ScalarAdvection.pyx`` (constructor
reads ``scalar_transport.order``; ``update(Gr, RS, PV, DV, Pa)`` adds
-alpha0_half * d(flux)/dx for every scalar in x, y, z order).
``MomentumAdvection`` mirrors ``MomentumAdvection.pyx`` (``momentum_transport.order``;
``update(Gr, RS, PV, Pa)`` adds the nine (advected, advecting) pairs).
Both call ``pycles_advection``; ``advect_fused`` issues one call for both,
which gives bit-identical tendencies with one pass over the velocities.
Orders 5, 7, 9 are supported; other orders raise (the reference silently
falls back to WENO5 for scalars, scalar_advection.h:722-724, and to second
order for momentum, momentum_advection.h:87-91).
"""
from __future__ import annotations

from . import _capi
from . import _device as D

SUPPORTED = (5, 7, 9)


def _scalar_names(PV):
    return [n for n, t in zip(PV.index_name, PV.var_type) if t == 1]


def _vel(PV):
    return [PV.field(PV.velocity_names_directional[d]).data_ptr() for d in range(3)]


def _base(PV, names):
    if PV.precision != D.F32 or not names:
        return None
    b = [PV.base_ptr(n) for n in names] + [None, None]
    return _capi.ScalarBase(b[0], b[1])


class ScalarAdvection:
    def __init__(self, namelist, LH=None, Pa=None):
        try:
            self.order = int(namelist["scalar_transport"]["order"])
        except KeyError:
            raise ValueError("scalar_transport order not given in namelist")
        if self.order not in SUPPORTED:
            raise ValueError(f"scalar_transport order {self.order} not supported (5, 7, 9)")

    def initialize(self, Gr, PV, NS=None, Pa=None):
        self.dims = D.dims_of(Gr)

    def update(self, Gr, RS, PV, DV=None, Pa=None, tend_zero=False):
        names = _scalar_names(PV)
        if not names:
            return
        ref = D.device_ref(RS)
        _capi.advection(self.dims, ref.c, _vel(PV), None, [PV.ptr(n) for n in names],
                        [PV.tptr(n) for n in names], _base(PV, names), self.order, self.order,
                        PV.precision, D.stream(), _capi.ADV_TEND_ZERO if tend_zero else 0)


class MomentumAdvection:
    def __init__(self, namelist, Pa=None):
        try:
            self.order = int(namelist["momentum_transport"]["order"])
        except KeyError:
            raise ValueError("momentum_transport order not given in namelist")
        if self.order not in SUPPORTED:
            raise ValueError(f"momentum_transport order {self.order} not supported (5, 7, 9)")

    def initialize(self, Gr, PV, NS=None, Pa=None):
        self.dims = D.dims_of(Gr)

    def update(self, Gr, RS, PV, Pa=None, tend_zero=False):
        ref = D.device_ref(RS)
        tv = [PV.tptr(PV.velocity_names_directional[d]) for d in range(3)]
        _capi.advection(self.dims, ref.c, _vel(PV), tv, [], [], None, self.order, self.order,
                        PV.precision, D.stream(), _capi.ADV_TEND_ZERO if tend_zero else 0)


def yz_fused(SA: ScalarAdvection, MA: MomentumAdvection, PV) -> bool:
    """Whether one call's y and z passes run as one fused kernel (the
    library does so for FP32 storage, WENO5 scalars and momentum, with at
    least one scalar); otherwise they are two kernels."""
    return PV.precision == D.F32 and SA.order == 5 and MA.order == 5 and bool(_scalar_names(PV))


def advect_rk3(SA: ScalarAdvection, MA: MomentumAdvection, TS, Gr, RS, PV, tend_zero=False, passes=0,
               arrays=None):
    """advect_fused followed by TS.update in one sweep (pycles_advection_rk3).
    arrays = (out, copy) from PV.rk_arrays(TS.rk_step) (three rotating value
    arrays; no stage copy written); None: the new values go to PV's second
    array and the copies to TS's copy array.  The caller makes the output
    array current afterwards."""
    names = _scalar_names(PV)
    ref = D.device_ref(RS)
    vn = PV.velocity_names_directional
    tv = [PV.tptr(vn[d]) for d in range(3)]
    if arrays is None:
        outs = [PV.ptr_next(vn[d]) for d in range(3)] + [PV.ptr_next(n) for n in names]
        cps = [TS.copy_ptr(PV, vn[d]) for d in range(3)] + [TS.copy_ptr(PV, n) for n in names]
    else:
        out, cp = arrays
        outs = [out[PV.get_nv(vn[d])].data_ptr() for d in range(3)] + [out[PV.get_nv(n)].data_ptr() for n in names]
        cps = None if cp is None else ([cp[PV.get_nv(vn[d])].data_ptr() for d in range(3)] +
                                       [cp[PV.get_nv(n)].data_ptr() for n in names])
    _capi.advection_rk3(SA.dims, ref.c, _vel(PV), tv, [PV.ptr(n) for n in names],
                        [PV.tptr(n) for n in names], _base(PV, names), SA.order, MA.order,
                        PV.precision, D.stream(), (_capi.ADV_TEND_ZERO if tend_zero else 0) | passes,
                        outs[:3], outs[3:], None if cps is None else cps[:3], None if cps is None else cps[3:],
                        TS.rk_step, TS.dt)


def advect_fused(SA: ScalarAdvection, MA: MomentumAdvection, Gr, RS, PV, tend_zero=False, passes=0):
    """SA.update followed by MA.update in one launch sequence.  tend_zero:
    the scalar, u and v tendencies are zero (or stale and to be replaced) on
    entry, so they are written rather than accumulated (same bits).  passes:
    0 for the full update, or an OR of _capi.ADV_PASS_X/Y/Z to run only those
    directional kernels (timing)."""
    names = _scalar_names(PV)
    ref = D.device_ref(RS)
    tv = [PV.tptr(PV.velocity_names_directional[d]) for d in range(3)]
    _capi.advection(SA.dims, ref.c, _vel(PV), tv, [PV.ptr(n) for n in names],
                    [PV.tptr(n) for n in names], _base(PV, names), SA.order, MA.order,
                    PV.precision, D.stream(), (_capi.ADV_TEND_ZERO if tend_zero else 0) | passes)
