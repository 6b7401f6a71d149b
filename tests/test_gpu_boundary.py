"""The drop-in boundary driven the way a caller holding the reference's
arrays drives it: reference-layout FP64 arrays ([i][j][k] with ghosts on every
side, Csrc/grid.h:2-24) cross into the device layout through
pycles_layout_ingress, one SSP-RK3 stage runs through the C ABI in the order
Simulation3d.run (Simulation3d.pyx:159-179) calls its operators, and
pycles_layout_egress hands reference-layout arrays back; they are compared with
the oracle's stage on the same arrays.  Also: the non-finite status of the CFL
read, and the pressure correction's z chunks on level counts whose chunk
split leaves a remainder."""
import numpy as np
import pytest
import torch

from tests.helpers import host_case, oracle_sim, rel_err

pytestmark = pytest.mark.gpu


def _periodic_image(a, gw):
    """The reference array with its x/y ghost columns replaced by the
    periodic images of the interior, as update_all_bcs leaves them on one
    rank (PrognosticVariables.pyx:183-233)."""
    out = a.copy()
    nx = a.shape[0] - 2 * gw
    ny = a.shape[1] - 2 * gw
    ii = (np.arange(a.shape[0]) - gw) % nx + gw
    jj = (np.arange(a.shape[1]) - gw) % ny + gw
    return out[ii][:, jj]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("gw", [3, 5])
def test_layout_round_trip_f64_bitwise(gpu, gw):
    from pycles_amd import _capi, _device as D
    nx, ny, nz = 24, 16, 20
    dims = _capi.Dims((nx, ny, nz), gw, (50.0, 50.0, 20.0))
    rng = np.random.default_rng(1)
    a = rng.normal(size=(nx + 2 * gw, ny + 2 * gw, nz + 2 * gw))
    src = _dev(a)
    dev = torch.empty((nz + 2 * gw, ny, nx), dtype=torch.float64, device=gpu)
    _capi.layout_ingress(dims, src.data_ptr(), dev.data_ptr(), None, D.F64, D.stream())
    # device layout: [k][j][i] of the interior columns, every level
    want_dev = np.transpose(a[gw:gw + nx, gw:gw + ny, :], (2, 1, 0))
    assert np.array_equal(dev.cpu().numpy(), want_dev)
    back = torch.zeros_like(src)
    _capi.layout_egress(dims, dev.data_ptr(), back.data_ptr(), None, D.F64, D.stream())
    assert np.array_equal(back.cpu().numpy(), _periodic_image(a, gw))


def test_layout_round_trip_f32_with_base(gpu):
    from pycles_amd import _capi, _device as D
    nx, ny, nz, gw = 16, 8, 30, 3
    dims = _capi.Dims((nx, ny, nz), gw, (50.0, 50.0, 20.0))
    rng = np.random.default_rng(2)
    base = 6800.0 + 3.0 * np.arange(nz + 2 * gw)
    a = base[None, None, :] + rng.uniform(-2.0, 2.0, size=(nx + 2 * gw, ny + 2 * gw, nz + 2 * gw))
    src, b = _dev(a), _dev(base)
    dev = torch.empty((nz + 2 * gw, ny, nx), dtype=torch.float32, device=gpu)
    _capi.layout_ingress(dims, src.data_ptr(), dev.data_ptr(), b.data_ptr(), D.F32, D.stream())
    dev_dev = np.transpose(a[gw:gw + nx, gw:gw + ny, :] - base[None, None, :], (2, 1, 0))
    # the deviation is rounded once to FP32
    assert np.array_equal(dev.cpu().numpy(), dev_dev.astype(np.float32))
    back = torch.zeros_like(src)
    _capi.layout_egress(dims, dev.data_ptr(), back.data_ptr(), b.data_ptr(), D.F32, D.stream())
    want = _periodic_image(a, gw)
    dev_ref = want - base[None, None, :]
    # within half an FP32 ulp of the deviation plus the FP64 rounding of the sum
    bound = 2.0 ** -24 * np.abs(dev_ref) + 2.0 ** -52 * np.abs(want)
    assert np.all(np.abs(back.cpu().numpy() - want) <= bound)


def test_layout_entry_points_reject_bad_arguments(gpu):
    from pycles_amd import _capi, _device as D
    dims = _capi.Dims((8, 8, 8), 3, (1.0, 1.0, 1.0))
    buf = torch.zeros(14 ** 3, dtype=torch.float64, device=gpu)
    with pytest.raises(_capi.PyclesError):
        _capi.layout_ingress(dims, None, buf.data_ptr(), None, D.F64, D.stream())
    with pytest.raises(_capi.PyclesError):
        _capi.layout_egress(dims, buf.data_ptr(), buf.data_ptr(), None, 5, D.stream())


def test_one_stage_through_the_c_abi_matches_the_oracle(gpu):
    """INTEGRATION.md's call order on reference-layout arrays: ingress,
    thermodynamics, advection, RK3 stage with the boundary fill, pressure
    projection, egress; FP64, against the oracle's stage (Simulation3d.pyx:
    141-179: boundary fill and thermodynamics before the loop, then stage 0)."""
    from pycles_amd import _capi, _device as D
    nl, Gr, RS, Th, fields = host_case(32, 16, 24, seed=21)
    osim = oracle_sim(Gr, RS, Th, fields, True, max_iter=0)
    osim.update_all_bcs()
    names = ("u", "v", "w", "s", "qt")
    ref_arrays = {n: osim.pv(n).copy() for n in names}  # the caller's PV.values
    d = Gr.dims
    nzg = d.nzg
    dims = D.dims_of(Gr)
    st = D.stream()
    ref = D.device_ref(RS)
    # ingress
    vals = torch.empty((5, nzg, d.n[1], d.n[0]), dtype=torch.float64, device=gpu)
    for q, n in enumerate(names):
        src = _dev(ref_arrays[n])
        _capi.layout_ingress(dims, src.data_ptr(), vals[q].data_ptr(), None, D.F64, st)
    tend = torch.zeros_like(vals)
    copies = torch.zeros_like(vals)
    dvn = ("temperature", "qv", "ql", "alpha", "buoyancy", "theta_rho", "buoyancy_frequency", "thetali")
    dv = {n: torch.zeros((nzg, d.n[1], d.n[0]), dtype=torch.float64, device=gpu) for n in dvn}
    Th._prepare(Gr)
    opt = _capi.SAOptions(True, 0, _capi.LH_VARIABLE, None, 1, None, False)
    p = lambda q: vals[q].data_ptr()
    tp = lambda q: tend[q].data_ptr()

    def thermo():
        _capi.thermo_sa(dims, ref.c, Th._lt, p(3), p(4), tp(2), *[dv[n].data_ptr() for n in dvn], None, opt,
                        D.F64, st)

    thermo()            # before the loop: the buoyancy stays in the w tendency
    thermo()            # stage 0's thermodynamics adds it again, as the reference does
    _capi.advection(dims, ref.c, [p(0), p(1), p(2)], [tp(0), tp(1), tp(2)], [p(3), p(4)], [tp(3), tp(4)],
                    _capi.ScalarBase(), 5, 5, D.F64, st)
    _capi.rk3_stage(dims, [p(q) for q in range(5)], [tp(q) for q in range(5)],
                    [copies[q].data_ptr() for q in range(5)], [1.0, 1.0, -1.0, 1.0, 1.0], 0, 1.0, D.F64, st, 1)
    pres = torch.zeros((nzg, d.n[1], d.n[0]), dtype=torch.float64, device=gpu)
    div = torch.zeros_like(pres)
    rho0 = np.ascontiguousarray(RS.rho0)
    rho0h = np.ascontiguousarray(RS.rho0_half)
    plan = _capi.PressurePlan(dims, rho0.ctypes.data, rho0h.ctypes.data, D.F64)
    plan.update(ref.c, p(0), p(1), p(2), pres.data_ptr(), div.data_ptr(), 1.0, None, st)
    # egress
    out = {}
    for q, n in enumerate(names):
        dst = torch.zeros(ref_arrays[n].shape, dtype=torch.float64, device=gpu)
        _capi.layout_egress(dims, p(q), dst.data_ptr(), None, D.F64, st)
        out[n] = dst
    pout = torch.zeros(ref_arrays["u"].shape, dtype=torch.float64, device=gpu)
    _capi.layout_egress(dims, pres.data_ptr(), pout.data_ptr(), None, D.F64, st)
    torch.cuda.synchronize()
    # the oracle: begin_run's thermodynamics (boundary fill above), then stage 0
    osim.thermo_update()
    osim.stage(0)
    errs = {n: rel_err(out[n].cpu().numpy(), osim.pv(n)) for n in names}
    errs["p"] = rel_err(pout.cpu().numpy(), _periodic_image(osim.dv["dynamic_pressure"], d.gw))
    print("one C-ABI stage vs oracle (reference layout, ghosts included)", errs)
    assert max(errs.values()) <= 1e-12, errs


def test_cfl_read_flags_a_nan_velocity(gpu):
    from pycles_amd import _capi, _device as D
    nl, Gr, RS, Th, fields = host_case(16, 8, 12, seed=10)
    d = Gr.dims
    u, v, w = (_dev(fields[n]) for n in ("u", "v", "w"))
    dims = D.dims_of(Gr)
    out = torch.zeros(1, dtype=torch.float64, device=gpu)
    _capi.cfl_max(dims, u.data_ptr(), v.data_ptr(), w.data_ptr(), 1.0, out.data_ptr(), D.F64, D.stream())
    c = _capi.cfl_read(out.data_ptr(), D.stream())
    assert np.isfinite(c) and c > 0.0
    v[d.gw + 3, 2, 5] = float("nan")
    _capi.cfl_max(dims, u.data_ptr(), v.data_ptr(), w.data_ptr(), 1.0, out.data_ptr(), D.F64, D.stream())
    with pytest.raises(_capi.PyclesError, match="error -4"):
        _capi.cfl_read(out.data_ptr(), D.stream())
    # the projection's fused CFL carries it too
    v[d.gw + 3, 2, 5] = float("inf")
    from pycles_amd.Simulation3d import Simulation3d
    f = dict(fields)
    f["v"] = v.cpu().numpy()
    sim = Simulation3d().initialize(nl, Gr=Gr, Ref=RS, Th=Th, fields=f)
    sim.Pr.update(sim.Gr, sim.Ref, sim.DV, sim.PV, sim.Pa, cfl_out=sim.TS.cfl_buffer(), dt=1.0)
    with pytest.raises(_capi.PyclesError, match="not finite"):
        sim.TS.compute_cfl_max(sim.Gr, sim.PV)
