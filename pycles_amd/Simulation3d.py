"""Driver of the device dynamical core (reference ``Simulation3d.pyx``).

``initialize(namelist)`` builds the operators in the reference's order
(``Simulation3d.pyx:40-138``) restricted to the hot path; ``run()`` follows
``:140-189``: one boundary fill and one thermodynamics update before the loop
(the latter adds buoyancy to the w tendency, which the first stage keeps), then
for each RK stage Ke.update, Th.update, SA.update, MA.update, Sur.update,
SGS.update, Damping.update, SD.update and MD.update, Fo.update (the SGS and
diffusion operators with an sgs section in the namelist; surface and forcing
for the Bomex case; damping with a damping section), TS.update (with the z
boundary fill), Pr.update and, on the last stage, TS.adjust_timestep and the
restart and statistics output.  The operators outside this path (radiation,
tracers, microphysics processes, other cases' surface and forcing) are not
part of this build and contribute nothing, as with a namelist that disables
them.
"""
from __future__ import annotations

import time

from . import _device as D
from .Advection import MomentumAdvection, ScalarAdvection, advect_fused, advect_rk3, yz_fused
from .Grid import Grid
from .ParallelMPI import NetCDFIO_Stats, ParallelMPI
from .PressureSolver import PressureSolver
from .PrognosticVariables import DiagnosticVariables, PrognosticVariables
from .ReferenceState import ReferenceState
from .Restart import Restart
from .Thermodynamics import LatentHeat
from .ThermodynamicsSA import ThermodynamicsFactory
from .TimeStepping import TimeStepping
from . import cases


class Simulation3d:
    def __init__(self, namelist=None):
        self.fuse_advection = True
        self.write_tendencies = True
        # the RK3 update runs inside the advection's z pass (needs the two above)
        self.fuse_rk = True
        self._wt_pending = False

    def initialize(self, namelist, Gr=None, Ref=None, Th=None, fields=None):
        """Build the run.  ``Gr``, ``Ref``, ``Th`` and ``fields`` (device-layout
        FP64 host arrays per prognostic name) may be passed in to start from an
        already built state instead of running the case initializer."""
        self.namelist = namelist
        precision = namelist.get("mi355x", {}).get("precision", "f64")
        self.Pa = ParallelMPI(namelist)
        self.Gr = Gr if Gr is not None else Grid(namelist, self.Pa)
        self.PV = PrognosticVariables(self.Gr)
        self.DV = DiagnosticVariables()
        self.Pr = PressureSolver()
        self.LH = LatentHeat(namelist, self.Pa)
        self.SA = ScalarAdvection(namelist, self.LH, self.Pa)
        self.MA = MomentumAdvection(namelist, self.Pa)
        self.Th = Th if Th is not None else ThermodynamicsFactory(namelist, None, self.LH, self.Pa)
        self.Ref = Ref if Ref is not None else ReferenceState(self.Gr)
        # kinematics, SGS closure and diffusion when the namelist has an
        # sgs section (Simulation3d.pyx:44-53)
        self.Ke = self.SGS = self.SD = self.MD = None
        if "sgs" in namelist:
            from .SGS import SGS, Kinematics, MomentumDiffusion, ScalarDiffusion
            self.Ke = Kinematics()
            self.SGS = SGS(namelist)
            self.SD = ScalarDiffusion(namelist, self.LH, self.DV, self.Pa)
            self.MD = MomentumDiffusion(self.DV, self.Pa)
        # case physics: surface, large-scale forcing and damping by the
        # namelist's casename and damping section (Simulation3d.pyx:54-62)
        from .CasePhysics import Damping, Dummy, Forcing, ForcingNone, SurfaceFactory, SurfaceNone
        self.Sur = SurfaceFactory(namelist, self.LH, self.Pa)
        self.Fo = Forcing(namelist, self.LH, self.Pa)
        self.Damping = Damping(namelist, self.Pa)
        self.case_physics = not (isinstance(self.Sur, SurfaceNone) and isinstance(self.Fo.scheme, ForcingNone)
                                 and isinstance(self.Damping.scheme, Dummy))
        self.TS = TimeStepping()
        self.StatsIO = NetCDFIO_Stats().initialize(namelist, None, None)
        self.Restart = Restart(namelist, self.Pa)
        self.Restart.initialize()
        self.PV.add_variable("u", "m/s", "u", "u velocity component", "sym", "velocity", self.Pa)
        self.PV.set_velocity_direction("u", 0, self.Pa)
        self.PV.add_variable("v", "m/s", "v", "v velocity component", "sym", "velocity", self.Pa)
        self.PV.set_velocity_direction("v", 1, self.Pa)
        self.PV.add_variable("w", "m/s", "w", "w velocity component", "asym", "velocity", self.Pa)
        self.PV.set_velocity_direction("w", 2, self.Pa)
        self.Th.initialize(self.Gr, self.PV, self.DV, self.StatsIO, self.Pa)
        self.PV.initialize(self.Gr, self.StatsIO, self.Pa, precision=precision)
        self.SA.initialize(self.Gr, self.PV, self.StatsIO, self.Pa)
        self.MA.initialize(self.Gr, self.PV, self.StatsIO, self.Pa)
        if self.SGS is not None:
            self.Ke.initialize(self.Gr, self.StatsIO, self.Pa, precision=precision)
            self.SGS.initialize(self.Gr, self.PV, self.StatsIO, self.Pa)
        self.TS.initialize(namelist, self.PV, self.Pa)
        if self.Restart.is_restart_run:
            # Simulation3d.pyx:103-121: state, time and reference from the file
            self.Restart.read(self.Pa)
            self.TS.t = self.Restart.restart_data["TS"]["t"]
            self.TS.dt = self.Restart.restart_data["TS"]["dt"]
            self.Ref.init_from_restart(self.Gr, self.Restart)
            self.PV.init_from_restart(self.Gr, self.Restart)
            if hasattr(self.Th, "init_from_restart"):
                self.Th.init_from_restart(self.Restart)
            if hasattr(self.Sur, "init_from_restart") and "surf" in self.Restart.restart_data:
                self.Sur.init_from_restart(self.Restart)
            self.Restart.last_restart_time = self.Restart.restart_data["last_restart_time"]
            self.Restart.free_memory()
            fields = {}
        elif fields is None:
            fields = cases.initialize(namelist, self.Gr, self.PV, self.Ref, self.Th)
        else:
            for name, arr in fields.items():
                if name in ("s", "qt"):
                    self.PV.set_base_from_device_layout(name, arr)
                self.PV.set_device_layout(name, arr)
        self.initial_fields = fields
        self.Sur.initialize(self.Gr, self.Ref, self.StatsIO, self.Pa)
        self.Fo.initialize(self.Gr, self.Ref, self.Th, self.StatsIO, self.Pa)
        self.Damping.initialize(self.Gr, self.Ref)
        self.Pr.initialize(namelist, self.Gr, self.Ref, self.DV, self.Pa)
        self.DV.initialize(self.Gr, self.StatsIO, self.Pa, precision=precision)
        self.walltimes = []
        return self

    # ---------------------------------------------------------------------
    def begin_run(self):
        """Simulation3d.run before its loop: boundary fill, thermodynamics."""
        self.TS._zbcs = D.dims_of(self.Gr)
        from . import _capi
        names = self.PV.index_name
        _capi.apply_z_bcs(self.TS._zbcs, [self.PV.ptr(n) for n in names],
                          [self.PV.bc_type[self.PV.get_nv(n)] for n in names], self.PV.precision,
                          D.stream())
        self.Th.update(self.Gr, self.Ref, self.PV, self.DV)
        # the buoyancy just added stays in the w tendency for the first stage
        # (the reference's run() does this on every start).  A restarted run
        # continues as the run that wrote the file would have, whose w
        # tendency started the stage at zero: it is cleared again.
        if self.Restart.is_restart_run:
            self.PV.zero_tendencies()
        self._wt_pending = not self.Restart.is_restart_run

    def _mark(self, name):
        """Record a HIP event on the working stream when timing is enabled."""
        if self.timers is not None:
            import torch
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
            self.timers.append((name, ev))

    def stage(self, rk_step):
        """One RK stage.  With ``write_tendencies`` (default) the RK update does
        not zero the tendencies; instead the first operator to touch each
        tendency in the next stage writes it (thermodynamics the w tendency,
        advection the others).  0 + x and 0 - x are exact, so this gives the
        same bits as zeroing and accumulating, with ten fewer field passes."""
        self.TS.rk_step = rk_step
        fast = self.write_tendencies
        if self.Ke is not None:
            self.Ke.update(self.Gr, self.PV)
        self._mark("thermo")
        self.Th.update(self.Gr, self.Ref, self.PV, self.DV, wt_overwrite=fast and not self._wt_pending)
        self._wt_pending = False
        # with SGS and diffusion the RK update runs after them (their
        # tendencies add to the advection's first, as in the reference)
        fused_rk = (self.fuse_rk and self.fuse_advection and fast and self.SGS is None
                    and not self.case_physics)
        if fused_rk:
            if self.timers is not None:
                from ._device import _capi
                if yz_fused(self.SA, self.MA, self.PV):
                    passes = (("adv_x", _capi.ADV_PASS_X), ("adv_yz_rk3", _capi.ADV_PASS_Y | _capi.ADV_PASS_Z))
                else:
                    passes = (("adv_x", _capi.ADV_PASS_X), ("adv_y", _capi.ADV_PASS_Y),
                              ("adv_z_rk3", _capi.ADV_PASS_Z))
            else:
                passes = (("advection_rk3", 0),)
            arrays = self.PV.rk_arrays(rk_step)
            for name, bit in passes:
                self._mark(name)
                advect_rk3(self.SA, self.MA, self.TS, self.Gr, self.Ref, self.PV, tend_zero=fast, passes=bit,
                           arrays=arrays)
            self.PV.set_values(arrays[0])
            self.TS.advance()
        elif self.fuse_advection and self.timers is not None:
            # the same kernels as the fused call, with an event between each
            # so every kernel is timed on its own
            from ._device import _capi
            if yz_fused(self.SA, self.MA, self.PV):
                passes = (("adv_x", _capi.ADV_PASS_X), ("adv_yz", _capi.ADV_PASS_Y | _capi.ADV_PASS_Z))
            else:
                passes = (("adv_x", _capi.ADV_PASS_X), ("adv_y", _capi.ADV_PASS_Y), ("adv_z", _capi.ADV_PASS_Z))
            for name, bit in passes:
                self._mark(name)
                advect_fused(self.SA, self.MA, self.Gr, self.Ref, self.PV, tend_zero=fast, passes=bit)
        elif self.fuse_advection:
            self._mark("advection")
            advect_fused(self.SA, self.MA, self.Gr, self.Ref, self.PV, tend_zero=fast)
        else:
            self._mark("advection")
            self.SA.update(self.Gr, self.Ref, self.PV, self.DV, self.Pa, tend_zero=fast)
            self.MA.update(self.Gr, self.Ref, self.PV, self.Pa, tend_zero=fast)
        # Simulation3d.pyx:165-171: Sur, SGS, Damping, SD, MD, Fo
        if self.case_physics:
            self._mark("surface")
            self.Sur.update(self.Gr, self.Ref, self.PV, self.DV, self.Pa, self.TS)
        if self.SGS is not None:
            self._mark("sgs")
            self.SGS.update(self.Gr, self.DV, self.PV, self.Ke, self.Sur, self.Pa)
        if self.case_physics:
            self._mark("damping")
            self.Damping.update(self.Gr, self.Ref, self.PV, self.DV, self.Pa)
        if self.SGS is not None:
            from .SGS import diffuse
            self._mark("diffusion")
            diffuse(self.SD, self.MD, self.Gr, self.Ref, self.PV, self.DV, self.Ke)
        if self.case_physics:
            self._mark("forcing")
            self.Fo.update(self.Gr, self.Ref, self.PV, self.DV, self.Pa)
        if not fused_rk:
            self._mark("rk3")
            self.TS.update(self.Gr, self.PV, self.Pa, zero_tendencies=not fast)
        last = rk_step == self.TS.n_rk_steps - 1
        self._mark("pressure")
        self.Pr.update(self.Gr, self.Ref, self.DV, self.PV, self.Pa,
                       cfl_out=self.TS.cfl_buffer() if last else None, dt=self.TS.dt)
        self._mark("end")
        self.TS.adjust_timestep(self.Gr, self.PV, self.DV, self.Pa)
        self.io()

    def io(self):
        """The restart part of Simulation3d.io (Simulation3d.pyx:191-283):
        on the last stage, shorten dt to land on the next restart time, and
        at that time dump the restart file (statistics output is the
        Statistics module's, called by the caller)."""
        TS, Re, NS = self.TS, self.Restart, self.StatsIO
        stats_on = NS.path is not None
        if not ((Re.enabled or stats_on) and TS.t > 0 and TS.rk_step == TS.n_rk_steps - 1):
            return
        dts = [TS.dt, TS.dt_max]
        if stats_on:
            dts += [NS.last_output_time + NS.frequency - TS.t, NS.frequency]
        if Re.enabled:
            dts += [Re.last_restart_time + Re.frequency - TS.t, Re.frequency]
        TS.dt = min(x for x in dts if x > 0.0)
        if stats_on and NS.last_output_time + NS.frequency == TS.t:
            NS.last_output_time = TS.t
            self.stats_io()
        if Re.enabled and Re.last_restart_time + Re.frequency == TS.t:
            self.write_restart()

    def stats_io(self):
        """The statistics output of Simulation3d.io (Simulation3d.pyx:227-256)
        for the operators on the device path: every reduction runs on the GPU
        and only the profiles come to the host."""
        NS = self.StatsIO
        NS.open_files(self.Pa)
        NS.write_simulation_time(self.TS.t, self.Pa)
        self.PV.stats_io(self.Gr, self.Ref, NS, self.Pa)
        self.DV.stats_io(self.Gr, NS, self.Pa)
        if hasattr(self.Th, "stats_io"):
            self.Th.stats_io(self.Gr, self.Ref, self.PV, self.DV, NS, self.Pa)
        self.Sur.stats_io(self.Gr, NS, self.Pa)
        return NS.close_files(self.Pa)

    def write_restart(self):
        """Dump the restart file for the current time (Simulation3d.pyx:269-283)."""
        Re = self.Restart
        Re.last_restart_time = self.TS.t
        self.Gr.restart(Re)
        self.Ref.restart(self.Gr, Re)
        self.PV.restart(self.Gr, Re)
        self.TS.restart(Re)
        if hasattr(self.Sur, "restart"):
            self.Sur.restart(Re)
        if hasattr(self.Th, "restart"):
            self.Th.restart(Re)
        return Re.write(self.Pa)

    timers = None

    def phase_times(self):
        """Milliseconds per phase summed over the recorded stages."""
        import torch
        torch.cuda.synchronize()
        out = {}
        t = self.timers or []
        for (name, ev), (_, nxt) in zip(t, t[1:]):
            if name == "end":
                continue
            out[name] = out.get(name, 0.0) + ev.elapsed_time(nxt)
        return out

    def step(self):
        for rk in range(self.TS.n_rk_steps):
            self.stage(rk)

    def run(self, max_steps=None):
        self.begin_run()
        n = 0
        while self.TS.t < self.TS.t_max and (max_steps is None or n < max_steps):
            t0 = time.time()
            self.step()
            self.walltimes.append(time.time() - t0)
            self.Pa.root_print(f"T = {self.TS.t} dt = {self.TS.dt} cfl_max = {self.TS.cfl_max} "
                               f"walltime = {self.walltimes[-1]}")
            n += 1
        self.Restart.cleanup()
        return self
