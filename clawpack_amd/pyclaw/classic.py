"""ClawSolver2D with kernel_language='HIP' (upstream pyclaw/src/pyclaw/classic/solver.py).

The unsplit classic step (step2/flux2/limiter with rpn2/rpt2) and the evolve_to_time loop run on
the GPU through libclawpack_amd.so: one fused kernel per step attempt plus a device-side
accept/reject/dt controller, so the host only synchronises at output times.  With a step_source
or before_step callback (upstream Solver.step: Godunov or Strang source splitting) the loop runs
on the host, one step launch per attempt, and the callbacks act on the device state through torch
(DeviceState).  A BC.custom side (user_bc_lower/user_bc_upper) also runs the host loop: each attempt
fills a ghost-padded copy of the state on the device in pyclaw's order, calling the user's function
on torch views of it, and steps from it (claw2d_step_padded; with dimensional_split the two
sweeps of claw2d_sweep_padded, with the same ghost fill of q* between them).  There is no CPU
compute path: a missing library or device raises.
"""
import math

import numpy as np

import torch

from .. import _lib
from ..device import DeviceGrid, make_desc
from .limiters import tvd
from .solver import BC, Solver

_KERNEL_LANGUAGES = ("HIP",)


def _fill_side(v, axis, side, bc, g, reflect):
    """pyclaw's _bc_lower/_bc_upper for one standard rule on the padded pyclaw-layout view v
    (num_eqn, mx+2g, my+2g): full slices along the other axis; reflect < 0 for aux (no negation)."""
    n = v.shape[axis]
    for i in range(g):
        dst = i if side == 0 else n - 1 - i
        if bc == BC.extrap:
            src = g if side == 0 else n - 1 - g
        elif bc == BC.periodic:
            src = n - 2 * g + i if side == 0 else 2 * g - 1 - i
        elif bc == BC.wall:
            src = 2 * g - 1 - i if side == 0 else n - 2 * g + i
        else:
            raise NotImplementedError("bc %r" % (bc,))
        v.select(axis, dst).copy_(v.select(axis, src))
        if bc == BC.wall and reflect >= 0:
            v[reflect].select(axis - 1, dst).neg_()


class DeviceState:
    """The state that step_source(solver, state, dt) and before_step(solver, state) receive on the
    HIP path.  q (num_eqn, mx, my) and aux (num_aux, mx, my) are torch device tensors in pyclaw's
    layout, views of the device planes: in-place updates (state.q[1] *= ..., state.q[1:3].copy_(...))
    change the device state, with no host copy.  problem_data, grid, patch and t are the host
    State's."""

    def __init__(self, state, grid):
        self.problem_data = state.problem_data
        self.grid = state.grid
        self.patch = state.patch
        self.index_capa = state.index_capa
        self.aux = None if grid.aux is None else grid.aux.permute(0, 2, 1)
        self.q = None
        self.t = state.t

    def bind(self, planes, t):
        self.q = planes.permute(0, 2, 1)
        self.t = t
        return self

    @property
    def num_eqn(self):
        return self.q.shape[0]

    @property
    def num_aux(self):
        return 0 if self.aux is None else self.aux.shape[0]


class ClawSolver2D(Solver):
    """Two-dimensional classic (wave-propagation) solver.

    Supported on the HIP path: dimensional_split True (step2ds, Godunov x-then-y splitting, the
    default) or False (unsplit step2 with transverse_waves 0/1/2), order 1 or 2, limiters 0-5 per
    wave, BC extrap/periodic/wall/custom (user_bc_lower/upper on torch views), the Riemann plugins shallow_roe_with_efix_2D,
    euler_4wave_2D and shallow_bathymetry_fwave_2D (with fwave = True, state.aux[0] = bathymetry
    and solver.aux_bc_lower/upper), solver.precision 'fp64' (default) or 'fp32'.
    """

    def __init__(self, riemann_solver=None, claw_package=None):
        super().__init__(riemann_solver, claw_package)
        self.num_dim = 2
        self.dimensional_split = True   # pyclaw's default: step2ds
        self.transverse_waves = 1
        self.order = 2
        self.limiters = tvd.minmod
        self.fwave = False
        self.kernel_language = "HIP"
        self.reflect_index = [1, 2]
        self.precision = "fp64"
        # FP32 shallow water: the device stores the depth as h - depth_ref (its FP32 rounding is
        # then relative to the deviation, not to h).  None: the mean initial depth, rounded to
        # FP32, when no user code sees the device state (no step_source, before_step or
        # BC.custom), else 0; a number sets it; 0 stores h.
        self.depth_ref = None
        self._auto_depth_ref = 0.0
        self.seg_h = 0          # rows per marching segment; 0 = occupancy-derived default
        self.autotune = False   # True: tile shape (seg_h, wg_waves, launch_order) from clawpack_amd.autotune
        self._grid = None
        self._desc_key = None

    # ---------------------------------------------------------------------------- set-up ---
    def _mthlim(self):
        lim = self.limiters if isinstance(self.limiters, (list, tuple)) else [self.limiters]
        if len(lim) == 1:
            lim = list(lim) * self.num_waves
        if len(lim) != self.num_waves:
            raise Exception("Length of solver.limiters must be 1 or num_waves")
        return [int(x) for x in lim]

    def _check_supported(self, state):
        if self.kernel_language not in _KERNEL_LANGUAGES:
            raise NotImplementedError(
                "kernel_language=%r is not available in clawpack_amd; the only compute path is "
                "'HIP' (the reference's f2py Fortran/Python kernels are not part of this package)"
                % (self.kernel_language,))
        pol = self._policy()
        rp_fwave = pol.fwave
        if bool(self.fwave) != rp_fwave:
            raise ValueError("solver.fwave must be %s for %s (its Riemann solver returns %s)"
                             % (rp_fwave, pol.name, "f-waves" if rp_fwave else "waves"))
        rp_aux = pol.num_aux
        capa = state.index_capa >= 0
        if capa:
            if state.index_capa >= state.num_aux:
                raise ValueError("state.index_capa = %d needs num_aux > %d (aux[index_capa] is the "
                                 "capacity function)" % (state.index_capa, state.index_capa))
            if self._has_custom_bc():
                raise NotImplementedError("a capacity function (index_capa) with BC.custom is not "
                                          "supported on the HIP path")
        if rp_aux == 0 and state.num_aux > 0 and not capa:
            raise NotImplementedError("%s reads no aux arrays; aux is supported with "
                                      "shallow_bathymetry_fwave_2D or a capacity function "
                                      "(state.index_capa) on the HIP path" % pol.name)
        if state.num_aux < rp_aux:
            raise ValueError("%s needs num_aux >= %d (aux[0] = bathymetry)" % (pol.name, rp_aux))
        if state.num_aux > 0:
            for k in range(2):
                lo, hi = self.aux_bc_lower[k], self.aux_bc_upper[k]
                if lo is None or hi is None:
                    raise Exception("Lower and upper aux boundary conditions must be specified for "
                                    "every dimension (solver.aux_bc_lower / solver.aux_bc_upper)")
                if BC.custom in (lo, hi):
                    raise NotImplementedError("BC.custom for aux is not supported on the HIP path")
                if (lo == BC.periodic) != (hi == BC.periodic):
                    raise Exception("Periodic aux boundary conditions must be set on both sides")
        if self.num_ghost != 2:
            raise NotImplementedError("num_ghost must be 2")
        if self.step_source is not None and self.source_split not in (1, 2):
            raise ValueError("source_split must be 1 (Godunov) or 2 (Strang)")
        if self.order not in (1, 2) or self.transverse_waves not in (0, 1, 2):
            raise ValueError("order must be 1 or 2 and transverse_waves 0, 1 or 2")
        for k in range(2):
            lo, hi = self.bc_lower[k], self.bc_upper[k]
            if lo is None or hi is None:
                raise Exception("Lower and upper boundary conditions must be specified for every "
                                "dimension (solver.bc_lower / solver.bc_upper)")
            if (lo == BC.periodic) != (hi == BC.periodic):
                raise Exception("Periodic boundary conditions must be set on both sides")
        custom = [BC.custom in (self.bc_lower[k], self.bc_upper[k]) for k in range(2)]
        if any(custom):
            if BC.custom in self.bc_lower and self.user_bc_lower is None:
                raise Exception("Lower custom BC function has not been set.")
            if BC.custom in self.bc_upper and self.user_bc_upper is None:
                raise Exception("Upper custom BC function has not been set.")
        if self.precision not in ("fp64", "fp32"):
            raise ValueError("solver.precision must be 'fp64' or 'fp32'")
        if state.num_eqn != pol.num_eqn:
            raise ValueError("state.num_eqn does not match the Riemann solver")
        for key in pol.problem_data_keys:
            if key not in state.problem_data:
                raise KeyError("state.problem_data[%r] must be set for %s" % (key, pol.name))

    def _policy(self):
        """The HIP solver for solver.rp, chosen by the module's name as pyclaw's riemann.static
        lookup does: this package's riemann modules and upstream's clawpack.riemann modules of the
        same name select the same kernel."""
        from ..riemann import static
        pol = None if self.rp is None else static.policy(self.rp)
        if pol is None:
            raise NotImplementedError(
                "Riemann solver %r not available on the HIP path: use shallow_roe_with_efix_2D, "
                "euler_4wave_2D or shallow_bathymetry_fwave_2D (from clawpack_amd.riemann or "
                "clawpack.riemann)" % (static.solver_name(self.rp) if self.rp is not None else None,))
        return pol

    def _make_desc(self, state):
        mx, my = state.grid.num_cells
        dx, dy = state.grid.delta
        pd = state.problem_data
        pol = self._policy()
        return make_desc(
            pol.kernel_id, _lib.FP64 if self.precision == "fp64" else _lib.FP32, mx, my, dx, dy,
            pol.num_eqn, pol.num_waves, order=self.order,
            transverse_waves=self.transverse_waves, mthlim=self._mthlim(),
            # a custom side's ghost cells come from the user's function through the padded step,
            # which does not read the state's BC codes; the descriptor carries extrap there
            bc_lower=[BC.extrap if b == BC.custom else b for b in self.bc_lower],
            bc_upper=[BC.extrap if b == BC.custom else b for b in self.bc_upper],
            reflect_index=self.reflect_index,
            grav=pd.get("grav", 1.0), gamma=pd.get("gamma", 1.4), cfl_max=self.cfl_max,
            cfl_desired=self.cfl_desired, dt_max=self.dt_max, dt_variable=self.dt_variable,
            seg_h=self.seg_h, num_aux=state.num_aux, fwave=bool(self.fwave),
            aux_bc_lower=[b if b is not None else 1 for b in self.aux_bc_lower],
            aux_bc_upper=[b if b is not None else 1 for b in self.aux_bc_upper],
            dry_tolerance=pd.get("dry_tolerance", 0.0), index_capa=state.index_capa,
            depth_ref=self._depth_ref(pol))

    def _depth_ref(self, pol):
        if self.precision != "fp32" or pol.kernel_id != _lib.SOLVER_SHALLOW_ROE_EFIX:
            return 0.0
        if self.step_source is not None or self.before_step is not None or self._has_custom_bc():
            return 0.0  # user functions see the device planes, which then hold h itself
        return float(self._auto_depth_ref if self.depth_ref is None else self.depth_ref)

    def setup(self, solution):
        state = solution.state
        self._check_supported(state)
        self._check_cfl_settings()
        self._auto_depth_ref = float(np.float32(np.mean(state.q[0]))) if state.q.size else 0.0
        if self.dt is None:
            self.dt = self.dt_initial
        self._device(state)
        self._is_set_up = True

    def _device(self, state):
        desc = self._make_desc(state)
        if self.autotune and self.seg_h == 0:
            from .. import autotune
            autotune.apply_shape(desc, autotune.tuned_shape(desc, split=bool(self.dimensional_split)))
        key = (desc.mx, desc.my, desc.num_eqn, desc.dtype, desc.num_aux)
        if self._grid is None or self._desc_key != key:
            self._grid = DeviceGrid(desc)
            self._desc_key = key
        else:
            self._grid.desc = desc
        self._grid.split = bool(self.dimensional_split)
        if state.num_aux > 0:
            self._grid.upload_aux(state.aux)  # aux is read-only on the device; refreshed per call
        return self._grid

    # ----------------------------------------------------------------------------- steps ---
    def _has_custom_bc(self):
        return BC.custom in list(self.bc_lower) + list(self.bc_upper)

    def _padded(self, g):
        """Device buffers of the padded path: qbc (num_eqn, my+4, mx+4), with dimensional_split a
        second one for q*, and, with aux, auxbc (num_aux, my+4, mx+4) filled once from the aux
        planes with the aux BCs."""
        d = g.desc
        qbc = torch.empty((d.num_eqn, d.my + 4, d.mx + 4), dtype=g.torch_dtype, device=g.device)
        qsbc = torch.empty_like(qbc) if self.dimensional_split else None
        auxbc = None
        if g.aux is not None:
            auxbc = torch.empty((d.num_aux, d.my + 4, d.mx + 4), dtype=g.torch_dtype, device=g.device)
            auxbc[:, 2:-2, 2:-2].copy_(g.aux)
            v = auxbc.permute(0, 2, 1)
            for k in range(2):
                _fill_side(v, k + 1, 0, self.aux_bc_lower[k], 2, -1)
                _fill_side(v, k + 1, 1, self.aux_bc_upper[k], 2, -1)
        return qbc, qsbc, auxbc

    def _fill_qbc(self, qbc, auxbc, src, ds, t):
        """pyclaw's _apply_bcs on the device: interior from the planes src (None: qbc's interior
        already holds the state), then per dimension (x, then y) the lower and the upper side,
        standard rules by slicing and BC.custom sides by the user's function(state, dim, t, qbc,
        auxbc, num_ghost) on pyclaw-layout torch views."""
        if src is None:
            src = qbc[:, 2:-2, 2:-2]
        else:
            qbc[:, 2:-2, 2:-2].copy_(src)
        v = qbc.permute(0, 2, 1)
        av = None if auxbc is None else auxbc.permute(0, 2, 1)
        ds.bind(src, t)
        for k, dim in enumerate(ds.grid.dimensions):
            for side, bc, fn in ((0, self.bc_lower[k], self.user_bc_lower), (1, self.bc_upper[k], self.user_bc_upper)):
                if bc == BC.custom:
                    fn(ds, dim, t, v, av, self.num_ghost)
                else:
                    _fill_side(v, k + 1, side, bc, 2, self.reflect_index[k])

    def _custom_step(self, g, bufs, src, ds, t, dt):
        """The hyperbolic step with ghost cells from _apply_bcs (user functions included), into the
        other state buffer: step2 from the filled qbc, or with dimensional_split the x-sweeps into
        q*, the same ghost fill of q* (same rules, same t) and the y-sweeps."""
        qbc, qsbc, auxbc = bufs
        self._fill_qbc(qbc, auxbc, src, ds, t)
        if not self.dimensional_split:
            g.step_padded(qbc, g.q[1 - g.cur], dt)
            return
        g.sweep_padded(qbc, qsbc, dt, 1)
        self._fill_qbc(qsbc, auxbc, None, ds, t)
        g.sweep_padded(qsbc, g.q[1 - g.cur], dt, 2)

    def step_hyperbolic(self, solution):
        """One step (step2, or step2ds when dimensional_split) with the current dt, no
        accept/reject, as ClawSolver2D.step_hyperbolic."""
        if not self._is_set_up:
            self.setup(solution)
        state = solution.state
        self._check_supported(state)
        g = self._device(state)
        g.upload(state.q)
        g.init_ctrl(solution.t, self.dt, target_steps=1)
        if self._has_custom_bc():
            self._custom_step(g, self._padded(g), g.q[g.cur], DeviceState(state, g), solution.t, self.dt)
            g.cur = 1 - g.cur
            c = g.read_ctrl()
            cfl = c.cfl
        else:
            cfl, c = g.step(self.dt)
        state.q[...] = g.download()
        self.cfl.update_global_max(cfl)
        if c.status & _lib.STATUS_BAD_STATE:
            raise FloatingPointError("HIP step produced a non-positive depth/density or NaN")
        return cfl

    def evolve_to_time(self, solution, tend=None):
        """Advance solution to tend (or by one step attempt if tend is None), with pyclaw's
        accept/reject and CFL-driven dt, on the device."""
        if not self._is_set_up:
            self.setup(solution)
        state = solution.state
        self._check_supported(state)
        self._check_cfl_settings()
        if self.step_source is not None or self.before_step is not None or self._has_custom_bc():
            return self._evolve_with_callbacks(solution, tend)
        g = self._device(state)
        tstart = solution.t
        take_one_step = tend is None
        self.status["cflmax"] = self.cfl.get_cached_max()
        self.status["dtmin"] = self.dt
        self.status["dtmax"] = self.dt
        self.status["numsteps"] = 0
        g.upload(state.q)
        if take_one_step:
            g.init_ctrl(tstart, self.dt, target_steps=1, max_steps=1)
        else:
            g.init_ctrl(tstart, self.dt, tend=tend, max_steps=self.max_steps)
        c = g.read_ctrl()
        attempts_left = self.max_steps if not take_one_step else 1
        while not c.done and attempts_left > 0:
            if take_one_step:
                chunk = 1
            else:
                est = (tend - c.t) / c.dt if c.dt > 0 else 1.0
                chunk = int(min(max(math.ceil(est) + 1, 1), 1024, attempts_left))
            g.advance(chunk)
            attempts_left -= chunk
            c = g.read_ctrl()
        g.sync_cur(c)
        state.q[...] = g.download()
        solution.t = c.t
        self.cfl.update_global_max(c.cfl)
        self.status["cflmax"] = max(self.status["cflmax"], c.cflmax)
        self.status["dtmin"] = min(self.status["dtmin"], c.dtmin)
        self.status["dtmax"] = max(self.status["dtmax"], c.dtmax)
        self.status["numsteps"] += int(c.numsteps)
        self.status["rejected"] = int(c.rejected)
        if c.status & _lib.STATUS_BAD_STATE:
            raise FloatingPointError("HIP step produced a non-positive depth/density or a NaN "
                                     "wave speed at t=%g" % c.t)
        if c.status & _lib.STATUS_CFL_FIXED_DT:
            raise Exception("CFL too large, giving up!")
        self.dt = c.dt
        if c.status & _lib.STATUS_MAX_STEPS:
            raise Exception("Maximum number of timesteps have been taken")
        return self.status

    def _evolve_with_callbacks(self, solution, tend):
        """evolve_to_time with before_step / step_source / BC.custom (upstream Solver.step and
        evolve_to_time): per attempt before_step, for Strang splitting a half source step, the
        ghost fill (with the user's BC functions) when a side is custom, the hyperbolic step, and
        after an accepted step the source over dt (Godunov) or dt/2 (Strang).  The accept/reject/dt rules are those of the device controller, applied on the
        host after each launch.  A rejected attempt leaves q^n untouched (the Strang half step
        works on a copy)."""
        state = solution.state
        g = self._device(state)
        g.upload(state.q)
        ds = DeviceState(state, g)
        take_one_step = tend is None
        strang = self.step_source is not None and self.source_split == 2
        tmp = torch.empty_like(g.q[0]) if strang else None
        custom = self._has_custom_bc()
        bufs = self._padded(g) if custom else None
        t = tstart = solution.t
        dt = self.dt
        if not take_one_step:
            if t + dt > tend and tstart < tend:
                dt = tend - t
            if tend - t - dt < 1.0e-14 * t:
                dt = tend - t
        g.init_ctrl(t, max(dt, 1e-300))  # zeroes the speed accumulators and status bits
        cflmax = self.cfl.get_cached_max()
        dtmin = dtmax = dt
        numsteps = rejected = 0
        done = False
        cfl = 0.0
        for attempt in range(1 if take_one_step else self.max_steps):
            if self.before_step is not None:
                self.before_step(self, ds.bind(g.q[g.cur], t))
            src = g.q[g.cur]
            if strang:
                tmp.copy_(src)
                self.step_source(self, ds.bind(tmp, t), dt / 2.0)
                src = tmp
            if custom:
                self._custom_step(g, bufs, src, ds, t, dt)
            else:
                g.step_into(src, g.q[1 - g.cur], dt)
            c = g.read_ctrl()
            cfl = c.cfl
            bad = bool(c.status & _lib.STATUS_BAD_STATE) or cfl != cfl
            if bad and not (cfl > self.cfl_max):
                solution.t = t
                raise FloatingPointError("HIP step produced a non-positive depth/density or a NaN "
                                         "wave speed at t=%g" % t)
            if c.status:
                c.status = 0
                g.write_ctrl(c)
            if not (cfl > self.cfl_max):
                cflmax = max(cfl, cflmax)
                numsteps += 1
                g.cur = 1 - g.cur
                if self.step_source is not None:
                    # pyclaw's Solver.step applies the source before evolve_to_time advances
                    # solution.t, so the source of an accepted step sees the step's start time
                    self.step_source(self, ds.bind(g.q[g.cur], t), dt / 2.0 if strang else dt)
                t = t + dt if self.dt_variable else tstart + numsteps * dt
            else:
                rejected += 1
                if not self.dt_variable:
                    cflmax = max(cfl, cflmax)
                    self.cfl.update_global_max(cfl)
                    raise Exception("CFL too large, giving up!")
            if take_one_step or t >= tend:
                done = True
                break
            if self.dt_variable:
                if cfl > 0.0:
                    dt = min(self.dt_max, dt * self.cfl_desired / cfl)
                    dtmin = min(dt, dtmin)
                    dtmax = max(dt, dtmax)
                else:
                    dt = self.dt_max
            if t + dt > tend and tstart < tend:
                dt = tend - t
            if tend - t - dt < 1.0e-14 * t:
                dt = tend - t
        state.q[...] = g.download()
        solution.t = t
        self.cfl.update_global_max(cfl)
        self.status["cflmax"] = cflmax
        self.status["dtmin"] = dtmin
        self.status["dtmax"] = dtmax
        self.status["numsteps"] = numsteps
        self.status["rejected"] = rejected
        self.dt = dt  # the dt of the last attempt (pyclaw leaves the loop before choosing a new one)
        # pyclaw raises only when every one of the max_steps attempts was accepted and tend is not
        # reached; attempts rejected on the way end the loop silently short of tend
        if not done and self.dt_variable and t < tend and numsteps == self.max_steps:
            raise Exception("Maximum number of timesteps have been taken")
        return self.status

    # ------------------------------------------------------------------------- extras ------
    def total(self, solution, m=0):
        """Compensated FP64 sum of q[m] over the grid, computed on the device in a fixed order
        (for FP32 states this is the conservation check of the FP32 path)."""
        if not self._is_set_up:
            self.setup(solution)
        g = self._device(solution.state)
        g.upload(solution.state.q)
        return g.mass_sum(m)
