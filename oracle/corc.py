"""ctypes bindings of ``oracle/claw2d_oracle.c`` (test infrastructure only).

Arrays follow pyclaw's conventions: ``q`` has shape ``(num_eqn, mx, my)`` and is passed to C in
Fortran order (component fastest), exactly like ``state.q`` handed to the f2py ``step2``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc2d.so")

SWE, EULER, SWE_BATHY = 0, 1, 2
BC_CUSTOM, BC_EXTRAP, BC_PERIODIC, BC_WALL = 0, 1, 2, 3


class Desc(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("num_eqn", C.c_int), ("num_waves", C.c_int),
        ("num_ghost", C.c_int), ("mx", C.c_int), ("my", C.c_int), ("order", C.c_int),
        ("transverse_waves", C.c_int), ("mthlim", C.c_int * 4), ("dx", C.c_double),
        ("dy", C.c_double), ("grav", C.c_double), ("gamma", C.c_double),
        ("num_aux", C.c_int), ("fwave", C.c_int), ("dry_tolerance", C.c_double),
        ("mcapa", C.c_int),
    ]


class Ctrl(C.Structure):
    _fields_ = [
        ("dt_initial", C.c_double), ("dt_max", C.c_double), ("cfl_max", C.c_double),
        ("cfl_desired", C.c_double), ("dt_variable", C.c_int), ("max_steps", C.c_int),
        ("bc_lower", C.c_int * 2), ("bc_upper", C.c_int * 2), ("reflect_index", C.c_int * 2),
        ("dimensional_split", C.c_int), ("aux_bc_lower", C.c_int * 2), ("aux_bc_upper", C.c_int * 2),
    ]


class Status(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt", C.c_double), ("cflmax", C.c_double), ("dtmin", C.c_double),
        ("dtmax", C.c_double), ("last_cfl", C.c_double), ("numsteps", C.c_long),
        ("rejected", C.c_long), ("attempts", C.c_long), ("error", C.c_int),
    ]


_lib = None


def build(force=True):
    """Compile liborc2d.so from claw2d_oracle.c with gcc (no reference sources involved); with
    force (the default) it recompiles even when the library is newer than the source."""
    subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build(force=False)
        L = C.CDLL(_LIB)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
        dpa = C.POINTER(C.c_double)
        L.orc_step2.argtypes = [C.POINTER(Desc), dpa, dpa, C.c_double, dpa]
        L.orc_step2_aux.argtypes = [C.POINTER(Desc), dpa, dpa, dpa, C.c_double, dpa]
        L.orc_step2ds_aux.argtypes = [C.POINTER(Desc), dpa, dpa, dpa, C.c_double, C.c_int, dpa]
        L.orc_fill_ghosts.argtypes = [C.POINTER(Desc), C.POINTER(Ctrl), dpa, dpa]
        L.orc_fill_ghosts.restype = None
        L.orc_fill_aux_ghosts.argtypes = [C.POINTER(Desc), C.POINTER(Ctrl), dpa, dpa]
        L.orc_fill_aux_ghosts.restype = None
        L.orc_step_hyperbolic.argtypes = [C.POINTER(Desc), C.POINTER(Ctrl), dpa, C.c_double, dpa]
        L.orc_step_hyperbolic_aux.argtypes = [C.POINTER(Desc), C.POINTER(Ctrl), dpa, dpa, C.c_double,
                                              dpa]
        L.orc_evolve.argtypes = [C.POINTER(Desc), C.POINTER(Ctrl), dpa, C.c_double, C.c_long,
                                 C.POINTER(Status)]
        L.orc_evolve_aux.argtypes = [C.POINTER(Desc), C.POINTER(Ctrl), dpa, dpa, C.c_double,
                                     C.c_long, C.POINTER(Status)]
        L.orc_rpn2.argtypes = [C.POINTER(Desc), C.c_int, C.c_int, dpa, dpa, dpa, dpa, dpa]
        L.orc_rpn2_aux.argtypes = [C.POINTER(Desc), C.c_int, C.c_int, dpa, dpa, dpa, dpa, dpa, dpa]
        L.orc_flux2_aux.argtypes = [C.POINTER(Desc), C.c_int, C.c_int, C.c_double, dpa, dpa, dpa,
                                    dpa, dpa, dpa]
        L.orc_rpt2.argtypes = [C.POINTER(Desc), C.c_int, C.c_int, dpa, dpa, dpa, dpa]
        L.orc_limiter.argtypes = [C.POINTER(Desc), C.c_int, dpa, dpa]
        L.orc_flux2.argtypes = [C.POINTER(Desc), C.c_int, C.c_int, C.c_double, dpa, dpa, dpa,
                                dpa, dpa]
        L.orc_philim.argtypes = [C.c_double, C.c_double, C.c_int]
        L.orc_philim.restype = C.c_double
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = None
        L.orc_component_sum.argtypes = [C.POINTER(Desc), dpa, C.c_int]
        L.orc_component_sum.restype = C.c_double
        for f in ("orc_step2", "orc_step2_aux", "orc_step2ds_aux", "orc_step_hyperbolic",
                  "orc_step_hyperbolic_aux", "orc_evolve", "orc_evolve_aux", "orc_rpn2",
                  "orc_rpn2_aux", "orc_rpt2", "orc_limiter", "orc_flux2", "orc_flux2_aux"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def set_threads(n):
    """Threads for the step (1 = sequential slice order, the default; n > 1 only for timing a
    multi-core CPU baseline: sums then associate differently at round-off level)."""
    lib().orc_set_threads(int(n))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _solver_id(solver):
    if solver in (SWE, "swe", "shallow_roe_with_efix_2D"):
        return SWE
    if solver in (SWE_BATHY, "bathy", "shallow_bathymetry_fwave_2D"):
        return SWE_BATHY
    if solver in (EULER, "euler", "euler_4wave_2D"):
        return EULER
    raise ValueError("unknown solver %r" % (solver,))


def make_desc(solver, mx, my, dx, dy, grav=1.0, gamma=1.4, order=2, transverse_waves=2,
              mthlim=(4, 4, 4, 4), num_ghost=2, num_aux=None, dry_tolerance=1e-3, fwave=None,
              index_capa=-1):
    """fwave and num_aux default to what the solver needs (f-waves and one aux field, the
    bathymetry, for shallow_bathymetry_fwave_2D).  index_capa >= 0: aux[index_capa] is the
    capacity function kappa (pyclaw's state.index_capa; the step divides each cell's flux
    differences by its kappa)."""
    d = Desc()
    d.solver = _solver_id(solver)
    d.num_eqn = 4 if d.solver == EULER else 3
    d.num_waves = d.num_eqn
    d.num_aux = int((1 if d.solver == SWE_BATHY else 0) if num_aux is None else num_aux)
    d.fwave = int(bool(d.solver == SWE_BATHY if fwave is None else fwave))
    d.dry_tolerance = float(dry_tolerance)
    d.num_ghost = num_ghost
    d.mx, d.my, d.order, d.transverse_waves = mx, my, order, transverse_waves
    lim = list(mthlim) + [mthlim[-1]] * (4 - len(mthlim))
    for k in range(4):
        d.mthlim[k] = int(lim[k])
    d.dx, d.dy, d.grav, d.gamma = dx, dy, grav, gamma
    if index_capa >= d.num_aux:
        raise ValueError("index_capa %d needs num_aux > %d" % (index_capa, index_capa))
    d.mcapa = int(index_capa) + 1
    return d


def make_ctrl(bc_lower=(1, 1), bc_upper=(1, 1), reflect_index=(1, 2), dt_initial=0.1,
              dt_max=1e99, cfl_max=1.0, cfl_desired=0.9, dt_variable=True, max_steps=10000,
              dimensional_split=False, aux_bc_lower=None, aux_bc_upper=None):
    """aux_bc_lower/upper default to bc_lower/bc_upper."""
    c = Ctrl()
    c.dt_initial, c.dt_max, c.cfl_max, c.cfl_desired = dt_initial, dt_max, cfl_max, cfl_desired
    c.dt_variable, c.max_steps = int(bool(dt_variable)), int(max_steps)
    c.dimensional_split = int(bool(dimensional_split))
    for k in range(2):
        c.bc_lower[k], c.bc_upper[k], c.reflect_index[k] = bc_lower[k], bc_upper[k], reflect_index[k]
        c.aux_bc_lower[k] = (bc_lower if aux_bc_lower is None else aux_bc_lower)[k]
        c.aux_bc_upper[k] = (bc_upper if aux_bc_upper is None else aux_bc_upper)[k]
    return c


def _aux_arg(d, aux):
    if d.num_aux == 0:
        return None
    if aux is None:
        raise ValueError("this solver needs aux (num_aux = %d)" % d.num_aux)
    return np.asfortranarray(aux, dtype=np.float64)


def fill_ghosts(d, c, q):
    q = np.asfortranarray(q, dtype=np.float64)
    g = d.num_ghost
    qbc = np.zeros((d.num_eqn, d.mx + 2 * g, d.my + 2 * g), order="F")
    lib().orc_fill_ghosts(C.byref(d), C.byref(c), _p(q), _p(qbc))
    return qbc


def fill_aux_ghosts(d, c, aux):
    aux = np.asfortranarray(aux, dtype=np.float64)
    g = d.num_ghost
    auxbc = np.zeros((d.num_aux, d.mx + 2 * g, d.my + 2 * g), order="F")
    lib().orc_fill_aux_ghosts(C.byref(d), C.byref(c), _p(aux), _p(auxbc))
    return auxbc


def step2(d, qbc, dt, auxbc=None):
    """One f2py-style step2 call on a ghost-filled qbc (and auxbc) -> (qnew_bc, cfl)."""
    qold = np.asfortranarray(qbc, dtype=np.float64)
    qnew = np.zeros_like(qold, order="F")
    cfl = C.c_double(0.0)
    rc = lib().orc_step2_aux(C.byref(d), _p(qold), _p(_aux_arg(d, auxbc)), _p(qnew), dt,
                             C.byref(cfl))
    assert rc == 0
    return qnew, cfl.value


def step2ds(d, qbc, dt, ids, auxbc=None):
    """One f2py-style step2ds call (ids 1: x-sweeps, 2: y-sweeps) -> (qnew_bc, cfl)."""
    qold = np.asfortranarray(qbc, dtype=np.float64)
    qnew = np.zeros_like(qold, order="F")
    cfl = C.c_double(0.0)
    rc = lib().orc_step2ds_aux(C.byref(d), _p(qold), _p(_aux_arg(d, auxbc)), _p(qnew), dt,
                               int(ids), C.byref(cfl))
    assert rc == 0
    return qnew, cfl.value


def step_hyperbolic(d, c, q, dt, aux=None):
    """ClawSolver2D.step_hyperbolic on interior q -> (qnew, cfl); q is not modified."""
    qq = np.array(q, dtype=np.float64, order="F", copy=True)
    cfl = C.c_double(0.0)
    rc = lib().orc_step_hyperbolic_aux(C.byref(d), C.byref(c), _p(qq), _p(_aux_arg(d, aux)), dt,
                                       C.byref(cfl))
    assert rc == 0
    return qq, cfl.value


def evolve(d, c, q, nsteps=None, tend=None, t0=0.0, dt=None, aux=None):
    """Solver.evolve_to_time semantics.  nsteps: stop after that many accepted steps."""
    qq = np.array(q, dtype=np.float64, order="F", copy=True)
    st = Status()
    st.t = t0
    st.dt = c.dt_initial if dt is None else dt
    st.cflmax = 0.0
    rc = lib().orc_evolve_aux(C.byref(d), C.byref(c), _p(qq), _p(_aux_arg(d, aux)),
                              float("inf") if tend is None else tend, int(nsteps or 0),
                              C.byref(st))
    assert rc == 0
    status = {k: getattr(st, k) for k, _ in Status._fields_}
    return qq, status


def slice_arrays(d, n):
    g = d.num_ghost
    L = n + 2 * g
    return L, g


def rpn2(d, ixy, q1d, aux1d=None):
    """q1d: (num_eqn, n+2g) [aux1d: (num_aux, n+2g)] -> wave (meqn, mwaves, L), s (mwaves, L),
    amdq, apdq (meqn, L)."""
    q1d = np.asfortranarray(q1d, dtype=np.float64)
    aux1d = _aux_arg(d, aux1d)
    L = q1d.shape[1]
    n = L - 2 * d.num_ghost
    wave = np.zeros((d.num_eqn, d.num_waves, L), order="F")
    s = np.zeros((d.num_waves, L), order="F")
    amdq = np.zeros((d.num_eqn, L), order="F")
    apdq = np.zeros((d.num_eqn, L), order="F")
    lib().orc_rpn2_aux(C.byref(d), ixy, n, _p(q1d), _p(aux1d), _p(wave), _p(s), _p(amdq), _p(apdq))
    return wave, s, amdq, apdq


def rpt2(d, ixy, q1d, asdq):
    q1d = np.asfortranarray(q1d, dtype=np.float64)
    asdq = np.asfortranarray(asdq, dtype=np.float64)
    L = q1d.shape[1]
    n = L - 2 * d.num_ghost
    bm = np.zeros((d.num_eqn, L), order="F")
    bp = np.zeros((d.num_eqn, L), order="F")
    lib().orc_rpt2(C.byref(d), ixy, n, _p(q1d), _p(asdq), _p(bm), _p(bp))
    return bm, bp


def limiter(d, wave, s):
    wave = np.array(wave, dtype=np.float64, order="F", copy=True)
    s = np.asfortranarray(s, dtype=np.float64)
    n = wave.shape[2] - 2 * d.num_ghost
    lib().orc_limiter(C.byref(d), n, _p(wave), _p(s))
    return wave


def flux2(d, ixy, q1d, dtdx, aux1d=None):
    q1d = np.asfortranarray(q1d, dtype=np.float64)
    aux1d = _aux_arg(d, aux1d)
    L = q1d.shape[1]
    n = L - 2 * d.num_ghost
    qadd = np.zeros((d.num_eqn, L), order="F")
    fadd = np.zeros((d.num_eqn, L), order="F")
    gadd = np.zeros((d.num_eqn, 2, L), order="F")
    cfl = C.c_double(0.0)
    rc = lib().orc_flux2_aux(C.byref(d), ixy, n, dtdx, _p(q1d), _p(aux1d), _p(qadd), _p(fadd),
                             _p(gadd), C.byref(cfl))
    assert rc == 0
    return qadd, fadd, gadd, cfl.value


def philim(a, b, meth):
    return lib().orc_philim(a, b, meth)


def component_sum(d, q, m):
    q = np.asfortranarray(q, dtype=np.float64)
    return lib().orc_component_sum(C.byref(d), _p(q), m)


def apply_bcs(d, c, q, t=0.0, user_bc_lower=None, user_bc_upper=None, auxbc=None):
    """Solver._apply_bcs (upstream pyclaw/src/pyclaw/solver.py, absent) in numpy: q copied into
    qbc (num_eqn, mx+2g, my+2g), then for x and then y the lower and the upper side, each over full
    slices of the other dimension: extrap copies the edge cell, periodic wraps, wall mirrors and
    negates the reflect_index component, custom (code 0) calls the user function
    f(None, idim, t, qbc, auxbc, num_ghost), which fills that side of qbc in place."""
    g = d.num_ghost
    qbc = np.zeros((d.num_eqn, d.mx + 2 * g, d.my + 2 * g), order="F")
    qbc[:, g:-g, g:-g] = q
    for idim in range(2):
        a = np.moveaxis(qbc, idim + 1, 1)  # a view, with this dimension as axis 1
        ri = c.reflect_index[idim]
        for side, bc in ((0, c.bc_lower[idim]), (1, c.bc_upper[idim])):
            if bc == BC_CUSTOM:
                (user_bc_lower if side == 0 else user_bc_upper)(None, idim, t, qbc, auxbc, g)
            elif side == 0:
                if bc == BC_EXTRAP:
                    for i in range(g):
                        a[:, i] = a[:, g]
                elif bc == BC_PERIODIC:
                    a[:, :g] = a[:, -2 * g:-g]
                elif bc == BC_WALL:
                    for i in range(g):
                        a[:, i] = a[:, 2 * g - 1 - i]
                        a[ri, i] = -a[ri, 2 * g - 1 - i]
            else:
                if bc == BC_EXTRAP:
                    for i in range(g):
                        a[:, -i - 1] = a[:, -g - 1]
                elif bc == BC_PERIODIC:
                    a[:, -g:] = a[:, g:2 * g]
                elif bc == BC_WALL:
                    for i in range(g):
                        a[:, -i - 1] = a[:, -2 * g + i]
                        a[ri, -i - 1] = -a[ri, -2 * g + i]
    return qbc


def evolve_with_source(d, c, q, source=None, source_split=1, before_step=None, nsteps=None,
                       tend=None, t0=0.0, dt=None, aux=None, user_bc_lower=None, user_bc_upper=None,
                       source_t=False):
    """Solver.evolve_to_time with Solver.step's source splitting (upstream pyclaw/src/pyclaw/
    solver.py, absent from the reference tree): per attempt before_step(q, t), for Strang
    splitting (source_split 2) a half step source(q, dt/2) before the hyperbolic step, and after
    an accepted step source(q, dt) (Godunov, source_split 1) or source(q, dt/2) (Strang); a
    rejected attempt restores q and t.  source and before_step act on q (num_eqn, mx, my) in
    place.  With custom BC codes (0) in c the ghost cells come from apply_bcs with the user
    functions and the step is orc_step2 on that qbc, or with dimensional splitting the x-sweeps of
    orc_step2ds, apply_bcs on q* and the y-sweeps.  The accept/reject, t and dt rules are orc_evolve's; nsteps stops after that many
    accepted steps (no clipping to tend), tend runs to that time.  The hyperbolic step is
    orc_step_hyperbolic.  With source_t the source is called as source(q, dt, t), t being the
    attempt's start time in both of its calls: Solver.step runs before evolve_to_time advances
    solution.t."""
    q = np.array(q, dtype=np.float64, order="F", copy=True)
    t = tstart = t0
    dt = c.dt_initial if dt is None else dt
    st = dict(numsteps=0, rejected=0, attempts=0, cflmax=0.0, dtmin=dt, dtmax=dt, error=0)
    to_tend = nsteps is None
    for _ in range(c.max_steps):
        if to_tend:
            if t + dt > tend and tstart < tend:
                dt = tend - t
            if tend - t - dt < 1.0e-14 * t:
                dt = tend - t
        qbackup = q.copy(order="F")
        if before_step is not None:
            before_step(q, t)
        src = source
        if source is not None and source_t:
            src = lambda qq, h, t_=t: source(qq, h, t_)  # noqa: E731
        if source is not None and source_split == 2:
            src(q, dt / 2.0)
        if BC_CUSTOM in tuple(c.bc_lower) + tuple(c.bc_upper):
            g = d.num_ghost
            auxbc = None if aux is None else fill_aux_ghosts(d, c, aux)
            qbc = apply_bcs(d, c, q, t, user_bc_lower, user_bc_upper, auxbc=auxbc)
            if c.dimensional_split:
                # orc_step_hyperbolic_aux's order: x-sweeps, the ghost fill of q*, y-sweeps
                qsbc, cx = step2ds(d, qbc, dt, 1, auxbc=auxbc)
                qsbc = apply_bcs(d, c, qsbc[:, g:-g, g:-g], t, user_bc_lower, user_bc_upper, auxbc=auxbc)
                qnbc, cy = step2ds(d, qsbc, dt, 2, auxbc=auxbc)
                cfl = max(cx, cy)
            else:
                qnbc, cfl = step2(d, qbc, dt, auxbc=auxbc)
            qn = np.asfortranarray(qnbc[:, g:-g, g:-g])
        else:
            qn, cfl = step_hyperbolic(d, c, q, dt, aux=aux)
        st["attempts"] += 1
        if not (cfl > c.cfl_max):
            q = qn
            st["cflmax"] = max(cfl, st["cflmax"])
            st["numsteps"] += 1
            if source is not None:
                src(q, dt / 2.0 if source_split == 2 else dt)
            t = t + dt if c.dt_variable else tstart + st["numsteps"] * dt
        else:
            st["rejected"] += 1
            q = qbackup
            if not c.dt_variable:
                st["error"] = 2
                break
        if (st["numsteps"] >= nsteps) if not to_tend else (t >= tend):
            break
        if c.dt_variable:
            if cfl > 0.0:
                dt = min(c.dt_max, dt * c.cfl_desired / cfl)
                st["dtmin"] = min(dt, st["dtmin"])
                st["dtmax"] = max(dt, st["dtmax"])
            else:
                dt = c.dt_max
    else:
        # as orc_evolve: short of tend only every attempt accepted is an error (Solver.evolve_to_time)
        if not to_tend or (c.dt_variable and t < tend and st["numsteps"] == c.max_steps):
            st["error"] = 1
    st["t"], st["dt"] = t, dt
    return q, st
