"""Capacity-form steps (state.index_capa: aux[index_capa] = kappa) on the GPU against the oracle's
capacity form (oracle/claw2d_oracle.c, pinned in tests/test_oracle_capa.py): single unsplit and
split steps for every Riemann solver over BCs (aux BCs different from the state's included),
method variants, the device time loop (the CFL number with kappa chooses dt), FP32, the PyClaw
front end, and kappa = 1 equal bit for bit to the same kernel family without a capacity function.
Tolerances: FP64 1e-12 max|q| per step (1e-11 over 100 steps), FP32 2e-6 per step."""
import numpy as np
import pytest

from clawpack_amd import _lib, problems, pyclaw, riemann
from clawpack_amd.device import DeviceGrid, make_desc
from oracle import corc

pytestmark = pytest.mark.gpu

G, GAMMA, TOL = 1.0, 1.4, 1e-3
SOLVERS = {"swe": (_lib.SOLVER_SHALLOW_ROE_EFIX, 3), "euler": (_lib.SOLVER_EULER_4WAVE, 4),
           "bathy": (_lib.SOLVER_SHALLOW_BATHY_FWAVE, 3)}


def kappa_field(mx, my, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(mx) + 0.5) / mx
    y = (np.arange(my) + 0.5) / my
    X, Y = np.meshgrid(x, y, indexing="ij")
    return 1.0 + 0.4 * np.sin(2 * np.pi * X + 1.0) * np.cos(2 * np.pi * Y) + 0.2 * rng.random((mx, my))


def problem(solver, mx, my, seed):
    """(q, aux, index_capa): aux holds kappa (and the bathymetry at aux[0] for the f-wave solver,
    kappa then at aux[1])."""
    kap = kappa_field(mx, my, seed + 1)
    if solver == "bathy":
        q, b = problems.random_bathymetry(mx, my, seed=seed)
        aux = np.asfortranarray(np.stack([b[0], kap]))
        return q, aux, 1
    q = (problems.random_smooth_swe(mx, my, seed, amp=0.2) if solver == "swe"
         else problems.random_euler(mx, my, seed))
    return q, np.asfortranarray(kap[None]), 0


def oracle_desc(solver, mx, my, dx, dy, index_capa, **kw):
    return corc.make_desc(solver, mx, my, dx, dy, grav=9.81 if solver == "bathy" else G, gamma=GAMMA,
                          num_aux=index_capa + 1, index_capa=index_capa, dry_tolerance=TOL, **kw)


def device(solver, q, aux, dx, dy, index_capa, bc, abc=None, split=False, dtype=_lib.FP64, **kw):
    sid, neq = SOLVERS[solver]
    mx, my = q.shape[1:]
    d = make_desc(sid, dtype, mx, my, dx, dy, neq, neq, bc_lower=bc[0], bc_upper=bc[1],
                  aux_bc_lower=None if abc is None else abc[0],
                  aux_bc_upper=None if abc is None else abc[1],
                  grav=9.81 if solver == "bathy" else G, gamma=GAMMA, dry_tolerance=TOL,
                  num_aux=aux.shape[0], index_capa=index_capa, **kw)
    g = DeviceGrid(d, split=split)
    g.upload(q)
    g.upload_aux(aux)
    return g


def speed_bound(solver, q):
    h = q[0]
    if solver == "euler":
        return 4.0
    return float(np.sqrt((9.81 if solver == "bathy" else G) * h.max()) + 2.0)


BCS = [(((1, 1), (1, 1)), None), (((2, 2), (2, 2)), None), (((3, 3), (3, 3)), None),
       (((3, 1), (3, 1)), ((1, 1), (1, 1))), (((1, 2), (3, 2)), ((3, 2), (1, 2)))]


@pytest.mark.parametrize("solver", ["swe", "euler", "bathy"])
@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("bcs", BCS)
def test_single_step(solver, split, bcs):
    mx, my = 131, 70
    bc, abc = bcs
    print("case capa step", solver, split, bcs, flush=True)
    q, aux, ic = problem(solver, mx, my, seed=5)
    dx, dy = 1.0 / mx, 1.2 / my
    dt = 0.25 * min(dx, dy) / speed_bound(solver, q)
    d = oracle_desc(solver, mx, my, dx, dy, ic)
    c = corc.make_ctrl(bc_lower=bc[0], bc_upper=bc[1], aux_bc_lower=None if abc is None else abc[0],
                       aux_bc_upper=None if abc is None else abc[1], dimensional_split=split)
    qo, cflo = corc.step_hyperbolic(d, c, q, dt, aux=aux)
    g = device(solver, q, aux, dx, dy, ic, bc, abc, split=split)
    g.init_ctrl(0.0, dt)
    cfl, st = g.step(dt)
    assert st.status == 0
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    assert err <= 1e-12, err
    assert abs(cfl - cflo) <= 1e-12 * cflo, (cfl, cflo)


@pytest.mark.parametrize("solver", ["swe", "euler"])
@pytest.mark.parametrize("order,tw,lim", [(1, 0, (0, 0, 0, 0)), (1, 2, (0, 0, 0, 0)),
                                          (2, 1, (1, 1, 1, 1)), (2, 2, (2, 3, 4, 4)),
                                          (2, 0, (4, 4, 4, 4)), (2, 2, (5, 5, 5, 5))])
def test_method_variants(solver, order, tw, lim):
    mx, my = 90, 66
    print("case capa method", solver, order, tw, lim, flush=True)
    q, aux, ic = problem(solver, mx, my, seed=11)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.25 * min(dx, dy) / speed_bound(solver, q)
    bc = ((1, 2), (1, 2))
    d = oracle_desc(solver, mx, my, dx, dy, ic, order=order, transverse_waves=tw, mthlim=lim)
    c = corc.make_ctrl(bc_lower=bc[0], bc_upper=bc[1])
    qo, cflo = corc.step_hyperbolic(d, c, q, dt, aux=aux)
    g = device(solver, q, aux, dx, dy, ic, bc, order=order, transverse_waves=tw, mthlim=lim)
    g.init_ctrl(0.0, dt)
    cfl, st = g.step(dt)
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    assert err <= 1e-12, err
    assert abs(cfl - cflo) <= 1e-12 * cflo


@pytest.mark.parametrize("solver", ["swe", "euler", "bathy"])
@pytest.mark.parametrize("split", [False, True])
def test_device_time_loop(solver, split):
    """100 accepted steps with the device controller: the capacity CFL number sets every dt, and
    the first attempt (dt_initial = 0.1) is rejected as the oracle's."""
    mx, my = 96, 80
    print("case capa loop", solver, split, flush=True)
    q, aux, ic = problem(solver, mx, my, seed=21)
    dx, dy = 1.0 / mx, 1.0 / my
    bc = ((3, 1), (3, 1))
    d = oracle_desc(solver, mx, my, dx, dy, ic)
    c = corc.make_ctrl(bc_lower=bc[0], bc_upper=bc[1], dimensional_split=split)
    qo, st = corc.evolve(d, c, q, nsteps=100, aux=aux)
    g = device(solver, q, aux, dx, dy, ic, bc, split=split)
    g.init_ctrl(0.0, 0.1, target_steps=100)
    g.advance(100 + st["rejected"] + 1)
    cc = g.sync_cur()
    assert cc.numsteps == 100 and cc.rejected == st["rejected"], (cc.numsteps, cc.rejected, st)
    assert cc.rejected >= 1
    assert abs(cc.t - st["t"]) <= 1e-12 * st["t"]
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    assert err <= 1e-11, err


@pytest.mark.parametrize("solver", ["swe", "euler", "bathy"])
@pytest.mark.parametrize("split", [False, True])
def test_fp32(solver, split):
    mx, my = 120, 77
    print("case capa fp32", solver, split, flush=True)
    q, aux, ic = problem(solver, mx, my, seed=3)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.25 * min(dx, dy) / speed_bound(solver, q)
    bc = ((2, 2), (2, 2))
    d = oracle_desc(solver, mx, my, dx, dy, ic)
    c = corc.make_ctrl(bc_lower=bc[0], bc_upper=bc[1], dimensional_split=split)
    qo, _ = corc.step_hyperbolic(d, c, q, dt, aux=aux)
    g = device(solver, q, aux, dx, dy, ic, bc, split=split, dtype=_lib.FP32)
    g.init_ctrl(0.0, dt)
    g.step(dt)
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    assert err <= 2e-6, err


@pytest.mark.parametrize("solver", ["swe", "euler"])
@pytest.mark.parametrize("split", [False, True])
def test_unit_capacity_equals_the_plain_kernel_bitwise(solver, split):
    """kappa = 1 in the capacity kernel: every factor 1/kappa is exactly 1, the averaged dt/dx of
    the correction is exactly dt/dx, so the results equal the plain kernel of the same method (the
    general-limiter kernel, selected here by mixed per-wave limiters) bit for bit."""
    mx, my = 150, 90
    print("case capa unit", solver, split, flush=True)
    q, _, _ = problem(solver, mx, my, seed=8)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.25 * min(dx, dy) / speed_bound(solver, q)
    lim = (4, 1, 4, 4)
    sid, neq = SOLVERS[solver]
    bc = ((3, 1), (3, 1))
    plain = DeviceGrid(make_desc(sid, _lib.FP64, mx, my, dx, dy, neq, neq, bc_lower=bc[0],
                                 bc_upper=bc[1], mthlim=lim, grav=G, gamma=GAMMA), split=split)
    plain.upload(q)
    plain.init_ctrl(0.0, dt)
    cfl_p, _ = plain.step(dt)
    ones = np.ones((1, mx, my), order="F")
    g = device(solver, q, ones, dx, dy, 0, bc, split=split, mthlim=lim)
    g.init_ctrl(0.0, dt)
    cfl_c, _ = g.step(dt)
    assert cfl_c == cfl_p
    assert np.array_equal(g.download(), plain.download())


def test_pyclaw_front_end_with_capacity():
    """ClawSolver2D with state.index_capa, through Solution/Controller-style evolve_to_time, against
    the oracle's evolve with the same kappa."""
    mx, my = 64, 48
    solver = pyclaw.ClawSolver2D(riemann.shallow_roe_with_efix_2D)
    solver.dimensional_split = False
    solver.transverse_waves = 2
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.bc_lower = [pyclaw.BC.wall, pyclaw.BC.extrap]
    solver.bc_upper = [pyclaw.BC.wall, pyclaw.BC.extrap]
    solver.aux_bc_lower = [pyclaw.BC.wall, pyclaw.BC.extrap]
    solver.aux_bc_upper = [pyclaw.BC.wall, pyclaw.BC.extrap]
    domain = pyclaw.Domain((0.0, 0.0), (1.0, 1.0), (mx, my))
    state = pyclaw.State(domain, 3, 1)
    state.problem_data["grav"] = G
    q, aux, ic = problem("swe", mx, my, seed=2)
    state.q[...] = q
    state.aux[...] = aux
    state.index_capa = 0
    sol = pyclaw.Solution(state, domain)
    solver.evolve_to_time(sol, 0.05)
    d = oracle_desc("swe", mx, my, 1.0 / mx, 1.0 / my, 0)
    c = corc.make_ctrl(bc_lower=(3, 1), bc_upper=(3, 1))
    qo, st = corc.evolve(d, c, q, tend=0.05, aux=aux)
    assert solver.status["numsteps"] == st["numsteps"]
    assert abs(sol.t - 0.05) < 1e-14
    assert np.abs(state.q - qo).max() <= 1e-12 * np.abs(q).max()
