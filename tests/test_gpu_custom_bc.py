"""BC.custom on the HIP path: pyclaw's user_bc_lower/user_bc_upper(state, dim, t, qbc, auxbc,
num_ghost) fill the ghost layers of a padded device copy of the state (torch views, pyclaw layout),
and claw2d_step_padded steps from it; with dimensional_split the x-sweeps of claw2d_sweep_padded
write q* into a second padded buffer, the same functions fill its ghost cells, and the y-sweeps
finish the step.  Checked against the oracle's numpy _apply_bcs with the same functions in numpy
(corc.apply_bcs, corc.evolve_with_source)."""
import numpy as np
import pytest
import torch

from clawpack_amd import problems, pyclaw, riemann
from oracle import corc

pytestmark = pytest.mark.gpu


def axis_of(dim):
    return 1 if dim.name == "x" else 2


def extrap_device(side):
    def fn(state, dim, t, qbc, auxbc, g):
        a = axis_of(dim)
        n = qbc.shape[a]
        for i in range(g):
            if side == 0:
                qbc.select(a, i).copy_(qbc.select(a, g))
            else:
                qbc.select(a, n - 1 - i).copy_(qbc.select(a, n - 1 - g))
    return fn


def inflow_device(state, dim, t, qbc, auxbc, g):
    qbc[0, :g, :] = 1.5
    qbc[1, :g, :] = 0.6
    qbc[2, :g, :] = 0.0


def inflow_np(state, idim, t, qbc, auxbc, g):
    qbc[0, :g, :] = 1.5
    qbc[1, :g, :] = 0.6
    qbc[2, :g, :] = 0.0


def swe_solution(mx, my, q0, bc_lower, bc_upper, split=False):
    solver = pyclaw.ClawSolver2D(riemann.shallow_roe_with_efix_2D)
    solver.dimensional_split = split
    solver.transverse_waves = 2
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.bc_lower, solver.bc_upper = list(bc_lower), list(bc_upper)
    domain = pyclaw.Domain((0.0, 0.0), (1.0, 1.0), (mx, my))
    state = pyclaw.State(domain, 3)
    state.problem_data["grav"] = 1.0
    state.q[...] = q0
    return solver, pyclaw.Solution(state, domain)


@pytest.mark.parametrize("split", [False, True])
def test_custom_function_writing_extrap_matches_bc_extrap(split):
    """Split: the fused one-pass kernel x-sweeps the ghost rows' source rows where the padded pair
    refills q*; for extrap the two agree (to rounding)."""
    print("case custom extrap split=%s" % split, flush=True)
    mx, my = 96, 80
    q0 = problems.random_smooth_swe(mx, my, 3)
    s1, sol1 = swe_solution(mx, my, q0, [pyclaw.BC.custom, pyclaw.BC.wall], [pyclaw.BC.custom, pyclaw.BC.wall], split)
    s1.user_bc_lower, s1.user_bc_upper = extrap_device(0), extrap_device(1)
    s1.evolve_to_time(sol1, 0.05)
    s2, sol2 = swe_solution(mx, my, q0, [pyclaw.BC.extrap, pyclaw.BC.wall], [pyclaw.BC.extrap, pyclaw.BC.wall], split)
    s2.evolve_to_time(sol2, 0.05)
    assert s1.status["numsteps"] == s2.status["numsteps"] and sol1.t == sol2.t
    assert np.abs(sol1.state.q - sol2.state.q).max() <= 1e-12 * np.abs(q0).max()


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("mx,my", [(80, 24), (131, 67)])
def test_inflow_matches_oracle(mx, my, split):
    print("case custom inflow", mx, my, "split=%s" % split, flush=True)
    q0 = np.zeros((3, mx, my), order="F")
    q0[0] = 1.0
    solver, sol = swe_solution(mx, my, q0, [pyclaw.BC.custom, pyclaw.BC.wall], [pyclaw.BC.extrap, pyclaw.BC.wall], split)
    solver.user_bc_lower = inflow_device
    solver.evolve_to_time(sol, 0.3)
    d = corc.make_desc("swe", mx, my, 1.0 / mx, 1.0 / my, grav=1.0)
    c = corc.make_ctrl(bc_lower=(0, 3), bc_upper=(1, 3), dimensional_split=split)
    qo, st = corc.evolve_with_source(d, c, q0, tend=0.3, user_bc_lower=inflow_np)
    assert solver.status["numsteps"] == st["numsteps"] and solver.status["rejected"] == st["rejected"]
    assert np.abs(sol.state.q - qo).max() <= 1e-12 * 1.5
    # one step_hyperbolic with the same boundary function
    solver.dt = 0.004
    solver.step_hyperbolic(sol)
    qbc = corc.apply_bcs(d, c, qo, 0.3, inflow_np)
    if split:
        qs, _ = corc.step2ds(d, qbc, 0.004, 1)
        qs = corc.apply_bcs(d, c, qs[:, 2:-2, 2:-2], 0.3, inflow_np)
        q1, _ = corc.step2ds(d, qs, 0.004, 2)
    else:
        q1, _ = corc.step2(d, qbc, 0.004)
    assert np.abs(sol.state.q - q1[:, 2:-2, 2:-2]).max() <= 1e-12 * 1.5


@pytest.mark.parametrize("split", [False, True])
def test_lake_at_rest_with_a_custom_sea_level_boundary(split):
    """Bathymetry, f-waves: the left side's ghost cells are set by a user function to the state at
    rest at sea level 0 over the ghost bathymetry it reads from auxbc; the lake stays at rest."""
    print("case custom sea level split=%s" % split, flush=True)
    mx, my = 120, 90
    _, aux = problems.random_bathymetry(mx, my, seed=5)
    h = -aux[0]
    q0 = np.asfortranarray(np.stack([h, 0 * h, 0 * h]))

    def sea_level(state, dim, t, qbc, auxbc, g):
        qbc[0, :g, :] = -auxbc[0, :g, :]
        qbc[1:, :g, :] = 0.0

    solver = pyclaw.ClawSolver2D(riemann.shallow_bathymetry_fwave_2D)
    solver.fwave = True
    solver.dimensional_split = split
    solver.bc_lower = [pyclaw.BC.custom, pyclaw.BC.wall]
    solver.bc_upper = [pyclaw.BC.wall, pyclaw.BC.wall]
    solver.aux_bc_lower = [pyclaw.BC.extrap, pyclaw.BC.wall]
    solver.aux_bc_upper = [pyclaw.BC.wall, pyclaw.BC.wall]
    solver.user_bc_lower = sea_level
    domain = pyclaw.Domain((0.0, 0.0), (1.0, 1.0), (mx, my))
    state = pyclaw.State(domain, 3, 1)
    state.problem_data.update(grav=9.81, dry_tolerance=1e-3)
    state.q[...] = q0
    state.aux[...] = aux
    sol = pyclaw.Solution(state, domain)
    solver.evolve_to_time(sol, 0.2)
    assert solver.status["numsteps"] > 20
    assert np.abs(state.q[0] + aux[0]).max() <= 1e-13
    assert np.abs(state.q[1:]).max() <= 1e-13


def test_split_fp32_inflow_matches_oracle():
    """FP32 state, split step, custom inflow side: within 1e-5 of the FP64 oracle over the run."""
    print("case custom inflow fp32 split", flush=True)
    mx, my = 100, 40
    q0 = np.zeros((3, mx, my), order="F")
    q0[0] = 1.0
    solver, sol = swe_solution(mx, my, q0, [pyclaw.BC.custom, pyclaw.BC.extrap], [pyclaw.BC.extrap, pyclaw.BC.extrap], True)
    solver.precision = "fp32"
    solver.dt_variable = False
    solver.dt_initial = 0.004
    solver.user_bc_lower = inflow_device
    solver.evolve_to_time(sol, 0.2)
    d = corc.make_desc("swe", mx, my, 1.0 / mx, 1.0 / my, grav=1.0)
    c = corc.make_ctrl(bc_lower=(0, 1), bc_upper=(1, 1), dimensional_split=True, dt_variable=False)
    qo, st = corc.evolve_with_source(d, c, q0, tend=0.2, dt=0.004, user_bc_lower=inflow_np)
    assert solver.status["numsteps"] == st["numsteps"] >= 49
    assert np.abs(sol.state.q - qo).max() <= 1e-5 * 1.5
