"""Source-term splitting on the HIP path (pyclaw's step_source / source_split / before_step): the
callbacks run on the device state through torch (DeviceState), the hyperbolic step is the HIP
kernel, and the time loop (accept/reject/dt) runs on the host per attempt.  Checked against the
oracle's source-splitting loop (corc.evolve_with_source) with the same source in numpy."""
import math

import numpy as np
import pytest

from clawpack_amd import problems, pyclaw, riemann
from oracle import corc

pytestmark = pytest.mark.gpu


def coriolis_np(f):
    def src(q, dt):
        c, s = math.cos(f * dt), math.sin(f * dt)
        hu = q[1].copy()
        q[1] = c * hu + s * q[2]
        q[2] = -s * hu + c * q[2]
    return src


def coriolis_device(f):
    def src(solver, state, dt):
        c, s = math.cos(f * dt), math.sin(f * dt)
        hu = state.q[1].clone()
        state.q[1].mul_(c).add_(state.q[2], alpha=s)
        state.q[2].mul_(c).add_(hu, alpha=-s)
    return src


def dam_break_solution(n, split):
    q0, dx, dy = problems.radial_dam_break(n, n)
    solver = pyclaw.ClawSolver2D(riemann.shallow_roe_with_efix_2D)
    solver.dimensional_split = split
    solver.transverse_waves = 2
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.bc_lower = [pyclaw.BC.periodic] * 2
    solver.bc_upper = [pyclaw.BC.periodic] * 2
    domain = pyclaw.Domain((-2.5, -2.5), (2.5, 2.5), (n, n))
    state = pyclaw.State(domain, 3)
    state.problem_data["grav"] = 1.0
    state.q[...] = q0
    return solver, pyclaw.Solution(state, domain), q0, dx, dy


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("source_split", [1, 2])
def test_coriolis_source_matches_oracle(source_split, split):
    n, f = 96, 2.0
    print("case source", source_split, split, flush=True)
    solver, sol, q0, dx, dy = dam_break_solution(n, split)
    solver.step_source = coriolis_device(f)
    solver.source_split = source_split
    solver.evolve_to_time(sol, 0.4)
    d = corc.make_desc("swe", n, n, dx, dy, grav=1.0)
    c = corc.make_ctrl(bc_lower=(2, 2), bc_upper=(2, 2), dimensional_split=split)
    qo, st = corc.evolve_with_source(d, c, q0, source=coriolis_np(f), source_split=source_split, tend=0.4)
    assert solver.status["numsteps"] == st["numsteps"] and solver.status["rejected"] == st["rejected"]
    assert abs(sol.t - st["t"]) <= 1e-14 and solver.dt == pytest.approx(st["dt"], rel=1e-12)
    assert np.abs(sol.state.q - qo).max() <= 1e-12 * np.abs(q0).max()
    assert np.abs(sol.state.q[2]).max() > 1e-3  # the rotation moved momentum into hv


def test_before_step_runs_per_attempt_and_one_step():
    n = 64
    solver, sol, q0, dx, dy = dam_break_solution(n, False)
    seen = []
    solver.before_step = lambda s, state: seen.append((state.t, float(state.q[0].sum())))
    solver.evolve_to_time(sol, 0.3)
    st = solver.status
    assert len(seen) == st["numsteps"] + st["rejected"] and st["rejected"] >= 1
    # the device loop without callbacks gives the same state
    solver2, sol2, _, _, _ = dam_break_solution(n, False)
    solver2.evolve_to_time(sol2, 0.3)
    assert np.array_equal(sol.state.q, sol2.state.q) and sol.t == sol2.t
    # one step attempt (tend=None), with a source
    solver.step_source = coriolis_device(1.0)
    dt0, t0 = solver.dt, sol.t
    solver.evolve_to_time(sol)
    assert solver.status["numsteps"] + solver.status["rejected"] == 1
    assert solver.dt == dt0 and (sol.t == t0 + dt0 or sol.t == t0)


def test_bathymetry_friction_source_reads_aux():
    """Sill, f-wave solver, and a depth-dependent linear friction k(b) = 0.5 (1 - b) computed on the
    device from state.aux, Godunov splitting, to t = 0.3."""
    mx, my = 120, 96
    q0, aux, dx, dy = problems.sill(mx, my)
    solver = pyclaw.ClawSolver2D(riemann.shallow_bathymetry_fwave_2D)
    solver.fwave = True
    solver.dimensional_split = False
    solver.transverse_waves = 2
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.bc_lower = solver.aux_bc_lower = [pyclaw.BC.extrap, pyclaw.BC.wall]
    solver.bc_upper = solver.aux_bc_upper = [pyclaw.BC.extrap, pyclaw.BC.wall]

    def friction(s, state, dt):
        k = 0.5 * (1.0 - state.aux[0])
        state.q[1:3] *= (-k * dt).exp()

    solver.step_source = friction
    domain = pyclaw.Domain((-1.0, -1.0), (1.0, 1.0), (mx, my))
    state = pyclaw.State(domain, 3, 1)
    state.problem_data.update(grav=9.81, dry_tolerance=1e-3, sea_level=0.0)
    state.q[...] = q0
    state.aux[...] = aux
    sol = pyclaw.Solution(state, domain)
    solver.evolve_to_time(sol, 0.3)

    def friction_np(q, dt):
        q[1:3] *= np.exp(-0.5 * (1.0 - aux[0]) * dt)

    d = corc.make_desc("bathy", mx, my, dx, dy, grav=9.81, mthlim=(4, 4, 4), dry_tolerance=1e-3)
    c = corc.make_ctrl(bc_lower=(1, 3), bc_upper=(1, 3))
    qo, st = corc.evolve_with_source(d, c, q0, source=friction_np, tend=0.3, aux=aux)
    assert solver.status["numsteps"] == st["numsteps"]
    assert np.abs(state.q - qo).max() <= 1e-12 * np.abs(q0).max()


@pytest.mark.parametrize("source_split", [1, 2])
def test_time_dependent_source_sees_the_step_start_time(source_split):
    """A source that reads state.t gets the attempt's start time t_n (pyclaw advances solution.t
    after Solver.step): on a uniform state at rest dh/dt = t adds dt * t_n per step, the oracle's
    run with the same source agrees, and the closed form h_0 + sum dt t_n holds."""
    n = 48
    print("case time source", source_split, flush=True)
    solver, sol, q0, dx, dy = dam_break_solution(n, False)
    sol.state.q[0] = 1.0
    q0 = sol.state.q.copy()
    solver.step_source = lambda s, state, dt: state.q[0].add_(dt * state.t)
    solver.source_split = source_split
    solver.evolve_to_time(sol, 0.5)
    d = corc.make_desc("swe", n, n, dx, dy, grav=1.0)
    c = corc.make_ctrl(bc_lower=(2, 2), bc_upper=(2, 2))

    def src(q, dt, t):
        q[0] += dt * t

    qo, st = corc.evolve_with_source(d, c, q0, source=src, source_split=source_split, tend=0.5,
                                     source_t=True)
    assert solver.status["numsteps"] == st["numsteps"] and abs(sol.t - st["t"]) <= 1e-14
    assert np.abs(sol.state.q - qo).max() <= 1e-13
    # every step is accepted at a CFL-chosen dt; without momentum the left sum is the exact
    # integral's left Riemann sum, below t^2/2 by sum(dt^2)/2
    assert 1.0 < sol.state.q[0, 0, 0] < 1.0 + 0.5 * 0.5 ** 2


@pytest.mark.parametrize("callbacks", [False, True])
def test_max_steps_spent_on_rejected_attempts(callbacks):
    """pyclaw raises "Maximum number of timesteps" only when all max_steps attempts were accepted
    short of tend.  The 200^2 dam break rejects its dt_initial = 0.1 attempt: with max_steps = 3
    it takes 2 steps and returns short of tend, on the device loop and on the host loop that a
    before_step callback selects, like the oracle; dt_initial = 0.002, whose three attempts are
    all accepted (the oracle's status says so), raises."""
    print("case max_steps callbacks", callbacks, flush=True)
    q0, dx, dy = problems.radial_dam_break(200, 200)
    d = corc.make_desc("swe", 200, 200, dx, dy, grav=1.0)
    _, st = corc.evolve(d, corc.make_ctrl(max_steps=3), q0, tend=1.0)
    assert st["error"] == 0 and st["numsteps"] == 2 and st["rejected"] == 1
    _, st2 = corc.evolve(d, corc.make_ctrl(max_steps=3, dt_initial=0.002), q0, tend=1.0)
    assert st2["error"] == 1 and st2["numsteps"] == 3 and st2["rejected"] == 0
    for dt0, raises in ((0.1, False), (0.002, True)):
        solver, sol, _, _, _ = dam_break_solution(200, False)
        solver.bc_lower = [pyclaw.BC.extrap] * 2
        solver.bc_upper = [pyclaw.BC.extrap] * 2
        sol.state.q[...] = q0
        solver.max_steps = 3
        solver.dt_initial = dt0
        if callbacks:
            solver.before_step = lambda s, state: None
        if raises:
            with pytest.raises(Exception, match="Maximum number of timesteps"):
                solver.evolve_to_time(sol, 1.0)
            assert solver.status["numsteps"] == 3
        else:
            solver.evolve_to_time(sol, 1.0)
            assert solver.status["numsteps"] == 2 and solver.status["rejected"] == 1
            assert abs(sol.t - st["t"]) <= 1e-14 and sol.t < 1.0
