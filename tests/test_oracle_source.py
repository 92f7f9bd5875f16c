"""Pins of the oracle's source-splitting loop (corc.evolve_with_source: Solver.evolve_to_time with
Solver.step's Godunov / Strang splitting of a step_source callback).

  * with no source it is orc_evolve bit for bit (to tend and for a step count);
  * a Coriolis source on a uniform periodic flow: the hyperbolic step leaves a uniform state
    unchanged, so the velocity must rotate by exactly f t (closed form), for both splittings;
  * Strang splitting applies half source steps on either side: with a damping source, k dt/2
    before and after each step, so a uniform state decays by exp(-k t) like Godunov's;
  * before_step runs once per attempt, rejected attempts included, and a rejected attempt
    leaves no trace of its Strang half step.
"""
import numpy as np
import pytest

from clawpack_amd import problems
from oracle import corc


def coriolis(f):
    def src(q, dt):
        c, s = np.cos(f * dt), np.sin(f * dt)
        hu = q[1].copy()
        q[1] = c * hu + s * q[2]
        q[2] = -s * hu + c * q[2]
    return src


def test_no_source_is_orc_evolve():
    q, dx, dy = problems.radial_dam_break(40, 36)
    d = corc.make_desc("swe", 40, 36, dx, dy, grav=1.0)
    for split in (False, True):
        c = corc.make_ctrl(dimensional_split=split)
        a, sa = corc.evolve(d, c, q, tend=0.35)
        b, sb = corc.evolve_with_source(d, c, q, tend=0.35)
        assert np.array_equal(a, b)
        assert (sa["numsteps"], sa["rejected"], sa["t"], sa["dt"]) == (sb["numsteps"], sb["rejected"], sb["t"], sb["dt"])


@pytest.mark.parametrize("split_src", [1, 2])
def test_coriolis_rotates_a_uniform_flow_exactly(split_src):
    mx, my, f = 24, 20, 3.0
    q = np.zeros((3, mx, my), order="F")
    q[0], q[1], q[2] = 1.5, 0.3, -0.2
    d = corc.make_desc("swe", mx, my, 1.0 / mx, 1.0 / my, grav=1.0)
    c = corc.make_ctrl(bc_lower=(2, 2), bc_upper=(2, 2), dt_initial=0.01)
    qn, st = corc.evolve_with_source(d, c, q, source=coriolis(f), source_split=split_src, tend=1.3)
    assert abs(st["t"] - 1.3) < 1e-14 and st["numsteps"] > 20
    th = f * st["t"]
    hu = np.cos(th) * 0.3 + np.sin(th) * (-0.2)
    hv = -np.sin(th) * 0.3 + np.cos(th) * (-0.2)
    assert np.abs(qn[0] - 1.5).max() == 0.0
    assert np.abs(qn[1] - hu).max() <= 1e-13 and np.abs(qn[2] - hv).max() <= 1e-13


def test_strang_half_steps_and_rejected_attempts():
    mx, my, k = 30, 30, 2.0
    q, dx, dy = problems.radial_dam_break(mx, my)
    q[1] = 0.2
    d = corc.make_desc("swe", mx, my, dx, dy, grav=1.0)
    c = corc.make_ctrl(bc_lower=(2, 2), bc_upper=(2, 2), dt_initial=0.5)  # the first attempt is rejected
    log = []

    def damp(q, dt):
        log.append(("src", dt))
        q[1:] *= np.exp(-k * dt)

    def before(q, t):
        log.append(("before", t))

    qn, st = corc.evolve_with_source(d, c, q, source=damp, source_split=2, before_step=before, nsteps=6)
    assert st["rejected"] >= 1 and st["numsteps"] == 6
    assert sum(1 for e in log if e[0] == "before") == st["attempts"]
    # two half steps per accepted attempt and one per rejected attempt
    assert sum(1 for e in log if e[0] == "src") == 2 * st["numsteps"] + st["rejected"]
    # total momentum decays exactly like exp(-k t) (the hyperbolic step conserves it, periodic)
    m0 = corc.component_sum(d, q, 1)
    m1 = corc.component_sum(d, qn, 1)
    assert abs(m1 - m0 * np.exp(-k * st["t"])) <= 1e-12 * abs(m0)


# ------------------------------------------------------------------------------------------------
# BC.custom: the numpy restatement of _apply_bcs with user functions
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("bcl,bcu,ri", [((1, 1), (1, 1), (1, 2)), ((2, 3), (2, 3), (1, 2)),
                                         ((3, 1), (3, 2), (2, 1)), ((3, 3), (1, 3), (1, 2))])
def test_apply_bcs_standard_rules_equal_orc_fill_ghosts(bcl, bcu, ri):
    if (bcl[1] == 2) != (bcu[1] == 2):
        bcu = (bcu[0], bcl[1])
    rng = np.random.default_rng(1)
    mx, my = 9, 7
    q = np.asfortranarray(rng.standard_normal((3, mx, my)))
    d = corc.make_desc("swe", mx, my, 0.1, 0.1)
    c = corc.make_ctrl(bc_lower=bcl, bc_upper=bcu, reflect_index=ri)
    assert np.array_equal(corc.apply_bcs(d, c, q), corc.fill_ghosts(d, c, q))


def extrap_user(side):
    """A user BC function that writes the extrapolation rule itself (x or y, lower or upper)."""
    def fn(state, idim, t, qbc, auxbc, g):
        a = np.moveaxis(qbc, idim + 1, 1)
        for i in range(g):
            if side == 0:
                a[:, i] = a[:, g]
            else:
                a[:, -i - 1] = a[:, -g - 1]
    return fn


@pytest.mark.parametrize("split", [False, True])
def test_custom_function_reproducing_extrap_gives_the_extrap_run(split):
    """With splitting the user function also fills q*'s ghost cells between the sweeps, as the
    standard rules are refilled there, so the runs agree bit for bit either way."""
    q, dx, dy = problems.radial_dam_break(40, 30)
    d = corc.make_desc("swe", 40, 30, dx, dy, grav=1.0)
    ref, sr = corc.evolve(d, corc.make_ctrl(bc_lower=(1, 3), bc_upper=(1, 3), dimensional_split=split),
                          q, tend=0.4)
    c = corc.make_ctrl(bc_lower=(0, 3), bc_upper=(0, 3), dimensional_split=split)
    got, sg = corc.evolve_with_source(d, c, q, tend=0.4, user_bc_lower=extrap_user(0),
                                      user_bc_upper=extrap_user(1))
    assert sg["numsteps"] == sr["numsteps"] and sg["t"] == sr["t"]
    assert np.array_equal(got, ref)


def test_split_calls_the_user_function_before_each_sweep():
    """Split step with a custom lower-x side: two calls per attempt (q^n, then q*), both for
    dimension 0 at the attempt's t, the second seeing the x-swept interior."""
    q, dx, dy = problems.radial_dam_break(24, 20)
    d = corc.make_desc("swe", 24, 20, dx, dy, grav=1.0)
    c = corc.make_ctrl(bc_lower=(0, 1), bc_upper=(1, 1), dimensional_split=True)
    calls = []
    ext = extrap_user(0)

    def fn(state, idim, t, qbc, auxbc, g):
        calls.append((idim, t, qbc[:, g:-g, g:-g].copy()))
        ext(state, idim, t, qbc, auxbc, g)

    dt = 0.01
    qn, st = corc.evolve_with_source(d, c, q, nsteps=3, dt=dt, user_bc_lower=fn)
    assert st["rejected"] == 0 and len(calls) == 2 * st["attempts"] == 6
    assert all(idim == 0 for idim, _, _ in calls)
    assert calls[0][1] == calls[1][1] == 0.0
    assert np.array_equal(calls[0][2], q)
    qbc = corc.apply_bcs(d, corc.make_ctrl(bc_lower=(1, 1), bc_upper=(1, 1)), q)
    qs, _ = corc.step2ds(d, qbc, dt, 1)
    assert np.array_equal(calls[1][2], qs[:, 2:-2, 2:-2])


@pytest.mark.parametrize("split", [False, True])
def test_custom_inflow_boundary_sends_a_bore(split):
    """An inflow state h = 1.5, hu = 0.6 imposed in the left ghost columns of still water h = 1
    drives a bore into the domain: mass grows by about the inflow rate times t, and the far right
    stays at rest."""
    mx, my = 80, 12
    q = np.zeros((3, mx, my), order="F")
    q[0] = 1.0
    d = corc.make_desc("swe", mx, my, 1.0 / mx, 1.0 / my, grav=1.0)
    c = corc.make_ctrl(bc_lower=(0, 3), bc_upper=(1, 3), dimensional_split=split)

    def inflow(state, idim, t, qbc, auxbc, g):
        qbc[0, :g, :] = 1.5
        qbc[1, :g, :] = 0.6
        qbc[2, :g, :] = 0.0

    qn, st = corc.evolve_with_source(d, c, q, tend=0.3, user_bc_lower=inflow)
    mass0, mass1 = q[0].sum() / (mx * my), qn[0].sum() / (mx * my)
    assert 0.5 * 0.6 * 0.3 < mass1 - mass0 < 1.5 * 0.6 * 0.3
    assert np.array_equal(qn[:, -10:], q[:, -10:])


@pytest.mark.parametrize("split_src", [1, 2])
def test_time_dependent_source_sees_the_step_start_time(split_src):
    """Solver.step runs before evolve_to_time advances solution.t, so a source that reads the time
    sees t_n in both of its calls.  On a uniform periodic state at rest (a fixed point of the
    hyperbolic step) the source dh/dt = t then adds dt * t_n per step, the left Riemann sum:
    h_N = h_0 + dt^2 N (N - 1) / 2 for a fixed dt; seeing t_{n+1} would add dt^2 N (N + 1) / 2."""
    mx, my, dt, n = 12, 10, 0.01, 30
    q = np.zeros((3, mx, my), order="F")
    q[0] = 1.0
    d = corc.make_desc("swe", mx, my, 1.0 / mx, 1.0 / my, grav=1.0)
    c = corc.make_ctrl(bc_lower=(2, 2), bc_upper=(2, 2), dt_initial=dt, dt_variable=False)

    def src(qq, h, t):
        qq[0] += h * t

    qn, st = corc.evolve_with_source(d, c, q, source=src, source_split=split_src, nsteps=n,
                                     source_t=True)
    assert st["numsteps"] == n and abs(st["t"] - n * dt) < 1e-15
    assert np.abs(qn[0] - (1.0 + dt * dt * n * (n - 1) / 2)).max() <= 1e-14
    assert np.all(qn[1:] == 0.0)


def test_max_steps_spent_on_rejected_attempts_is_not_an_error():
    """Solver.evolve_to_time raises "Maximum number of timesteps" only when all max_steps attempts
    were accepted short of tend.  Here the dt_initial = 0.1 attempt is rejected, so 3 attempts give
    2 steps: the loop ends short of tend without an error, in orc_evolve and the source loop alike;
    with a dt that is never rejected the same max_steps is an error."""
    q, dx, dy = problems.radial_dam_break(40, 40)
    d = corc.make_desc("swe", 40, 40, dx, dy, grav=1.0)
    c = corc.make_ctrl(max_steps=3)
    for run in (corc.evolve, corc.evolve_with_source):
        _, st = run(d, c, q, tend=1.0)
        assert st["error"] == 0 and st["numsteps"] == 2 and st["rejected"] == 1 and st["t"] < 1.0
        _, st = run(d, c, q, tend=1.0, dt=0.01)
        assert st["error"] == 1 and st["numsteps"] == 3 and st["rejected"] == 0
