"""FP32 state/flux path against the FP64 path, and the compensated FP64 mass sum.

Bar stated per result: FP32 after N steps vs FP64 on the same dt sequence, relative to max|q|.
"""
import numpy as np
import pytest

from clawpack_amd import problems
from oracle import corc

pytestmark = pytest.mark.gpu


def _grid(dtype, q, dx, dy, bc=(2, 2), depth_ref=0.0):
    from clawpack_amd.device import DeviceGrid, make_desc
    neq, mx, my = q.shape
    d = make_desc(0, dtype, mx, my, dx, dy, 3, 3, bc_lower=bc, bc_upper=bc, grav=1.0, depth_ref=depth_ref)
    g = DeviceGrid(d)
    g.upload(q)
    return g


def test_fp32_vs_fp64_fixed_dt_sequence():
    n, steps = 512, 200
    print("case fp32 vs fp64 humps", n, steps, flush=True)
    q, dx, dy = problems.gaussian_humps(n, n, 2024)
    g64, g32 = _grid(0, q, dx, dy), _grid(1, q, dx, dy)
    g64.init_ctrl(0.0, 1e-3, target_steps=steps)
    g64.advance(steps + 2)
    c = g64.sync_cur()
    dt = c.t / steps
    # FP32 on the same (uniform) dt sequence: fixed-dt steps
    g64b = _grid(0, q, dx, dy)
    for g in (g32, g64b):
        g.init_ctrl(0.0, dt, target_steps=steps)
        for _ in range(steps):
            g.enqueue_step(dt)
    q64, q32 = g64b.download(), g32.download()
    scale = np.abs(q64).max()
    linf = np.abs(q32 - q64).max() / scale
    l1 = np.abs(q32 - q64).mean() / np.abs(q64).mean()
    print("fp32 vs fp64: Linf %.3e  L1 %.3e" % (linf, l1))
    assert linf <= 1e-5 and l1 <= 1e-6


def test_mass_sum_compensated_and_conserved():
    n = 384
    q, dx, dy = problems.gaussian_humps(n, n, 11)
    exact = corc.component_sum(corc.make_desc("swe", n, n, dx, dy), q, 0)
    for dtype in (0, 1):
        g = _grid(dtype, q, dx, dy)
        s0 = g.mass_sum(0)
        ref = exact if dtype == 0 else float(np.asarray(q[0], dtype=np.float32).astype(np.float64).sum())
        assert abs(s0 - ref) <= 1e-13 * abs(ref)
        g.init_ctrl(0.0, 1e-3, target_steps=100)
        g.advance(102)
        g.sync_cur()
        s1 = g.mass_sum(0)
        drift = abs(s1 - s0) / s0
        print("dtype %d mass drift after 100 periodic steps: %.3e" % (dtype, drift))
        assert drift <= (1e-13 if dtype == 0 else 1e-6)
        # determinism: same sum twice, bit for bit
        assert g.mass_sum(0) == s1


def test_fp32_step_matches_oracle_loosely():
    q = problems.random_smooth_swe(70, 33, 4)
    d = corc.make_desc("swe", 70, 33, 0.1, 0.1, grav=1.0)
    qo, cflo = corc.step_hyperbolic(d, corc.make_ctrl(), q, 0.01)
    g = _grid(1, q, 0.1, 0.1, bc=(1, 1))
    g.init_ctrl(0.0, 0.01)
    cfl, _ = g.step(0.01)
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    assert err <= 2e-6, err
    assert abs(cfl - cflo) <= 1e-6 * cflo


@pytest.mark.parametrize("bc", [(1, 1), (2, 2), (1, 3)])
def test_euler_fp32_specialised_kernel_matches_oracle(bc):
    """The FP32 Euler kernel that the bench times (uniform MC, order 2, transverse_waves 2: the
    method fixed at compile time) against the FP64 oracle: one step of random smooth data, and 20
    controller steps of the four-quadrant problem at 192^2 with the oracle's dt sequence.  FP32
    rounding bounds: 2e-6 max|q| for the step, 1e-4 after 20 steps through the shocks."""
    from clawpack_amd.device import DeviceGrid, make_desc
    print("case euler fp32 specialised", bc, flush=True)
    q = problems.random_euler(70, 33, 4)
    d = make_desc(1, 1, 70, 33, 0.1, 0.1, 4, 4, bc_lower=bc, bc_upper=bc, gamma=1.4,
                  mthlim=(4, 4, 4, 4), order=2, transverse_waves=2)
    g = DeviceGrid(d)
    g.upload(q)
    g.init_ctrl(0.0, 0.005)
    cfl, _ = g.step(0.005)
    od = corc.make_desc("euler", 70, 33, 0.1, 0.1, gamma=1.4)
    qo, cflo = corc.step_hyperbolic(od, corc.make_ctrl(bc_lower=bc, bc_upper=bc), q, 0.005)
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    print("  one step: rel err %.3e" % err, flush=True)
    assert err <= 2e-6, err
    assert abs(cfl - cflo) <= 1e-6 * cflo
    n, steps = 192, 20
    q, dx, dy = problems.euler_quadrants(n, n, 1)
    od = corc.make_desc("euler", n, n, dx, dy, gamma=1.4)
    c = corc.make_ctrl(bc_lower=bc, bc_upper=bc)
    qo, st = corc.evolve(od, c, q, nsteps=steps)
    d = make_desc(1, 1, n, n, dx, dy, 4, 4, bc_lower=bc, bc_upper=bc, gamma=1.4)
    g = DeviceGrid(d)
    g.upload(q)
    dts = np.diff(np.r_[0.0, _oracle_times(od, c, q, steps)])
    g.init_ctrl(0.0, float(dts[0]), target_steps=steps)
    for dt in dts:
        g.enqueue_step(float(dt))
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    print("  %d steps: rel err %.3e" % (steps, err), flush=True)
    assert err <= 1e-4, err


def _oracle_times(od, c, q, steps):
    """The oracle's accepted step times t_1..t_steps (one controller step per call)."""
    out, qq, t, dt = [], q, 0.0, c.dt_initial
    while len(out) < steps:
        qn, cfl = corc.step_hyperbolic(od, c, qq, dt)
        if cfl <= c.cfl_max:
            qq, t = qn, t + dt
            out.append(t)
        dt = min(c.dt_max, dt * c.cfl_desired / cfl)
    return np.array(out)


def test_fp32_per_component_error_with_stored_depth_deviation():
    """The FP32 path with the depth stored as h - 1 (depth_ref = 1, the humps' still-water level)
    against the FP64 path after 1000 fixed-dt steps at 2048^2: every component, relative to its
    own scale (L1 = sum|dq_m| / sum|q_m|, Linf = max|dq_m| / max|q_m|), within 1e-5.  The same run
    storing h itself is reported beside it (its momentum error is the larger one)."""
    import math
    import torch
    n, steps = 2048, 1000
    print("case fp32 per component", n, steps, flush=True)
    q, dx, dy = problems.gaussian_humps(n, n, 2024)
    dt = 0.8 * min(dx, dy) / (math.sqrt(float(q[0].max())) + 0.5)
    out = {}
    for name, dtype, ref in (("fp64", 0, 0.0), ("fp32", 1, 1.0), ("fp32_h", 1, 0.0)):
        g = _grid(dtype, q, dx, dy, depth_ref=ref)
        g.init_ctrl(0.0, dt, target_steps=steps)
        for _ in range(steps):
            g.enqueue_step(dt)
        c = g.read_ctrl()
        assert c.status == 0 and c.cfl <= 1.0
        p = g.planes().to(torch.float64).clone()
        p[0] += ref
        out[name] = p
        del g
    a = out["fp64"]
    errs = {}
    for name in ("fp32", "fp32_h"):
        b = out[name]
        errs[name] = [(float((b[m] - a[m]).abs().sum() / a[m].abs().sum()),
                       float((b[m] - a[m]).abs().max() / a[m].abs().max())) for m in range(3)]
        print("  %-7s L1/Linf per component: %s" % (name, " ".join("%.2e/%.2e" % e for e in errs[name])),
              flush=True)
    for l1, linf in errs["fp32"]:
        assert l1 <= 1e-5 and linf <= 1e-5, errs["fp32"]
    # the deviation storage is the better of the two on the momenta
    assert max(e[1] for e in errs["fp32"][1:]) < max(e[1] for e in errs["fp32_h"][1:])


def test_fp32_depth_ref_front_end_and_mass_sum():
    """The PyClaw front end picks depth_ref (the mean initial depth) for FP32 shallow water; the
    state it returns is h, and the device mass sum is the sum of h."""
    from clawpack_amd import pyclaw, riemann
    n = 96
    q, dx, dy = problems.gaussian_humps(n, n, 7)
    domain = pyclaw.Domain([pyclaw.Dimension(-2.5, 2.5, n, name="x"), pyclaw.Dimension(-2.5, 2.5, n, name="y")])
    state = pyclaw.State(domain, 3)
    state.problem_data["grav"] = 1.0
    state.q[...] = q
    solver = pyclaw.ClawSolver2D(riemann.shallow_roe_with_efix_2D)
    solver.precision = "fp32"
    solver.dimensional_split = False
    solver.bc_lower = solver.bc_upper = [pyclaw.BC.periodic, pyclaw.BC.periodic]
    sol = pyclaw.Solution(state, domain)
    solver.setup(sol)
    ref = solver._depth_ref(solver._policy())
    assert ref == float(np.float32(q[0].mean())) and ref > 1.0
    total0 = solver.total(sol, 0)
    assert abs(total0 - q[0].sum()) <= 1e-6 * q[0].sum()
    solver.evolve_to_time(sol, 0.2)
    assert abs(solver.total(sol, 0) - total0) <= 2e-6 * total0
    assert abs(state.q[0].mean() - q[0].mean()) <= 1e-6
    # the front end's method: pyclaw's defaults, minmod and transverse_waves = 1
    d64 = corc.make_desc("swe", n, n, dx, dy, grav=1.0, transverse_waves=1, mthlim=(1, 1, 1))
    c = corc.make_ctrl(bc_lower=(2, 2), bc_upper=(2, 2), dimensional_split=False)
    qo, _ = corc.evolve(d64, c, q, tend=0.2)
    err = np.abs(state.q - qo).max() / np.abs(qo).max()
    print("front end fp32 (depth_ref %.6f) vs oracle at t=0.2: %.2e" % (ref, err))
    assert err <= 1e-5
