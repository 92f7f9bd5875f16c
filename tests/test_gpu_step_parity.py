"""HIP step vs the CPU FP64 oracle: one fixed-dt step on small grids, every BC, both solvers.

Bar: max|dq| <= 1e-12 * max|q| (FP64), CFL number to 1e-13 relative.
"""
import numpy as np
import pytest

from oracle import corc
from clawpack_amd import problems

pytestmark = pytest.mark.gpu

BCS = [(1, 1), (2, 2), (3, 3), (1, 3), (3, 2)]


def _grid(solver, mx, my, dx, dy, bc, dtype=0, order=2, tw=2, lim=(4, 4, 4, 4), seg_h=0):
    from clawpack_amd.device import DeviceGrid, make_desc
    neq = 3 if solver == "swe" else 4
    d = make_desc(0 if solver == "swe" else 1, dtype, mx, my, dx, dy, neq, neq, order=order,
                  transverse_waves=tw, mthlim=lim, bc_lower=bc, bc_upper=bc, grav=1.0, gamma=1.4,
                  seg_h=seg_h)
    return DeviceGrid(d)


def _oracle_step(solver, q, dx, dy, bc, dt, order=2, tw=2, lim=(4, 4, 4, 4)):
    mx, my = q.shape[1:]
    d = corc.make_desc(solver, mx, my, dx, dy, grav=1.0, gamma=1.4, order=order,
                       transverse_waves=tw, mthlim=lim)
    c = corc.make_ctrl(bc_lower=bc, bc_upper=bc)
    return corc.step_hyperbolic(d, c, q, dt)


@pytest.mark.parametrize("solver", ["swe", "euler"])
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("shape", [(23, 17), (64, 9), (130, 75)])
def test_single_step(solver, bc, shape):
    mx, my = shape
    print("case", solver, bc, shape, flush=True)
    q = problems.random_smooth_swe(mx, my, 1) if solver == "swe" else problems.random_euler(mx, my, 2)
    dx, dy, dt = 0.1, 0.13, 0.02
    qo, cflo = _oracle_step(solver, q, dx, dy, bc, dt)
    g = _grid(solver, mx, my, dx, dy, bc, seg_h=7)
    g.init_ctrl(0.0, dt)
    g.upload(q)
    cfl, _ = g.step(dt)
    qg = g.download()
    err = np.abs(qg - qo).max() / np.abs(q).max()
    assert err <= 1e-12, err
    assert abs(cfl - cflo) <= 1e-13 * cflo


@pytest.mark.parametrize("order,tw", [(1, 0), (1, 1), (2, 0), (2, 1), (2, 2)])
@pytest.mark.parametrize("lim", [(1, 1, 1, 1), (2, 2, 2, 2), (3, 3, 3, 3), (4, 4, 4, 4), (0, 0, 0, 0),
                                 (5, 5, 5, 5), (4, 1, 0, 3)])
def test_method_variants(order, tw, lim):
    mx, my = 61, 33
    q = problems.random_smooth_swe(mx, my, 3)
    dx, dy, dt = 0.1, 0.1, 0.015
    qo, cflo = _oracle_step("swe", q, dx, dy, (1, 1), dt, order=order, tw=tw, lim=lim)
    g = _grid("swe", mx, my, dx, dy, (1, 1), order=order, tw=tw, lim=lim)
    g.init_ctrl(0.0, dt)
    g.upload(q)
    cfl, _ = g.step(dt)
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    assert err <= 1e-12, err


def test_dam_break_100_steps():
    """Workload of the CPU-runnable config: radial dam break, 200x200, MC, CFL 0.9, extrap,
    100 accepted steps with the device-side dt controller (first attempt dt=0.1 is rejected)."""
    q, dx, dy = problems.radial_dam_break(200, 200)
    d = corc.make_desc("swe", 200, 200, dx, dy, grav=1.0)
    c = corc.make_ctrl()
    qo, st = corc.evolve(d, c, q, nsteps=100)
    g = _grid("swe", 200, 200, dx, dy, (1, 1))
    g.init_ctrl(0.0, 0.1, target_steps=100)
    g.upload(q)
    g.advance(101)
    ctrl = g.sync_cur()
    print("ctrl: numsteps %d attempts %d rejected %d status %d t %.17g" %
          (ctrl.numsteps, ctrl.attempts, ctrl.rejected, ctrl.status, ctrl.t))
    assert ctrl.numsteps == 100 and ctrl.rejected == st["rejected"] == 1
    assert abs(ctrl.t - st["t"]) <= 1e-12 * st["t"]
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    print("dam break max rel err", err)
    assert err <= 1e-12, err
