"""Shallow water over bathymetry (shallow_bathymetry_fwave_2D, f-waves, aux(1) = b) on the GPU
against the oracle: single unsplit and split steps over BCs (aux BCs different from the state's
included), orders, transverse modes, limiters and dry cells; the device time loop and the
PyClaw front end on a sill problem; the lake at rest and depth conservation at full size; FP32.
Tolerances: FP64 1e-12 max|q| per step (1e-11 over hundreds of steps), FP32 2e-6 per step."""
import numpy as np
import pytest

from clawpack_amd import _lib, problems, pyclaw, riemann
from clawpack_amd.device import DeviceGrid, make_desc
from oracle import corc

pytestmark = pytest.mark.gpu

G, TOL = 9.81, 1e-3


def oracle_step(q, aux, dx, dy, dt, bc, abc=None, order=2, tw=2, lim=(4, 4, 4), split=False):
    mx, my = q.shape[1:]
    d = corc.make_desc("bathy", mx, my, dx, dy, grav=G, order=order, transverse_waves=tw,
                       mthlim=lim, dry_tolerance=TOL)
    c = corc.make_ctrl(bc_lower=bc[0], bc_upper=bc[1], aux_bc_lower=None if abc is None else abc[0],
                       aux_bc_upper=None if abc is None else abc[1], dimensional_split=split)
    return corc.step_hyperbolic(d, c, q, dt, aux=aux)


def device_grid(q, aux, dx, dy, bc, abc=None, order=2, tw=2, lim=(4, 4, 4), split=False,
                dtype=_lib.FP64, seg_h=0):
    mx, my = q.shape[1:]
    d = make_desc(_lib.SOLVER_SHALLOW_BATHY_FWAVE, dtype, mx, my, dx, dy, 3, 3, order=order,
                  transverse_waves=tw, mthlim=lim, bc_lower=bc[0], bc_upper=bc[1],
                  aux_bc_lower=None if abc is None else abc[0],
                  aux_bc_upper=None if abc is None else abc[1], grav=G, dry_tolerance=TOL,
                  seg_h=seg_h)
    g = DeviceGrid(d, split=split)
    g.upload(q)
    g.upload_aux(aux)
    return g


BCS = [(((1, 1), (1, 1)), None), (((2, 2), (2, 2)), None), (((3, 3), (3, 3)), None),
       (((3, 1), (3, 1)), ((1, 1), (1, 1))), (((1, 2), (3, 2)), ((3, 2), (1, 2)))]


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("bcs", BCS)
@pytest.mark.parametrize("shape", [(37, 29), (130, 71), (61, 200)])
def test_single_step(shape, bcs, split):
    mx, my = shape
    bc, abc = bcs
    print("case bathy step", shape, bcs, split, flush=True)
    q, aux = problems.random_bathymetry(mx, my, seed=mx + my)
    dx, dy = 1.0 / mx, 1.2 / my
    dt = 0.2 * min(dx, dy) / (np.sqrt(G * 1.6) + 1.5)
    qo, cflo = oracle_step(q, aux, dx, dy, dt, bc, abc, split=split)
    g = device_grid(q, aux, dx, dy, bc, abc, split=split)
    g.init_ctrl(0.0, dt)
    cfl, c = g.step(dt)
    assert c.status == 0
    assert np.abs(g.download() - qo).max() <= 1e-12 * np.abs(q).max()
    assert abs(cfl - cflo) <= 1e-12 * cflo


# Beam-warming (limiter 5) is tested without walls: its limited wave is the upwind wave whenever
# the local f-wave is nonzero and 0 when it is exactly zero, so at a wall, where the shear f-wave
# vanishes only up to rounding, the result depends on the last bit of that rounding.  The
# unlimited second-order methods (0 and 5) are not positivity preserving at a wet/dry front (the
# depth goes negative, a bad state), so they run without the dry patch.
@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("order,tw,lim,bc,dry", [
    (1, 0, (0, 0, 0), ((3, 1), (3, 1)), True), (1, 2, (0, 0, 0), ((3, 1), (3, 1)), True),
    (2, 1, (1, 1, 1), ((3, 1), (3, 1)), True), (2, 2, (2, 3, 4), ((3, 1), (3, 1)), True),
    (2, 0, (4, 4, 4), ((3, 1), (3, 1)), True), (2, 1, (2, 2, 2), ((2, 3), (2, 3)), True),
    (2, 2, (2, 3, 5), ((3, 1), (3, 1)), False), (2, 2, (0, 0, 0), ((3, 1), (3, 1)), False),
    (2, 2, (5, 5, 5), ((1, 2), (1, 2)), False),
])
def test_method_variants(order, tw, lim, bc, dry, split):
    mx, my = 90, 66
    print("case bathy method", order, tw, lim, bc, dry, split, flush=True)
    q, aux = problems.random_bathymetry(mx, my, seed=7, dry_patch=dry)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.15 * min(dx, dy) / (np.sqrt(G * 1.6) + 1.5)
    qo, cflo = oracle_step(q, aux, dx, dy, dt, bc, order=order, tw=tw, lim=lim, split=split)
    g = device_grid(q, aux, dx, dy, bc, order=order, tw=tw, lim=lim, split=split, seg_h=9)
    g.init_ctrl(0.0, dt)
    cfl, c = g.step(dt)
    assert c.status == 0
    qg = g.download()
    assert qg[0].min() >= 0.0
    assert np.abs(qg - qo).max() <= 1e-12 * np.abs(q).max()
    assert abs(cfl - cflo) <= 1e-12 * cflo


def test_dam_break_onto_dry_bed():
    mx, my = 120, 40
    h = np.zeros((mx, my))
    h[:30] = 1.0
    q = np.asfortranarray(np.stack([h, 0 * h, 0 * h]))
    aux = np.zeros((1, mx, my), order="F")
    dx, dy = 1.0 / mx, 1.0 / my
    d = corc.make_desc("bathy", mx, my, dx, dy, grav=G, mthlim=(1, 1, 1), dry_tolerance=TOL)
    c = corc.make_ctrl(bc_lower=(3, 3), bc_upper=(3, 3), dt_initial=1e-3)
    qo, st = corc.evolve(d, c, q, nsteps=40, aux=aux)
    g = device_grid(q, aux, dx, dy, ((3, 3), (3, 3)), lim=(1, 1, 1))
    g.init_ctrl(0.0, 1e-3, target_steps=40, max_steps=1000)
    g.advance(200)
    cc = g.sync_cur()
    assert cc.done and cc.numsteps == 40 and cc.rejected == st["rejected"]
    qg = g.download()
    assert np.abs(qg - qo).max() <= 1e-11 * np.abs(q).max()
    assert np.all(qg[:, 80:] == 0.0)


@pytest.mark.parametrize("split", [False, True])
def test_sill_through_front_end(split):
    """PyClaw usage: ClawSolver2D(riemann.shallow_bathymetry_fwave_2D), fwave = True, state.aux[0]
    = bathymetry, aux BCs; 160 x 120 to t = 0.5 against the oracle's evolve_to_time."""
    mx, my = 160, 120
    q0, aux, dx, dy = problems.sill(mx, my)
    solver = pyclaw.ClawSolver2D(riemann.shallow_bathymetry_fwave_2D)
    solver.fwave = True
    solver.dimensional_split = split
    solver.transverse_waves = 2
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.bc_lower = [pyclaw.BC.extrap, pyclaw.BC.wall]
    solver.bc_upper = [pyclaw.BC.extrap, pyclaw.BC.wall]
    solver.aux_bc_lower = [pyclaw.BC.extrap, pyclaw.BC.wall]
    solver.aux_bc_upper = [pyclaw.BC.extrap, pyclaw.BC.wall]
    domain = pyclaw.Domain((-1.0, -1.0), (1.0, 1.0), (mx, my))
    state = pyclaw.State(domain, 3, 1)
    state.problem_data.update(grav=G, dry_tolerance=TOL, sea_level=0.0)
    state.q[...] = q0
    state.aux[...] = aux
    sol = pyclaw.Solution(state, domain)
    solver.evolve_to_time(sol, 0.5)
    d = corc.make_desc("bathy", mx, my, dx, dy, grav=G, mthlim=(4, 4, 4), dry_tolerance=TOL)
    c = corc.make_ctrl(bc_lower=(1, 3), bc_upper=(1, 3), dimensional_split=split)
    qo, st = corc.evolve(d, c, q0, tend=0.5, aux=aux)
    assert solver.status["numsteps"] == st["numsteps"]
    assert abs(sol.t - 0.5) <= 1e-14
    assert np.abs(state.q - qo).max() <= 1e-11 * np.abs(q0).max()
    assert np.abs(state.q[1]).max() > 1e-3  # the hump has moved water


@pytest.mark.parametrize("dtype,tol", [(_lib.FP64, 1e-13), (_lib.FP32, 1e-5)])
@pytest.mark.parametrize("split", [False, True])
def test_lake_at_rest_full_size(dtype, tol, split):
    """1024 x 768, MC, walls, 300 accepted steps of the device time loop: the surface stays flat
    and the water at rest to round-off (absolute, depths ~ 0.4 .. 1.2).  FP32 rounding of the
    hydrostatic balance leaves about 4e-8 of momentum per step (about 2e-6 seen after 300 steps),
    a scheme that is not well balanced about 1e-3."""
    mx, my = 1024, 768
    _, aux = problems.random_bathymetry(mx, my, seed=3)
    h = -aux[0]
    q = np.asfortranarray(np.stack([h, 0 * h, 0 * h]))
    g = device_grid(q, aux, 1.0 / mx, 1.0 / my, ((3, 3), (3, 3)), split=split, dtype=dtype)
    g.init_ctrl(0.0, 1e-4, target_steps=300, max_steps=10000)
    g.advance(320)
    c = g.sync_cur()
    assert c.done and c.numsteps == 300 and c.status == 0
    qg = g.download()
    assert np.abs(qg[0] + aux[0]).max() <= tol
    assert np.abs(qg[1:]).max() <= tol


@pytest.mark.parametrize("dtype,tol", [(_lib.FP64, 1e-13), (_lib.FP32, 1e-6)])
def test_depth_conserved_periodic(dtype, tol):
    mx, my = 700, 500
    q, aux = problems.random_bathymetry(mx, my, seed=12, amp=0.1)
    g = device_grid(q, aux, 2.0 / mx, 2.0 / my, ((2, 2), (2, 2)), dtype=dtype)
    m0 = g.mass_sum(0)
    g.init_ctrl(0.0, 1e-3, target_steps=200, max_steps=10000)
    g.advance(260)
    c = g.sync_cur()
    assert c.done and c.numsteps == 200 and c.status == 0
    assert abs(g.mass_sum(0) - m0) <= tol * abs(m0)


def test_fp32_single_step_parity():
    mx, my = 257, 190
    q, aux = problems.random_bathymetry(mx, my, seed=21)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.2 * min(dx, dy) / (np.sqrt(G * 1.6) + 1.5)
    bc = ((1, 3), (1, 3))
    qo, cflo = oracle_step(q, aux, dx, dy, dt, bc)
    g = device_grid(q, aux, dx, dy, bc, dtype=_lib.FP32)
    g.init_ctrl(0.0, dt)
    cfl, c = g.step(dt)
    err = np.abs(g.download() - qo).max() / np.abs(q).max()
    assert err <= 2e-6, err
    assert abs(cfl - cflo) <= 1e-6 * cflo
