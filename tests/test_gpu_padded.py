"""The padded entry points against the oracle from ghost layers that no boundary rule would make.

claw2d_step_padded (unsplit step2) and claw2d_sweep_padded (the two sweeps of step2ds) read their
ghost cells from the caller.  Here qbc is one random state on the padded grid (its ghost cells
continue the interior smoothly, which no boundary rule does) and the ghost ring of q* is a second,
independent random state, so a kernel that re-derives a ghost value from the interior, or reads a
ghost cell from the wrong place, fails.
Compared with corc.step2 / corc.step2ds on the same padded arrays:

  * sweep 1 writes q* on the interior of its padded output and leaves the ghost ring untouched;
  * sweep 2 from the oracle's own q* isolates the y-sweep; the chain through the device's q*
    checks the pair;
  * ctrl.cfl after the pair is max(cfl_x, cfl_y) of the oracle's two sweeps.

FP64 to 1e-12 * max|q|, FP32 to 1e-5 * max|q| (single sweeps of an FP32 state)."""
import numpy as np
import pytest
import torch

from clawpack_amd import _lib, problems
from clawpack_amd.device import DeviceGrid, make_desc
from oracle import corc

pytestmark = pytest.mark.gpu

IDS = {"swe": _lib.SOLVER_SHALLOW_ROE_EFIX, "euler": _lib.SOLVER_EULER_4WAVE,
       "bathy": _lib.SOLVER_SHALLOW_BATHY_FWAVE}


def padded_state(solver, mx, my, seed):
    """A positive-depth / positive-pressure state on the padded grid (num_eqn, mx+4, my+4)."""
    if solver == "euler":
        return np.asfortranarray(problems.random_euler(mx + 4, my + 4, seed))
    return np.asfortranarray(problems.random_smooth_swe(mx + 4, my + 4, seed))


def to_dev(a, dt):
    """pyclaw layout (neq, mx+4, my+4) -> device planes (neq, my+4, mx+4)."""
    return torch.from_numpy(np.ascontiguousarray(a.transpose(0, 2, 1))).to("cuda:0", dt)


def to_host(t):
    """device planes (neq, ny, nx) -> pyclaw layout (neq, nx, ny), FP64."""
    return np.asfortranarray(t.double().cpu().numpy().transpose(0, 2, 1))


def setup(solver, mx, my, lim, order, seed, fp32=False):
    dx, dy = 1.0 / mx, 1.3 / my
    neq = 4 if solver == "euler" else 3
    grav = 9.81 if solver == "bathy" else 1.0
    aux = auxbc = None
    qbc = padded_state(solver, mx, my, seed)
    ghosts = padded_state(solver, mx, my, seed + 100)   # the ghost ring q* is given
    abl, abu = (3, 1), (1, 3)
    if solver == "bathy":
        # a smooth bathymetry b in [-1, -0.6] under the (wet) random depth
        aux = np.asfortranarray(-0.8 + 0.2 * np.sin(6 * np.linspace(0, 1, mx))[None, :, None]
                                * np.cos(5 * np.linspace(0, 1, my))[None, None, :])
    od = corc.make_desc(solver, mx, my, dx, dy, grav=grav, order=order, mthlim=tuple(lim))
    if solver == "bathy":
        oc = corc.make_ctrl(aux_bc_lower=abl, aux_bc_upper=abu)
        auxbc = corc.fill_aux_ghosts(od, oc, aux)
    desc = make_desc(IDS[solver], _lib.FP32 if fp32 else _lib.FP64, mx, my, dx, dy, neq, neq,
                     order=order, transverse_waves=2, mthlim=tuple(lim), grav=grav,
                     aux_bc_lower=abl, aux_bc_upper=abu)
    g = DeviceGrid(desc, "cuda:0")
    if aux is not None:
        g.upload_aux(aux)
    return od, g, qbc, ghosts, auxbc


def sweep_dt(od, qbc, ghosts, auxbc):
    """dt for CFL 0.8 over both sweeps (CFL numbers are linear in dt; the y-sweep of qbc and of
    the ghost state bound the y-sweep of q*)."""
    c = max(corc.step2ds(od, q, 1.0, ids, auxbc=auxbc)[1] for q in (qbc, ghosts) for ids in (1, 2))
    return 0.8 / c


CASES = [("swe", 64, 48, (4, 4, 4), 2), ("swe", 61, 5, (1, 1, 1), 2), ("swe", 123, 77, (2, 3, 5), 2),
         ("swe", 37, 29, (4, 4, 4), 1), ("euler", 70, 66, (4, 4, 4, 4), 2),
         ("euler", 59, 130, (3, 1, 2, 4), 2), ("bathy", 90, 70, (4, 4, 4), 2),
         ("bathy", 45, 121, (1, 1, 1), 2), ("swe", 2, 2, (4, 4, 4), 2), ("swe", 121, 3, (4, 4, 4), 2)]


@pytest.mark.parametrize("solver,mx,my,lim,order", CASES)
def test_sweeps_from_arbitrary_ghosts_match_step2ds(solver, mx, my, lim, order):
    print("case sweep_padded", solver, mx, my, lim, order, flush=True)
    od, g, qbc, ghosts, auxbc = setup(solver, mx, my, lim, order, seed=mx + 7 * my)
    dt = sweep_dt(od, qbc, ghosts, auxbc)
    qs_o, cx = corc.step2ds(od, qbc, dt, 1, auxbc=auxbc)
    qs_in = ghosts.copy(order="F")
    qs_in[:, 2:-2, 2:-2] = qs_o[:, 2:-2, 2:-2]
    qn_o, cy = corc.step2ds(od, qs_in, dt, 2, auxbc=auxbc)
    scale = np.abs(qbc).max()

    g.init_ctrl(0.0, dt, target_steps=1)
    qs_d = to_dev(ghosts, torch.float64)        # the ghost ring q* will see; the interior is overwritten
    out = torch.empty_like(g.q[0])
    g.sweep_padded(to_dev(qbc, torch.float64), qs_d, dt, 1)
    g.sweep_padded(to_dev(qs_in, torch.float64), out, dt, 2)  # y-sweep from the oracle's q*
    c = g.read_ctrl()
    qs_h = to_host(qs_d)
    ring = np.ones(qs_h.shape[1:], bool)
    ring[2:-2, 2:-2] = False
    assert np.array_equal(qs_h[:, ring], ghosts[:, ring]), "sweep 1 wrote a ghost cell"
    assert np.abs(qs_h[:, 2:-2, 2:-2] - qs_o[:, 2:-2, 2:-2]).max() <= 1e-12 * scale
    assert np.abs(to_host(out) - qn_o[:, 2:-2, 2:-2]).max() <= 1e-12 * scale
    assert c.status == 0
    assert abs(c.cfl - max(cx, cy)) <= 1e-12 * max(cx, cy)
    assert abs(c.smax_last[0] * dt / od.dx - cx) <= 1e-12 * cx
    assert abs(c.smax_last[1] * dt / od.dy - cy) <= 1e-12 * max(cy, 1e-300)

    # the pair chained through the device's own q*, twice (the control block closes each pair)
    for _ in range(2):
        g.sweep_padded(to_dev(qbc, torch.float64), qs_d, dt, 1)
        g.sweep_padded(qs_d, out, dt, 2)
        c = g.read_ctrl()
        assert np.abs(to_host(out) - qn_o[:, 2:-2, 2:-2]).max() <= 1e-12 * scale
        assert abs(c.cfl - max(cx, cy)) <= 1e-12 * max(cx, cy) and c.arrive == 0


@pytest.mark.parametrize("solver,mx,my,lim,order", CASES[:3] + CASES[4:7])
@pytest.mark.parametrize("tw", [2, 1])
def test_step_padded_from_arbitrary_ghosts_matches_step2(solver, mx, my, lim, order, tw):
    print("case step_padded", solver, mx, my, lim, order, tw, flush=True)
    od, g, qbc, _, auxbc = setup(solver, mx, my, lim, order, seed=3 * mx + my)
    od.transverse_waves = tw
    g.desc.transverse_waves = tw
    dt = 0.8 / corc.step2(od, qbc, 1.0, auxbc=auxbc)[1]
    qn_o, cfl = corc.step2(od, qbc, dt, auxbc=auxbc)
    g.init_ctrl(0.0, dt, target_steps=1)
    out = torch.empty_like(g.q[0])
    g.step_padded(to_dev(qbc, torch.float64), out, dt)
    c = g.read_ctrl()
    assert np.abs(to_host(out) - qn_o[:, 2:-2, 2:-2]).max() <= 1e-12 * np.abs(qbc).max()
    assert c.status == 0 and abs(c.cfl - cfl) <= 1e-12 * cfl


@pytest.mark.parametrize("solver,mx,my,lim", [("swe", 96, 80, (4, 4, 4)), ("euler", 77, 65, (4, 4, 4, 4)),
                                              ("bathy", 64, 90, (4, 4, 4))])
def test_fp32_sweeps_and_step_from_arbitrary_ghosts(solver, mx, my, lim):
    print("case padded fp32", solver, mx, my, flush=True)
    od, g, qbc, ghosts, auxbc = setup(solver, mx, my, lim, 2, seed=mx + my, fp32=True)
    dt = sweep_dt(od, qbc, ghosts, auxbc)
    qs_o, cx = corc.step2ds(od, qbc, dt, 1, auxbc=auxbc)
    qs_in = ghosts.copy(order="F")
    qs_in[:, 2:-2, 2:-2] = qs_o[:, 2:-2, 2:-2]
    qn_o, cy = corc.step2ds(od, qs_in, dt, 2, auxbc=auxbc)
    scale = np.abs(qbc).max()
    g.init_ctrl(0.0, dt, target_steps=1)
    qs_d = to_dev(ghosts, torch.float32)
    out = torch.empty_like(g.q[0])
    g.sweep_padded(to_dev(qbc, torch.float32), qs_d, dt, 1)
    g.sweep_padded(qs_d, out, dt, 2)
    c = g.read_ctrl()
    assert np.abs(to_host(out) - qn_o[:, 2:-2, 2:-2]).max() <= 1e-5 * scale
    assert abs(c.cfl - max(cx, cy)) <= 1e-5 * max(cx, cy)
    qn_u, cu = corc.step2(od, qbc, dt, auxbc=auxbc)
    g.init_ctrl(0.0, dt, target_steps=1)
    g.step_padded(to_dev(qbc, torch.float32), out, dt)
    c = g.read_ctrl()
    assert np.abs(to_host(out) - qn_u[:, 2:-2, 2:-2]).max() <= 1e-5 * scale
    assert abs(c.cfl - cu) <= 1e-5 * cu
