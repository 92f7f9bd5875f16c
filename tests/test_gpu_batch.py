"""Batched grids: one persistent kernel per attempt over every grid, each grid with its own dt.

Bars: batched == one-grid-at-a-time launches bit for bit (same arithmetic, other work split);
batched vs the CPU oracle per instance <= 1e-12 * max|q|; HIP-graph replays == eager launches.
"""
import numpy as np
import pytest
import torch

from clawpack_amd import problems
from oracle import corc

pytestmark = pytest.mark.gpu


def _batch(ngrids, dtype, seed=5, nmin=40, nmax=160, seg_h=0):
    from clawpack_amd.batch import BatchedGrids
    from clawpack_amd.device import make_desc
    grids = problems.batch_workload(ngrids, seed, nmin=nmin, nmax=nmax)
    d = make_desc(0, dtype, 8, 8, 1.0, 1.0, 3, 3, bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0)
    return grids, BatchedGrids(d, grids, seg_h=seg_h)


@pytest.mark.parametrize("dtype", [0, 1])
def test_batched_equals_sequential(dtype):
    print("case batched vs sequential dtype", dtype, flush=True)
    grids, b1 = _batch(24, dtype)
    _, b2 = _batch(24, dtype)
    steps = 12
    for b in (b1, b2):
        b.init(0.0, 0.1, target_steps=steps)
    torch.cuda.synchronize()
    print("  init done", flush=True)
    b1.advance(steps + 3)
    torch.cuda.synchronize()
    print("  batched advance done", flush=True)
    b2.advance_sequential(steps + 3)
    torch.cuda.synchronize()
    print("  sequential advance done", flush=True)
    c1, c2 = b1.read_ctrls(), b2.read_ctrls()
    for g in range(len(grids)):
        assert c1[g].numsteps == c2[g].numsteps == steps
        assert c1[g].t == c2[g].t and c1[g].dt == c2[g].dt
        assert np.array_equal(b1.download(g, c1), b2.download(g, c2)), g


def test_batched_matches_oracle_per_instance():
    print("case batched vs oracle", flush=True)
    grids, b = _batch(10, 0, seed=9, nmin=24, nmax=90, seg_h=16)
    steps = 10
    b.init(0.0, 0.1, target_steps=steps)
    b.advance(steps + 3)
    torch.cuda.synchronize()
    print("  batched advance done", flush=True)
    ctrls = b.read_ctrls()
    for g, (q, dx, dy) in enumerate(grids):
        mx, my = q.shape[1:]
        d = corc.make_desc("swe", mx, my, dx, dy, grav=1.0)
        c = corc.make_ctrl(bc_lower=(2, 2), bc_upper=(2, 2))
        qo, st = corc.evolve(d, c, q, nsteps=steps)
        assert ctrls[g].numsteps == steps and ctrls[g].rejected == st["rejected"]
        assert abs(ctrls[g].t - st["t"]) <= 1e-12 * st["t"]
        err = np.abs(b.download(g, ctrls) - qo).max() / np.abs(q).max()
        assert err <= 1e-12, (g, err)


def test_graph_replays_match_eager():
    print("case batch graph replay", flush=True)
    grids, eager = _batch(16, 0, seed=3)
    _, graphed = _batch(16, 0, seed=3)
    per_graph, replays = 4, 3
    for b in (eager, graphed):
        b.init(0.0, 0.1, target_steps=10 ** 6)
    h = graphed.make_graph(per_graph)
    print("  graph captured", flush=True)
    for r in range(replays):
        eager.advance(per_graph)
        graphed.replay(h)
        torch.cuda.synchronize()
        print("  replay", r, "done", flush=True)
        # eager work of the program between replays
        _ = (torch.ones(1000, device="cuda") * 2).sum().item()
        ce, cg = eager.read_ctrls(), graphed.read_ctrls()
        for g in range(len(grids)):
            assert ce[g].attempts == cg[g].attempts == per_graph * (r + 1)
            assert ce[g].t == cg[g].t
            assert np.array_equal(eager.download(g, ce), graphed.download(g, cg)), (r, g)


def _gentle_quadrants(mx, my, seed, gamma=1.4):
    """A four-quadrant Euler problem on [0,1]^2 with rho, p ~ U[0.5, 1.5] and u, v ~ U[-0.3, 0.3]
    per quadrant: it stays positive next to walls (the faster quadrants of the bench workload
    empty a wall's neighbourhood into near vacuum, where the Roe solver fails)."""
    rng = np.random.default_rng(seed)
    rho, p = rng.uniform(0.5, 1.5, 4), rng.uniform(0.5, 1.5, 4)
    u, v = rng.uniform(-0.3, 0.3, 4), rng.uniform(-0.3, 0.3, 4)
    xc, dx = problems.centers(0.0, 1.0, mx)
    yc, dy = problems.centers(0.0, 1.0, my)
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    k = (X >= 0.5).astype(int) + 2 * (Y >= 0.5).astype(int)
    q = np.zeros((4, mx, my), order="F")
    q[0], q[1], q[2] = rho[k], rho[k] * u[k], rho[k] * v[k]
    q[3] = p[k] / (gamma - 1.0) + 0.5 * rho[k] * (u[k] ** 2 + v[k] ** 2)
    return q, dx, dy


def _batch_method(solver, dtype, bc_lower, bc_upper, mthlim, ngrids=10, seed=11):
    """Grids of uneven sizes for one method: gentle Euler quadrants or 16-hump shallow water,
    with the given BCs and per-wave limiters."""
    from clawpack_amd.batch import BatchedGrids
    from clawpack_amd.device import make_desc
    rng = np.random.default_rng(seed)
    sizes = rng.integers(24, 150, size=(ngrids, 2))
    if solver == "euler":
        grids = [_gentle_quadrants(int(mx), int(my), 100 + g) for g, (mx, my) in enumerate(sizes)]
        sid, neq = 1, 4
    else:
        grids = [problems.gaussian_humps(int(mx), int(my), seed + g)
                 for g, (mx, my) in enumerate(sizes)]
        sid, neq = 0, 3
    d = make_desc(sid, dtype, 8, 8, 1.0, 1.0, neq, neq, bc_lower=bc_lower, bc_upper=bc_upper,
                  grav=1.0, gamma=1.4, mthlim=mthlim)
    return grids, (lambda: BatchedGrids(d, grids))


# (solver, bc_lower, bc_upper, mthlim): Euler with the specialised MC kernel under extrap/wall
# sides, and the run-time-limiter kernel (mixed per-wave limiters) for both solvers with walls
BATCH_METHODS = [
    ("euler", (1, 3), (3, 1), (4, 4, 4, 4)),
    ("euler", (3, 3), (3, 3), (1, 4, 2, 3)),
    ("swe", (3, 1), (1, 3), (2, 4, 1)),
]


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("solver,bcl,bcu,lim", BATCH_METHODS)
def test_batched_methods_equal_sequential_and_oracle(solver, bcl, bcu, lim, dtype):
    """Batched Euler and run-time-limiter kernels with non-periodic BCs: bit for bit equal to one
    launch per grid, and per grid equal to the oracle (FP64 to 1e-12 max|q|; FP32 to 1e-4, the
    single-precision rounding over 10 steps of shocks)."""
    print("case batch", solver, bcl, bcu, lim, "dtype", dtype, flush=True)
    grids, make = _batch_method(solver, dtype, bcl, bcu, lim)
    b1, b2 = make(), make()
    steps = 10
    for b in (b1, b2):
        b.init(0.0, 0.1, target_steps=steps)
    b1.advance(steps + 3)
    b2.advance_sequential(steps + 3)
    torch.cuda.synchronize()
    print("  advances done", flush=True)
    c1, c2 = b1.read_ctrls(), b2.read_ctrls()
    for g, (q, dx, dy) in enumerate(grids):
        assert c1[g].numsteps == c2[g].numsteps == steps
        assert c1[g].t == c2[g].t and c1[g].dt == c2[g].dt
        qb = b1.download(g, c1)
        assert np.array_equal(qb, b2.download(g, c2)), g
        mx, my = q.shape[1:]
        d = corc.make_desc(solver, mx, my, dx, dy, grav=1.0, gamma=1.4, mthlim=lim)
        c = corc.make_ctrl(bc_lower=bcl, bc_upper=bcu)
        qo, st = corc.evolve(d, c, q, nsteps=steps)
        assert c1[g].rejected == st["rejected"]
        err = np.abs(qb - qo).max() / np.abs(q).max()
        assert err <= (1e-12 if dtype == 0 else 1e-4), (g, err)


def test_fp32_batch_with_stored_depth_deviation():
    """FP32 batches store the depth as h - depth_ref as single grids do: batched equals one grid at
    a time bit for bit, and both stay within FP32 rounding of the FP64 batch (per grid, relative to
    the grid's own max|q|), closer than the same batch storing h."""
    from clawpack_amd.batch import BatchedGrids
    from clawpack_amd.device import make_desc
    print("case fp32 batch depth_ref", flush=True)
    grids = problems.batch_workload(12, 21, nmin=40, nmax=140)
    steps = 30
    runs = {}
    for name, dtype, ref, seq in (("fp64", 0, 0.0, False), ("ref", 1, 1.0, False), ("ref_seq", 1, 1.0, True),
                                  ("h", 1, 0.0, False)):
        d = make_desc(0, dtype, 8, 8, 1.0, 1.0, 3, 3, bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0,
                      depth_ref=ref, dt_variable=False)
        b = BatchedGrids(d, grids)
        b.init(0.0, 2e-3, target_steps=steps)
        if seq:
            b.advance_sequential(steps + 2)
        else:
            b.advance(steps + 2)
        torch.cuda.synchronize()
        c = b.read_ctrls()
        assert all(x.numsteps == steps and x.status == 0 for x in c)
        runs[name] = [b.download(g, c) for g in range(len(grids))]
    for g in range(len(grids)):
        assert np.array_equal(runs["ref"][g], runs["ref_seq"][g]), g
    err = lambda k: max(np.abs(runs[k][g][1:] - runs["fp64"][g][1:]).max() / np.abs(runs["fp64"][g]).max()
                        for g in range(len(grids)))
    e_ref, e_h = err("ref"), err("h")
    print("  momentum error vs fp64: stored h - 1 %.2e, stored h %.2e" % (e_ref, e_h), flush=True)
    assert e_ref <= 1e-6 and e_ref < e_h
