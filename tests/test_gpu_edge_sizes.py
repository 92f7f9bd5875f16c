"""Edge sizes for the marching-strip kernel against the oracle: the smallest grids (2 x 2, 2 x N,
N x 2), widths around the 60-column strip (59, 60, 61, 120, 121), one-row and few-row segments
(seg_h 1..3), for shallow water (unsplit and split), Euler and the bathymetry solver, FP64, one
step each, over extrap, periodic and wall boundaries."""
import numpy as np
import pytest

from clawpack_amd import _lib, problems
from clawpack_amd.device import DeviceGrid, make_desc
from oracle import corc

pytestmark = pytest.mark.gpu

SHAPES = [(2, 2), (2, 9), (9, 2), (3, 3), (59, 3), (60, 5), (61, 5), (120, 7), (121, 7), (700, 2), (2, 300)]
BCS = [(1, 1), (2, 2), (3, 3)]


def step_both(solver, q, dx, dy, bc, dt, split=False, seg_h=0, aux=None):
    mx, my = q.shape[1:]
    sid, neq = {"swe": (0, 3), "euler": (1, 4), "bathy": (2, 3)}[solver]
    grav = 9.81 if solver == "bathy" else 1.0
    d = corc.make_desc(solver, mx, my, dx, dy, grav=grav, gamma=1.4, dry_tolerance=1e-3)
    c = corc.make_ctrl(bc_lower=bc, bc_upper=bc, dimensional_split=split)
    qo, cflo = corc.step_hyperbolic(d, c, q, dt, aux=aux)
    dd = make_desc(sid, _lib.FP64, mx, my, dx, dy, neq, neq, bc_lower=bc, bc_upper=bc, grav=grav,
                   gamma=1.4, seg_h=seg_h, dry_tolerance=1e-3)
    g = DeviceGrid(dd, split=split)
    g.upload(q)
    if aux is not None:
        g.upload_aux(aux)
    g.init_ctrl(0.0, dt)
    cfl, cc = g.step(dt)
    return g.download(), cfl, qo, cflo


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("shape", SHAPES)
def test_swe_edge_shapes(shape, bc, split):
    mx, my = shape
    print("case edge swe", shape, bc, split, flush=True)
    q = problems.random_smooth_swe(mx, my, seed=mx * 7 + my)
    dx, dy = 1.0 / mx, 1.0 / my
    dt = 0.1 * min(dx, dy)
    qg, cfl, qo, cflo = step_both("swe", q, dx, dy, bc, dt, split)
    assert np.abs(qg - qo).max() <= 1e-12 * np.abs(q).max()
    assert abs(cfl - cflo) <= 1e-12 * cflo


@pytest.mark.parametrize("seg_h", [1, 2, 3])
@pytest.mark.parametrize("split", [False, True])
def test_few_row_segments(seg_h, split):
    mx, my = 75, 13
    print("case edge seg_h", seg_h, split, flush=True)
    q = problems.random_smooth_swe(mx, my, seed=seg_h)
    dt = 0.1 / mx
    qg, cfl, qo, cflo = step_both("swe", q, 1.0 / mx, 1.0 / my, (3, 3), dt, split, seg_h=seg_h)
    assert np.abs(qg - qo).max() <= 1e-12 * np.abs(q).max()


@pytest.mark.parametrize("shape", [(2, 2), (3, 7), (61, 4), (121, 3)])
def test_euler_and_bathymetry_edge_shapes(shape):
    mx, my = shape
    print("case edge euler/bathy", shape, flush=True)
    q = problems.random_euler(mx, my, seed=mx + my)
    dt = 0.05 * min(1.0 / mx, 1.0 / my)
    qg, cfl, qo, cflo = step_both("euler", q, 1.0 / mx, 1.0 / my, (1, 1), dt)
    assert np.abs(qg - qo).max() <= 1e-12 * np.abs(q).max()
    qb, aux = problems.random_bathymetry(mx, my, seed=mx * my)
    dt = 0.05 * min(1.0 / mx, 1.0 / my)
    qg, cfl, qo, cflo = step_both("bathy", qb, 1.0 / mx, 1.0 / my, (3, 3), dt, aux=aux)
    assert np.abs(qg - qo).max() <= 1e-12 * np.abs(qb).max()
