"""Segment sizing changes launch shapes only, never results.

The library picks the rows per marching segment from the grid size, the SIMD count and the
kernel's occupancy: short segments for half the SIMDs on small grids, at most 2 wavefronts per
SIMD when segments would be under 32 rows, and no launch just over a whole number of wavefronts
per SIMD.  Every interface is solved by its owner with the same arithmetic wherever segments
start and end, so the default launch must match explicit segment heights bit for bit, through
several controller-driven steps (state, t, dt, step counts).  Sizes are chosen so that each rule
applies: 256^2 (small grid), 1024^2 (1152 -> 972 wavefronts), 2048^2 FP32 and split (4 -> 2 per
SIMD), and a non-square grid.
"""
import numpy as np
import pytest

from clawpack_amd import _lib, problems
from clawpack_amd.device import DeviceGrid, make_desc

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def run(q, dx, dy, dtype, split, seg_h, steps=4):
    mx, my = q.shape[1], q.shape[2]  # pyclaw layout (num_eqn, mx, my)
    d = make_desc(0, dtype, mx, my, dx, dy, 3, 3, order=2, transverse_waves=2, mthlim=(4, 4, 4),
                  bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0, seg_h=seg_h)
    g = DeviceGrid(d, DEV, split=split)
    g.upload(q)
    g.init_ctrl(0.0, 1e-4, target_steps=10 ** 9)
    g.advance(steps)
    c = g.read_ctrl()
    assert c.status == 0
    out = g.download()
    return out, (c.t, c.dt, c.numsteps, c.attempts)


@pytest.mark.parametrize("mx,my,dtype,split,explicit", [
    (256, 256, _lib.FP64, False, (16, 7)),
    (1024, 1024, _lib.FP64, False, (16, 24)),
    (1024, 1024, _lib.FP32, False, (16,)),
    (2048, 2048, _lib.FP32, False, (18,)),
    (2048, 2048, _lib.FP32, True, (18,)),
    (2048, 2048, _lib.FP64, True, (24,)),
    (700, 300, _lib.FP64, False, (16, 5)),
])
def test_default_sizing_matches_explicit_segments(mx, my, dtype, split, explicit):
    print("case seg sizing", mx, my, dtype, split, explicit, flush=True)
    q, dx, dy = problems.gaussian_humps(mx, my, 11)
    ref, cref = run(q, dx, dy, dtype, split, 0)
    assert np.isfinite(ref).all()
    for seg in explicit:
        got, cgot = run(q, dx, dy, dtype, split, seg)
        assert cgot == cref, (seg, cgot, cref)
        assert np.array_equal(got, ref), seg
