"""Writes the golden fixtures of this directory from the CPU FP64 oracle.

The reference tree holds no test vectors for this path (its pyclaw/riemann submodules are empty),
so these fixtures are the oracle's own output on the workload shapes, pinned by the invariant
tests in tests/test_oracle_*.py.  They freeze the oracle (CPU regression) and give the GPU tests
an on-disk target that does not need the oracle at run time.

    python tests/golden/make_fixtures.py            # writes the fixtures that do not exist yet
    python tests/golden/make_fixtures.py --all      # rewrites every fixture
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from clawpack_amd import problems  # noqa: E402
from oracle import corc  # noqa: E402

EULER_SEED = 1  # screened for positivity on the oracle (seed 5 goes negative, for example)

CASES = {
    # radial dam break on the pyclaw example's domain (coarser), MC, extrap, dt_initial 0.1
    "swe_dam_break_64x64_20steps": dict(
        solver="swe", make=lambda: problems.radial_dam_break(64, 64), bc=(1, 1), steps=20),
    # Euler four-quadrant Riemann problem, MC, extrap
    "euler_quadrants_48x48_20steps": dict(
        solver="euler", make=lambda: problems.euler_quadrants(48, 48, EULER_SEED), bc=(1, 1),
        steps=20),
    # periodic Gaussian humps (the precision workload's initial data), non-square
    "swe_humps_periodic_64x48_15steps": dict(
        solver="swe", make=lambda: problems.gaussian_humps(64, 48, 7), bc=(2, 2), steps=15),
    # walls on all sides
    "swe_humps_wall_40x40_15steps": dict(
        solver="swe", make=lambda: problems.gaussian_humps(40, 40, 3), bc=(3, 3), steps=15),
    # shallow water over a sill (f-wave solver, aux plane b), g = 9.81, x extrap / y wall, unsplit
    # with transverse_waves 2 and the dimensionally split step
    "swe_bathy_sill_64x48_20steps": dict(
        solver="bathy", make=lambda: problems.sill(64, 48), bc=((1, 3), (1, 3)), steps=20,
        grav=9.81),
    "swe_bathy_sill_split_64x48_20steps": dict(
        solver="bathy", make=lambda: problems.sill(64, 48), bc=((1, 3), (1, 3)), steps=20,
        grav=9.81, split=True),
}

DRY_TOLERANCE = 1e-3


def bcs(c):
    """(bc_lower, bc_upper) of a case ("bc" is one pair used for both sides, or a pair of pairs)."""
    bc = c["bc"]
    return (tuple(bc[0]), tuple(bc[1])) if isinstance(bc[0], tuple) else (tuple(bc), tuple(bc))


def run_case(name):
    """(q0, q after the case's accepted steps, status, dx, dy, aux or None)."""
    c = CASES[name]
    made = c["make"]()
    if len(made) == 4:
        q, aux, dx, dy = made
    else:
        (q, dx, dy), aux = made, None
    mx, my = q.shape[1:]
    lo, hi = bcs(c)
    d = corc.make_desc(c["solver"], mx, my, dx, dy, grav=c.get("grav", 1.0), gamma=1.4,
                       dry_tolerance=DRY_TOLERANCE)
    ctrl = corc.make_ctrl(bc_lower=lo, bc_upper=hi, dimensional_split=c.get("split", False))
    qn, st = corc.evolve(d, ctrl, q, nsteps=c["steps"], aux=aux)
    return q, qn, st, dx, dy, aux


def main():
    rewrite = "--all" in sys.argv[1:]
    for name, c in CASES.items():
        path = os.path.join(HERE, name + ".npz")
        if os.path.exists(path) and not rewrite:
            continue
        q0, qn, st, dx, dy, aux = run_case(name)
        extra = {} if aux is None else dict(aux=aux, grav=c["grav"], dry_tolerance=DRY_TOLERANCE,
                                            split=bool(c.get("split", False)))
        np.savez_compressed(path, q0=q0, q=qn, t=st["t"], dt=st["dt"], cflmax=st["cflmax"],
                            numsteps=st["numsteps"], rejected=st["rejected"], dx=dx, dy=dy,
                            bc=np.array(c["bc"]), solver=c["solver"], **extra)
        print(name, os.path.getsize(path), "bytes", st)


if __name__ == "__main__":
    main()
