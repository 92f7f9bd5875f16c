"""PyClaw-compatible front end of the HIP 2D classic solver (upstream package clawpack.pyclaw).

    from clawpack_amd import pyclaw, riemann
    solver = pyclaw.ClawSolver2D(riemann.shallow_roe_with_efix_2D)
    solver.dimensional_split = False
    solver.transverse_waves = 2
    solver.limiters = pyclaw.limiters.tvd.MC
    solver.bc_lower = [pyclaw.BC.extrap, pyclaw.BC.extrap]
    solver.bc_upper = [pyclaw.BC.extrap, pyclaw.BC.extrap]
    domain = pyclaw.Domain([pyclaw.Dimension(-2.5, 2.5, 200, name='x'),
                            pyclaw.Dimension(-2.5, 2.5, 200, name='y')])
    state = pyclaw.State(domain, solver.num_eqn)
    state.problem_data['grav'] = 1.0
    ...
    claw = pyclaw.Controller(); claw.solution = pyclaw.Solution(state, domain); claw.solver = solver
    claw.run()
"""
from . import limiters  # noqa: F401
from .classic import ClawSolver2D  # noqa: F401
from .controller import Controller  # noqa: F401
from .geometry import Dimension, Domain, Grid, Patch  # noqa: F401
from .solution import Solution  # noqa: F401
from .solver import BC, CFL, Solver  # noqa: F401
from .state import State  # noqa: F401
