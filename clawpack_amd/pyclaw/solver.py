"""Solver base (upstream pyclaw/src/pyclaw/solver.py): BC codes, CFL cache, defaults.

The time loop of evolve_to_time lives on the device (ClawSolver2D); this module keeps
pyclaw's attribute names, defaults and error behaviour.
"""


class BC:
    """Boundary-condition codes of pyclaw (solver.BC)."""
    custom = 0
    extrap = 1
    periodic = 2
    wall = 3


class CFL:
    """pyclaw.cfl.CFL: holds the CFL number of the last step."""

    def __init__(self, global_max=0.0):
        self._global_max = global_max
        self._local_max = global_max

    def get_global_max(self):
        return self._global_max

    def get_cached_max(self):
        return self._global_max

    def set_local_max(self, new_local_max):
        self._local_max = new_local_max

    def update_global_max(self, new_local_max):
        self._global_max = new_local_max


class Solver:
    """Attributes and defaults shared by pyclaw solvers."""

    def __init__(self, riemann_solver=None, claw_package=None):
        self.num_ghost = 2
        self.dt_initial = 0.1
        self.dt_max = 1e99
        self.max_steps = 10000
        self.dt_variable = True
        self.cfl_max = 1.0
        self.cfl_desired = 0.9
        self.bc_lower = [None, None]
        self.bc_upper = [None, None]
        self.aux_bc_lower = [None, None]
        self.aux_bc_upper = [None, None]
        self.user_bc_lower = None
        self.user_bc_upper = None
        self.before_step = None
        self.step_source = None
        self.source_split = 1
        self.verbosity = 0
        self.logger = None
        self.dt = None
        self.cfl = CFL(0.0)
        self.status = {"cflmax": 0.0, "dtmin": None, "dtmax": None, "numsteps": 0}
        self._is_set_up = False
        self.rp = None
        self.num_eqn = None
        self.num_waves = None
        if riemann_solver is not None:
            self.rp = riemann_solver
            name = riemann_solver.__name__.split(".")[-1]
            from ..riemann import static
            self.num_eqn = static.num_eqn.get(name, getattr(riemann_solver, "num_eqn", None))
            self.num_waves = static.num_waves.get(name, getattr(riemann_solver, "num_waves", None))

    def _check_cfl_settings(self):
        if not self.dt_variable and self.cfl_max < self.cfl_desired:
            raise Exception("Fixed time step with cfl_desired > cfl_max")
        if self.dt_variable and self.cfl_desired > self.cfl_max:
            raise Exception("Variable time-stepping and desired CFL > maximum CFL")

    def get_dt_new(self):
        cfl = self.cfl.get_cached_max()
        return min(self.dt_max, self.dt * self.cfl_desired / cfl) if cfl > 0 else self.dt_max

    def accept_reject_step(self, state):
        return not (self.cfl.get_cached_max() > self.cfl_max)
