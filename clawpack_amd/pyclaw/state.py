"""State (upstream pyclaw/src/pyclaw/state.py): q(num_eqn, mx, my), problem_data, t."""
import numpy as np

from .geometry import Domain, Patch


class State:
    def __init__(self, geom, num_eqn, num_aux=0):
        if isinstance(geom, Domain):
            self.patch = geom.patch
        elif isinstance(geom, Patch):
            self.patch = geom
        else:
            raise TypeError("State(geom, num_eqn): geom must be a Domain or Patch")
        self.grid = self.patch.grid
        self.problem_data = {}
        self.t = 0.0
        self.index_capa = -1
        self.keep_gauges = False
        shape = [num_eqn] + self.grid.num_cells
        self.q = np.zeros(shape, order="F")
        self.aux = np.zeros([num_aux] + self.grid.num_cells, order="F") if num_aux > 0 else None

    @property
    def num_eqn(self):
        return self.q.shape[0]

    @property
    def num_aux(self):
        return 0 if self.aux is None else self.aux.shape[0]

    @property
    def mp(self):
        return 0

    def is_valid(self):
        return bool(np.isfinite(self.q).all())

    def get_q_global(self):
        return self.q.copy()

    def __copy__(self):
        other = State(self.patch, self.num_eqn, self.num_aux)
        other.q = self.q.copy(order="F")
        other.aux = None if self.aux is None else self.aux.copy(order="F")
        other.t = self.t
        other.problem_data = dict(self.problem_data)
        return other

    def __deepcopy__(self, memo):
        return self.__copy__()
