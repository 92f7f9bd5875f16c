"""Controller (upstream pyclaw/src/pyclaw/controller.py): output-time loop around
solver.evolve_to_time.  Frames are kept in memory (keep_copy); writing frame files is not
part of this package."""
import copy

import numpy as np


class Controller:
    def __init__(self):
        self.solution = None
        self.solver = None
        self.tfinal = 1.0
        self.num_output_times = 10
        self.output_style = 1
        self.out_times = None
        self.keep_copy = False
        self.output_format = None
        self.outdir = "./_output"
        self.frames = []
        self.verbosity = 0

    def _times(self):
        t0 = self.solution.t
        if self.output_style == 1:
            return np.linspace(t0, self.tfinal, self.num_output_times + 1)
        if self.output_style == 2:
            return np.asarray(self.out_times, dtype=float)
        raise NotImplementedError("output_style %r" % self.output_style)

    def run(self):
        if self.solution is None or self.solver is None:
            raise Exception("Controller.run needs solution and solver")
        if self.output_format is not None:
            raise NotImplementedError("frame output is not part of clawpack_amd; set output_format=None")
        self.solver.setup(self.solution)
        if self.keep_copy:
            self.frames.append(copy.deepcopy(self.solution))
        totals = {"cflmax": 0.0, "dtmin": None, "dtmax": None, "numsteps": 0}
        for t in self._times()[1:]:
            st = self.solver.evolve_to_time(self.solution, t)
            totals["cflmax"] = max(totals["cflmax"], st["cflmax"])
            totals["dtmin"] = st["dtmin"] if totals["dtmin"] is None else min(totals["dtmin"], st["dtmin"])
            totals["dtmax"] = st["dtmax"] if totals["dtmax"] is None else max(totals["dtmax"], st["dtmax"])
            totals["numsteps"] += st["numsteps"]
            if self.keep_copy:
                self.frames.append(copy.deepcopy(self.solution))
        return totals
