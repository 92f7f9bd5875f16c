"""Controller (upstream pyclaw/src/pyclaw/controller.py): output-time loop around
solver.evolve_to_time.  Frames are written to outdir in the ascii format (output_format='ascii',
the default), the binary format ('binary'), or both (a list); None writes nothing.  Frames can
also be kept in memory (keep_copy)."""
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
        self.output_format = "ascii"
        self.output_file_prefix = None
        self.output_options = {}
        self.write_aux_init = False
        self.write_aux_always = False
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
        self.solver.setup(self.solution)
        frame = 0
        self._output(frame, self.write_aux_init or self.write_aux_always)
        totals = {"cflmax": 0.0, "dtmin": None, "dtmax": None, "numsteps": 0}
        for t in self._times()[1:]:
            st = self.solver.evolve_to_time(self.solution, t)
            totals["cflmax"] = max(totals["cflmax"], st["cflmax"])
            totals["dtmin"] = st["dtmin"] if totals["dtmin"] is None else min(totals["dtmin"], st["dtmin"])
            totals["dtmax"] = st["dtmax"] if totals["dtmax"] is None else max(totals["dtmax"], st["dtmax"])
            totals["numsteps"] += st["numsteps"]
            frame += 1
            self._output(frame, self.write_aux_always)
        return totals

    def _output(self, frame, write_aux):
        if self.output_format is not None:
            self.solution.write(frame, self.outdir, self.output_format, self.output_file_prefix,
                                write_aux, self.output_options)
        if self.keep_copy:
            self.frames.append(copy.deepcopy(self.solution))
