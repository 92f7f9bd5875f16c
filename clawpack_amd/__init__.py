"""clawpack_amd: MI355X-native (gfx950 HIP) 2D unsplit wave-propagation step with a
PyClaw-compatible front end.

    from clawpack_amd import pyclaw, riemann
    solver = pyclaw.ClawSolver2D(riemann.shallow_roe_with_efix_2D)
    solver.kernel_language = 'HIP'      # the only compute path; 'Fortran' is not available

The compute path is libclawpack_amd.so (csrc/claw2d.hip); see include/clawpack_amd.h.
"""
__version__ = "0.1.0"
