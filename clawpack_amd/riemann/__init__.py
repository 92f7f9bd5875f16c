"""Riemann-solver plugins available to the HIP path (upstream package ``clawpack.riemann``).

The solvers on the path are provided; each module carries the ``num_eqn``/``num_waves``
of upstream ``riemann.static`` and the kernel id of include/clawpack_amd.h.  Upstream passes the
f2py ``rpn2``/``rpt2`` procedure pointers down to step2; here the solver is a compile-time policy
of the fused kernel and the module is the handle that selects it.
"""
from . import euler_4wave_2D, shallow_bathymetry_fwave_2D, shallow_roe_with_efix_2D, static  # noqa: F401

__all__ = ["shallow_roe_with_efix_2D", "euler_4wave_2D", "shallow_bathymetry_fwave_2D", "static"]
