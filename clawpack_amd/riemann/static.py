"""Properties of the Riemann solvers by module name (upstream riemann/riemann/static.py).

pyclaw looks a solver's num_eqn/num_waves up by the trailing part of the module's ``__name__``,
so any module object of that name selects the solver: this package's handle modules and the
f2py-built ``clawpack.riemann.<name>`` modules of upstream alike (their rpn2/rpt2 procedures are
not called; the HIP kernel carries its own solver).  ``policy(rp)`` returns that entry.
"""
from collections import namedtuple

num_eqn = {
    "shallow_roe_with_efix_2D": 3,
    "euler_4wave_2D": 4,
    "shallow_bathymetry_fwave_2D": 3,
}
num_waves = {
    "shallow_roe_with_efix_2D": 3,
    "euler_4wave_2D": 4,
    "shallow_bathymetry_fwave_2D": 3,
}
num_aux = {
    "shallow_roe_with_efix_2D": 0,
    "euler_4wave_2D": 0,
    "shallow_bathymetry_fwave_2D": 1,
}
# the HIP path's solver ids (include/clawpack_amd.h CLAW_SOLVER_*), whether rpn2 returns f-waves,
# and the problem_data keys its Fortran cparam block reads
kernel_id = {
    "shallow_roe_with_efix_2D": 0,
    "euler_4wave_2D": 1,
    "shallow_bathymetry_fwave_2D": 2,
}
fwave = {
    "shallow_roe_with_efix_2D": False,
    "euler_4wave_2D": False,
    "shallow_bathymetry_fwave_2D": True,
}
problem_data_keys = {
    "shallow_roe_with_efix_2D": ("grav",),
    "euler_4wave_2D": ("gamma",),
    "shallow_bathymetry_fwave_2D": ("grav", "dry_tolerance"),
}

Policy = namedtuple("Policy", "name kernel_id num_eqn num_waves num_aux fwave problem_data_keys")


def solver_name(rp):
    """The trailing part of a Riemann module's __name__ (pyclaw's key into these tables)."""
    return str(getattr(rp, "__name__", "")).split(".")[-1]


def policy(rp):
    """The HIP solver selected by Riemann module rp, or None when the module names a solver the
    HIP path does not have."""
    name = solver_name(rp)
    if name not in kernel_id:
        return None
    return Policy(name, kernel_id[name], num_eqn[name], num_waves[name], num_aux[name], fwave[name],
                  problem_data_keys[name])
