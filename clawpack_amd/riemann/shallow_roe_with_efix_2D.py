"""2D shallow water, Roe solver with Harten-Hyman entropy fix (upstream
riemann/src/rpn2_shallow_roe_with_efix.f90 and rpt2_shallow_roe_with_efix.f90).

q = (h, hu, hv); problem_data['grav'] is the gravitational constant.
Wave speeds u-c, u, u+c with c = sqrt(grav*h) at the Roe average.
"""
num_eqn = 3
num_waves = 3
kernel_id = 0  # CLAW_SOLVER_SHALLOW_ROE_EFIX
problem_data_keys = ("grav",)
depth = 0
x_momentum = 1
y_momentum = 2
