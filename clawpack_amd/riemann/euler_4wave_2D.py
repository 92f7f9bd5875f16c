"""2D Euler equations, 4-wave Roe solver with entropy fix (upstream
riemann/src/rpn2_euler_4wave.f90 and rpt2_euler.f90).

q = (rho, rho u, rho v, E); problem_data['gamma'] is the ratio of specific heats
(gamma1 = gamma - 1 is derived).  Waves: u-c, entropy (u), shear (u), u+c.
"""
num_eqn = 4
num_waves = 4
kernel_id = 1  # CLAW_SOLVER_EULER_4WAVE
problem_data_keys = ("gamma",)
density = 0
x_momentum = 1
y_momentum = 2
energy = 3
