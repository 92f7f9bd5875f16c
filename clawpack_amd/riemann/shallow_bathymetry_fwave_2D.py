"""2D shallow water over variable bathymetry, f-wave Roe solver (upstream
riemann/src/rpn2_shallow_bathymetry_fwave.f90 and rpt2_shallow_bathymetry_fwave.f90).

q = (h, hu, hv), aux[0] = bathymetry b (the free surface is h + b).  The solver returns f-waves,
so the solver must be run with solver.fwave = True.  problem_data: 'grav' (gravitational constant)
and 'dry_tolerance' (a cell with h <= dry_tolerance has zero velocity); 'sea_level' is accepted
and not used by the Riemann solver.  Roe speeds u-c, u, u+c.
"""
num_eqn = 3
num_waves = 3
num_aux = 1
fwave = True
kernel_id = 2  # CLAW_SOLVER_SHALLOW_BATHY_FWAVE
problem_data_keys = ("grav", "dry_tolerance")
depth = 0
x_momentum = 1
y_momentum = 2
