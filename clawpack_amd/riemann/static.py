"""num_eqn / num_waves by solver module name (upstream riemann/riemann/static.py)."""
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
