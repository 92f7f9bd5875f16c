"""num_eqn / num_waves by solver module name (upstream riemann/riemann/static.py)."""
num_eqn = {
    "shallow_roe_with_efix_2D": 3,
    "euler_4wave_2D": 4,
}
num_waves = {
    "shallow_roe_with_efix_2D": 3,
    "euler_4wave_2D": 4,
}
