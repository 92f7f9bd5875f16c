"""Wave-limiter ids (upstream pyclaw/src/pyclaw/limiters/tvd.py; the classic Fortran limiter
interprets them).  The HIP kernel implements ids 0-5."""
minmod = 1
superbee = 2
vanleer = 3
MC = 4
beam_warming = 5
# no limiting
unlimited = 0
