"""PyClaw-compatible front end on the CPU: containers, defaults and the refusal paths.
Nothing here computes a step; the HIP path refuses to run without a device."""
import numpy as np
import pytest

from clawpack_amd import pyclaw, riemann
from conftest import gpu_available


def make(nx=20, ny=10, solver_mod=None):
    solver_mod = solver_mod or riemann.shallow_roe_with_efix_2D
    solver = pyclaw.ClawSolver2D(solver_mod)
    solver.dimensional_split = False
    solver.bc_lower = [pyclaw.BC.extrap, pyclaw.BC.extrap]
    solver.bc_upper = [pyclaw.BC.extrap, pyclaw.BC.extrap]
    x = pyclaw.Dimension(-2.5, 2.5, nx, name="x")
    y = pyclaw.Dimension(-2.5, 2.5, ny, name="y")
    domain = pyclaw.Domain([x, y])
    state = pyclaw.State(domain, solver.num_eqn)
    state.problem_data["grav"] = 1.0
    state.problem_data["gamma"] = 1.4
    state.q[0] = 1.0
    if solver.num_eqn == 4:
        state.q[3] = 2.5
    return solver, state, domain, pyclaw.Solution(state, domain)


def test_geometry_and_state():
    solver, state, domain, sol = make()
    assert state.q.shape == (3, 20, 10)
    assert domain.grid.x.delta == 0.25 and domain.grid.num_cells == [20, 10]
    X, Y = state.grid.p_centers
    assert X.shape == (20, 10) and X[0, 0] == -2.375 and Y[0, 1] == -1.75
    assert sol.t == 0.0 and sol.state is state
    d2 = pyclaw.Domain((0.0, 0.0), (1.0, 2.0), (4, 8))
    assert d2.grid.delta == [0.25, 0.25]


def test_defaults_follow_pyclaw():
    s = pyclaw.ClawSolver2D(riemann.euler_4wave_2D)
    assert (s.num_eqn, s.num_waves) == (4, 4)
    assert s.dimensional_split is True and s.transverse_waves == 1 and s.order == 2
    assert s.limiters == pyclaw.limiters.tvd.minmod
    assert (s.cfl_max, s.cfl_desired, s.dt_initial, s.max_steps) == (1.0, 0.9, 0.1, 10000)
    assert s.kernel_language == "HIP"


@pytest.mark.parametrize("attr,value,exc", [
    ("kernel_language", "Fortran", NotImplementedError),
    ("kernel_language", "Python", NotImplementedError),
    ("fwave", True, ValueError),  # shallow_roe_with_efix returns waves, not f-waves
    ("order", 3, ValueError),
    ("precision", "fp16", ValueError),
])
def test_unsupported_configurations_raise(attr, value, exc):
    solver, state, domain, sol = make()
    setattr(solver, attr, value)
    with pytest.raises(exc):
        solver.setup(sol)


def test_bc_checks():
    solver, state, domain, sol = make()
    solver.bc_lower = [pyclaw.BC.custom, pyclaw.BC.extrap]
    with pytest.raises(Exception, match="Lower custom BC function has not been set"):
        solver.setup(sol)
    solver.user_bc_lower = lambda state, dim, t, qbc, auxbc, num_ghost: None
    for split in (True, False):  # custom sides run with both step forms
        solver.dimensional_split = split
        solver._check_supported(state)
    solver.bc_lower = [pyclaw.BC.extrap, pyclaw.BC.extrap]
    solver.bc_upper = [pyclaw.BC.extrap, pyclaw.BC.custom]
    with pytest.raises(Exception, match="Upper custom BC function has not been set"):
        solver.setup(sol)
    solver.bc_upper = [pyclaw.BC.extrap, pyclaw.BC.extrap]
    solver.bc_lower = [pyclaw.BC.periodic, pyclaw.BC.extrap]
    with pytest.raises(Exception, match="Periodic"):
        solver.setup(sol)
    solver.bc_lower = [None, pyclaw.BC.extrap]
    with pytest.raises(Exception, match="boundary conditions"):
        solver.setup(sol)


def test_problem_data_required():
    solver, state, domain, sol = make()
    del state.problem_data["grav"]
    with pytest.raises(KeyError):
        solver.setup(sol)


def test_limiter_list_length():
    solver, state, domain, sol = make()
    solver.limiters = [1, 4]
    with pytest.raises(Exception, match="limiters"):
        solver.setup(sol)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device refusal")
def test_no_cpu_fallback_without_device():
    solver, state, domain, sol = make()
    with pytest.raises(RuntimeError, match="HIP device"):
        solver.evolve_to_time(sol, 0.1)


def make_bathy(nx=20, ny=10):
    solver = pyclaw.ClawSolver2D(riemann.shallow_bathymetry_fwave_2D)
    solver.fwave = True
    solver.bc_lower = [pyclaw.BC.wall, pyclaw.BC.extrap]
    solver.bc_upper = [pyclaw.BC.wall, pyclaw.BC.extrap]
    solver.aux_bc_lower = [pyclaw.BC.wall, pyclaw.BC.extrap]
    solver.aux_bc_upper = [pyclaw.BC.wall, pyclaw.BC.extrap]
    domain = pyclaw.Domain((0.0, 0.0), (1.0, 1.0), (nx, ny))
    state = pyclaw.State(domain, solver.num_eqn, 1)
    state.problem_data.update(grav=9.81, dry_tolerance=1e-3, sea_level=0.0)
    state.aux[0] = -1.0
    state.q[0] = 1.0
    return solver, state, domain, pyclaw.Solution(state, domain)


def test_bathymetry_plugin_metadata():
    rp = riemann.shallow_bathymetry_fwave_2D
    assert (rp.num_eqn, rp.num_waves, rp.num_aux, rp.fwave) == (3, 3, 1, True)
    assert riemann.static.num_aux["shallow_bathymetry_fwave_2D"] == 1
    s = pyclaw.ClawSolver2D(rp)
    assert (s.num_eqn, s.num_waves) == (3, 3)


@pytest.mark.parametrize("change,exc,match", [
    (lambda s, st: setattr(s, "fwave", False), ValueError, "fwave"),
    (lambda s, st: setattr(st, "aux", None), ValueError, "num_aux"),
    (lambda s, st: s.aux_bc_lower.__setitem__(0, None), Exception, "aux boundary"),
    (lambda s, st: s.aux_bc_upper.__setitem__(1, pyclaw.BC.custom), NotImplementedError, "custom"),
    (lambda s, st: s.aux_bc_lower.__setitem__(1, pyclaw.BC.periodic), Exception, "Periodic aux"),
    (lambda s, st: st.problem_data.pop("dry_tolerance"), KeyError, "dry_tolerance"),
    (lambda s, st: setattr(st, "index_capa", 1), ValueError, "index_capa"),
    (lambda s, st: (setattr(st, "index_capa", 0), s.bc_lower.__setitem__(0, pyclaw.BC.custom),
                    setattr(s, "user_bc_lower", lambda *a: None)), NotImplementedError, "capacity"),
])
def test_bathymetry_configuration_checks(change, exc, match):
    solver, state, domain, sol = make_bathy()
    change(solver, state)
    with pytest.raises(exc, match=match):
        solver.setup(sol)


def test_aux_with_a_solver_that_reads_none_is_refused():
    solver, state, domain, sol = make()
    state.aux = np.zeros((1, 20, 10), order="F")
    with pytest.raises(NotImplementedError, match="aux"):
        solver.setup(sol)


def upstream_module(name):
    """A stand-in for an f2py-built upstream Riemann module (clawpack.riemann.<name>): only what
    such a module carries, the module name and the rpn2/rpt2 procedure objects with their _cpointer
    (here placeholders; the HIP path never calls them)."""
    import types
    m = types.ModuleType("clawpack.riemann." + name)

    class _Proc:
        _cpointer = None

    m.rpn2, m.rpt2 = _Proc(), _Proc()
    return m


@pytest.mark.parametrize("name", ["shallow_roe_with_efix_2D", "euler_4wave_2D",
                                  "shallow_bathymetry_fwave_2D"])
def test_upstream_riemann_module_selects_the_same_solver(name):
    """An upstream module object selects the HIP solver by its name, as riemann.static does: the
    same num_eqn/num_waves and the same kernel descriptor, byte for byte, as this package's module."""
    ours = getattr(riemann, name)
    up = upstream_module(name)
    a, b = pyclaw.ClawSolver2D(ours), pyclaw.ClawSolver2D(up)
    assert (a.num_eqn, a.num_waves) == (b.num_eqn, b.num_waves) == (ours.num_eqn, ours.num_waves)
    assert a._policy() == b._policy() and a._policy().kernel_id == ours.kernel_id
    domain = pyclaw.Domain((0.0, 0.0), (1.0, 1.0), (12, 10))
    state = pyclaw.State(domain, ours.num_eqn, getattr(ours, "num_aux", 0))
    state.problem_data.update(grav=1.0, gamma=1.4, dry_tolerance=1e-3)
    for s in (a, b):
        s.fwave = bool(getattr(ours, "fwave", False))
        s.bc_lower = s.bc_upper = s.aux_bc_lower = s.aux_bc_upper = [pyclaw.BC.extrap] * 2
        s._check_supported(state)
    assert bytes(a._make_desc(state)) == bytes(b._make_desc(state))


def test_unknown_riemann_module_is_refused():
    solver, state, domain, sol = make()
    solver.rp = upstream_module("acoustics_2D")
    with pytest.raises(NotImplementedError, match="acoustics_2D"):
        solver.setup(sol)
