"""C ABI of libclawpack_amd.so: loads without a GPU, exports every function declared in
include/clawpack_amd.h, struct layouts match the binding, descriptor validation (no compute)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "clawpack_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(claw2d_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("claw2d_step", "claw2d_advance", "claw2d_validate", "claw2d_ctrl_init",
                 "claw2d_mass_sum", "claw2d_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from clawpack_amd import _lib
    lib = _lib.load()
    raw = C.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(raw, name), name
        assert name in _lib.SIGNATURES, "binding lacks %s" % name
    assert lib is not None


def test_struct_layouts():
    from clawpack_amd import _lib
    lib = _lib.load()
    assert lib.claw2d_abi_struct_size(0) == C.sizeof(_lib.ClawDesc)
    assert lib.claw2d_abi_struct_size(1) == C.sizeof(_lib.ClawCtrl)


def _desc(**kw):
    from clawpack_amd.device import make_desc
    args = dict(solver_id=0, dtype=0, mx=32, my=16, dx=0.1, dy=0.1, num_eqn=3, num_waves=3)
    args.update(kw)
    return make_desc(**args)


@pytest.mark.parametrize("bad,code", [
    (dict(mx=1), -5), (dict(num_eqn=4), -3), (dict(solver_id=7), -2), (dict(dtype=3), -4),
    (dict(num_ghost=3), -6), (dict(order=3), -7), (dict(transverse_waves=3), -8),
    (dict(mthlim=(9, 4, 4)), -9), (dict(bc_lower=(0, 1)), -10), (dict(bc_lower=(2, 1)), -11),
    (dict(reflect_index=(5, 2)), -12), (dict(dx=0.0), -13),
])
def test_validate_rejects(bad, code):
    from clawpack_amd import _lib
    lib = _lib.load()
    d = _desc(**bad)
    assert lib.claw2d_validate(C.byref(d)) == code
    assert lib.claw2d_last_error()


def test_validate_accepts_and_sizes():
    from clawpack_amd import _lib
    lib = _lib.load()
    d = _desc()
    assert lib.claw2d_validate(C.byref(d)) == 0
    assert lib.claw2d_state_bytes(C.byref(d)) == 3 * 32 * 16 * 8
    d32 = _desc(dtype=1, solver_id=1, num_eqn=4, num_waves=4)
    assert lib.claw2d_state_bytes(C.byref(d32)) == 4 * 32 * 16 * 4


def test_ctrl_init_clips_to_tend():
    from clawpack_amd import _lib
    lib = _lib.load()
    d = _desc()
    c = _lib.ClawCtrl()
    assert lib.claw2d_ctrl_init(C.byref(d), C.byref(c), 0.0, 0.1, 0.05, 1, 0, 0) == 0
    assert c.dt == 0.05 and c.done == 0 and c.flags & _lib.FLAG_RUN_TO_TEND
    assert lib.claw2d_ctrl_init(C.byref(d), C.byref(c), 0.0, 0.1, 0.0, 0, 5, 0) == 0
    assert c.target_steps == 5 and c.dt == 0.1 and not (c.flags & _lib.FLAG_RUN_TO_TEND)
    assert lib.claw2d_ctrl_init(C.byref(d), C.byref(c), 0.0, -1.0, 0.0, 0, 5, 0) == -21


def test_adjoint_workspace_and_validation():
    from clawpack_amd import _lib
    lib = _lib.load()
    d = _desc()
    assert lib.claw2d_adjoint_workspace_bytes(C.byref(d)) == 6 * (32 + 4) * (16 + 4) * 8
    # Euler and FP32 have no adjoint: size 0 and a message
    for bad in (_desc(solver_id=1, num_eqn=4, num_waves=4), _desc(dtype=1)):
        assert lib.claw2d_adjoint_workspace_bytes(C.byref(bad)) == 0
        assert b"adjoint" in lib.claw2d_last_error()
        assert lib.claw2d_adjoint_step(C.byref(bad), None, None, None, None, 0.1, None) < 0
    # null buffers and a non-positive dt are refused before any launch
    assert lib.claw2d_adjoint_step(C.byref(d), None, None, None, None, 0.1, None) < 0


def _bathy(**kw):
    args = dict(solver_id=2, num_aux=1, fwave=1, dry_tolerance=1e-3)
    args.update(kw)
    return _desc(**args)


@pytest.mark.parametrize("bad,code", [
    (dict(num_aux=0), -16), (dict(fwave=0), -17), (dict(dry_tolerance=-1.0), -18),
    (dict(aux_bc_lower=(0, 1)), -19), (dict(aux_bc_lower=(2, 1)), -11),
])
def test_validate_rejects_bathymetry(bad, code):
    from clawpack_amd import _lib
    lib = _lib.load()
    d = _bathy(**bad)
    assert lib.claw2d_validate(C.byref(d)) == code
    assert lib.claw2d_last_error()


def test_aux_and_fwave_only_with_the_bathymetry_solver():
    from clawpack_amd import _lib
    lib = _lib.load()
    assert lib.claw2d_validate(C.byref(_bathy())) == 0
    assert lib.claw2d_validate(C.byref(_desc(num_aux=1))) == -16
    assert lib.claw2d_validate(C.byref(_desc(fwave=1))) == -17
    assert lib.claw2d_state_bytes(C.byref(_bathy(dtype=1))) == 3 * 32 * 16 * 4


def test_entry_points_without_aux_refuse_aux_solvers():
    """Refused before any device call, so this runs without a GPU (the pointers are never used)."""
    from clawpack_amd import _lib
    lib = _lib.load()
    d = _bathy()
    qa, qb, ctrl, aux = C.c_void_p(4096), C.c_void_p(8192), C.c_void_p(12288), C.c_void_p(16384)
    assert lib.claw2d_step(C.byref(d), qa, qb, 0.1, ctrl, None) == -26
    assert b"claw2d_step_aux" in lib.claw2d_last_error()
    assert lib.claw2d_advance(C.byref(d), qa, qb, ctrl, 1, None) == -26
    assert lib.claw2d_step_ds(C.byref(d), qa, qb, 0.1, ctrl, None) == -26
    assert lib.claw2d_advance_ds(C.byref(d), qa, qb, ctrl, 1, None) == -26
    assert lib.claw2d_step_aux(C.byref(d), qa, qb, None, 0.1, ctrl, None) == -22
    assert lib.claw2d_advance_ds_aux(C.byref(d), qa, qb, None, ctrl, 1, None) == -22
    assert lib.claw2d_step_aux(C.byref(d), qa, qa, aux, 0.1, ctrl, None) == -23
    assert lib.claw2d_step_ds_aux(C.byref(d), qa, qb, aux, -0.1, ctrl, None) == -21
    # batched grids and the adjoint take no aux solvers
    assert lib.claw2d_adjoint_workspace_bytes(C.byref(d)) == 0


def test_step_padded_accepts_custom_codes_and_checks_arguments():
    """claw2d_step_padded takes the state's ghost cells from its input, so BC code 0 (custom) is
    accepted there and only there; argument errors are refused before any device call."""
    from clawpack_amd import _lib
    lib = _lib.load()
    d = _desc(bc_lower=(0, 1), bc_upper=(1, 0))
    assert lib.claw2d_validate(C.byref(d)) == -10
    qa, qb, ctrl = C.c_void_p(4096), C.c_void_p(8192), C.c_void_p(12288)
    assert lib.claw2d_step_padded(C.byref(d), None, qb, None, 0.1, ctrl, None) == -22
    assert lib.claw2d_step_padded(C.byref(d), qa, qa, None, 0.1, ctrl, None) == -23
    assert lib.claw2d_step_padded(C.byref(d), qa, qb, None, 0.0, ctrl, None) == -21
    bad = _desc(bc_lower=(7, 1))
    assert lib.claw2d_step_padded(C.byref(bad), qa, qb, None, 0.1, ctrl, None) == -10
    aux_d = _desc(solver_id=2, num_aux=1, fwave=1, bc_lower=(0, 3), aux_bc_lower=(1, 3), aux_bc_upper=(1, 1))
    assert lib.claw2d_step_padded(C.byref(aux_d), qa, qb, None, 0.1, ctrl, None) == -22  # null aux
    # aux boundary conditions are still index maps: custom is refused for them
    aux_custom = _desc(solver_id=2, num_aux=1, fwave=1, bc_lower=(0, 3), aux_bc_lower=(0, 3), aux_bc_upper=(1, 1))
    assert lib.claw2d_step_padded(C.byref(aux_custom), qa, qb, C.c_void_p(16384), 0.1, ctrl, None) == -19


def test_sweep_padded_checks_arguments():
    """claw2d_sweep_padded: the padded descriptor rules of claw2d_step_padded plus the sweep
    direction (1 or 2); every error is refused before any device call."""
    from clawpack_amd import _lib
    lib = _lib.load()
    d = _desc(bc_lower=(0, 1), bc_upper=(1, 0))
    qa, qb, ctrl = C.c_void_p(4096), C.c_void_p(8192), C.c_void_p(12288)
    for ids in (0, 3, -1):
        assert lib.claw2d_sweep_padded(C.byref(d), qa, qb, None, 0.1, ids, ctrl, None) == -27
    for ids in (1, 2):
        assert lib.claw2d_sweep_padded(C.byref(d), None, qb, None, 0.1, ids, ctrl, None) == -22
        assert lib.claw2d_sweep_padded(C.byref(d), qa, qb, None, 0.1, ids, None, None) == -22
        assert lib.claw2d_sweep_padded(C.byref(d), qa, qa, None, 0.1, ids, ctrl, None) == -23
        assert lib.claw2d_sweep_padded(C.byref(d), qa, qb, None, -1.0, ids, ctrl, None) == -21
    assert lib.claw2d_sweep_padded(None, qa, qb, None, 0.1, 1, ctrl, None) == -1
    bad = _desc(bc_lower=(7, 1))
    assert lib.claw2d_sweep_padded(C.byref(bad), qa, qb, None, 0.1, 1, ctrl, None) == -10
    aux_d = _desc(solver_id=2, num_aux=1, fwave=1, bc_lower=(0, 3), aux_bc_lower=(1, 3), aux_bc_upper=(1, 1))
    assert lib.claw2d_sweep_padded(C.byref(aux_d), qa, qb, None, 0.1, 2, ctrl, None) == -22


def test_capacity_descriptor_rules():
    """mcapa (method(6), 1-based aux plane of kappa): any solver takes aux arrays with a capacity
    function, mcapa must name an aux plane, and the padded, batched and adjoint entry points refuse
    it before any device call."""
    from clawpack_amd import _lib
    lib = _lib.load()
    ok = _desc(num_aux=1, index_capa=0)
    assert ok.mcapa == 1 and lib.claw2d_validate(C.byref(ok)) == 0
    assert lib.claw2d_validate(C.byref(_desc(solver_id=1, num_eqn=4, num_waves=4, num_aux=2,
                                             index_capa=1))) == 0
    assert lib.claw2d_validate(C.byref(_bathy(num_aux=2, index_capa=1))) == 0
    bad = _desc(num_aux=1)
    bad.mcapa = 2
    assert lib.claw2d_validate(C.byref(bad)) == -29
    bad.mcapa = -1
    assert lib.claw2d_validate(C.byref(bad)) == -29
    assert lib.claw2d_validate(C.byref(_desc(num_aux=1))) == -16  # aux without kappa: no reader
    assert lib.claw2d_validate(C.byref(_desc(num_aux=1, index_capa=0, aux_bc_lower=(0, 1)))) == -19
    qa, qb, ctrl, aux = C.c_void_p(4096), C.c_void_p(8192), C.c_void_p(12288), C.c_void_p(16384)
    assert lib.claw2d_step_padded(C.byref(ok), qa, qb, aux, 0.1, ctrl, None) == -28
    assert lib.claw2d_sweep_padded(C.byref(ok), qa, qb, aux, 0.1, 1, ctrl, None) == -28
    assert lib.claw2d_step(C.byref(ok), qa, qb, 0.1, ctrl, None) == -26  # needs the _aux form
    assert lib.claw2d_adjoint_workspace_bytes(C.byref(ok)) == 0
    assert lib.claw2d_batch_advance(C.byref(ok), qa, 1, 1, qa, qb, ctrl, aux, 1, None) == -25



@pytest.mark.parametrize("kw,ok", [
    (dict(dtype=1, depth_ref=1.0), True), (dict(dtype=0, depth_ref=1.0), False),
    (dict(dtype=1, depth_ref=float("nan")), False),
    (dict(solver_id=1, num_eqn=4, num_waves=4, dtype=1, depth_ref=1.0), False),
])
def test_depth_ref_only_for_fp32_shallow_water(kw, ok):
    """A stored depth deviation (h - depth_ref) is an FP32 shallow-water option."""
    from clawpack_amd import _lib
    lib = _lib.load()
    d = _desc(**kw)
    rc = lib.claw2d_validate(C.byref(d))
    assert (rc == 0) if ok else (rc == -35), rc
