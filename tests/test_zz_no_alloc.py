"""claw2d_mass_sum_noalloc (the mass sum with its partials in the library's own device array)
against claw2d_mass_sum with a caller-allocated workspace, through the C ABI: bit for bit, FP64 and
FP32, on two calls in a row, the second on a smaller grid (the array keeps the first call's
contents), then again after other work on the same GPU."""
import ctypes as C

import numpy as np
import pytest

from clawpack_amd import _lib, problems
from clawpack_amd.device import make_desc

pytestmark = pytest.mark.gpu


def _state(mx, my, dtype_id, seed):
    import torch
    q, dx, dy = problems.gaussian_humps(mx, my, seed)
    tdt = torch.float64 if dtype_id == _lib.FP64 else torch.float32
    qd = torch.tensor(np.ascontiguousarray(np.transpose(q, (0, 2, 1))), dtype=tdt, device="cuda:0")
    # momenta of both signs and a spread of magnitudes, so the compensation term matters
    qd[1] = torch.randn(my, mx, dtype=torch.float64, device="cuda:0").to(tdt) * 1e3
    qd[2] = torch.randn(my, mx, dtype=torch.float64, device="cuda:0").to(tdt) * 1e-3
    d = make_desc(0, dtype_id, mx, my, dx, dy, 3, 3, bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0)
    return d, qd


def _sums(lib, d, qd):
    import torch
    s = torch.cuda.current_stream()
    sp = C.c_void_p(s.cuda_stream)
    ws = torch.full(((lib.claw2d_mass_sum_workspace_bytes(C.byref(d)) + 7) // 8,), float("nan"),
                    dtype=torch.float64, device="cuda:0")
    out_ws = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda:0")
    out_na = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda:0")
    for m in range(3):
        assert lib.claw2d_mass_sum(C.byref(d), C.c_void_p(qd.data_ptr()), m,
                                   C.c_void_p(out_ws.data_ptr() + 8 * m), C.c_void_p(ws.data_ptr()), sp) == 0
        rc = lib.claw2d_mass_sum_noalloc(C.byref(d), C.c_void_p(qd.data_ptr()), m,
                                         C.c_void_p(out_na.data_ptr() + 8 * m), sp)
        assert rc == 0, lib.claw2d_last_error()
    torch.cuda.synchronize()
    return out_ws.cpu().numpy(), out_na.cpu().numpy()


@pytest.mark.parametrize("dtype_id", [_lib.FP64, _lib.FP32])
def test_noalloc_mass_sum_matches_workspace_version(dtype_id):
    lib = _lib.load()
    print("noalloc mass sum, dtype %d: 1024x768 then 96x40" % dtype_id, flush=True)
    for (mx, my, seed) in [(1024, 768, 1), (96, 40, 2), (1024, 768, 3)]:
        d, qd = _state(mx, my, dtype_id, seed)
        a, b = _sums(lib, d, qd)
        assert np.all(np.isfinite(a))
        assert a.tobytes() == b.tobytes(), (mx, my, a, b)
        # and against a plain FP64 sum of the same values
        ref = qd.double().sum(dim=(1, 2)).cpu().numpy()
        assert np.allclose(a, ref, rtol=1e-12, atol=1e-9 * np.abs(ref).max())


def test_noalloc_mass_sum_rejects_bad_arguments():
    lib = _lib.load()
    d, qd = _state(64, 32, _lib.FP64, 4)
    out = C.c_void_p(qd.data_ptr())
    assert lib.claw2d_mass_sum_noalloc(C.byref(d), C.c_void_p(qd.data_ptr()), 3, out, None) < 0
    assert b"component" in lib.claw2d_last_error()
    assert lib.claw2d_mass_sum_noalloc(C.byref(d), None, 0, out, None) < 0
    assert lib.claw2d_mass_sum_noalloc_max_cells() == 1 << 31
