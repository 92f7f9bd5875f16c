"""Device-resident grid: SoA state planes, control block and the calls into libclawpack_amd.so.

Memory is owned by PyTorch (HIP device tensors); the library only receives pointers.  The state
lives as two ping-pong buffers of shape (num_eqn, my, mx), x fastest; ``cur`` in the device
control block says which one holds q^n, so a rejected step costs nothing to undo.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import ClawCtrl, ClawDesc, check

_CTRL_BYTES = C.sizeof(ClawCtrl)


def _stream_handle(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


def make_desc(solver_id, dtype, mx, my, dx, dy, num_eqn, num_waves, order=2, transverse_waves=2,
              mthlim=(4, 4, 4, 4), bc_lower=(1, 1), bc_upper=(1, 1), reflect_index=(1, 2),
              grav=1.0, gamma=1.4, cfl_max=1.0, cfl_desired=0.9, dt_max=1e99, dt_variable=True,
              seg_h=0, num_ghost=2, num_aux=None, fwave=None, aux_bc_lower=None, aux_bc_upper=None,
              dry_tolerance=1e-3, index_capa=-1, depth_ref=0.0, wg_waves=0, launch_order=0):
    """num_aux and fwave default to what the solver needs (one aux field, the bathymetry, and
    f-waves for shallow_bathymetry_fwave; with a capacity function at least index_capa + 1 aux
    fields); aux BCs default to the state's.  index_capa >= 0: aux[index_capa] is the capacity
    function kappa (pyclaw's state.index_capa), the step is the capacity form.  depth_ref (FP32
    shallow water only): the device state stores the depth as h - depth_ref; upload, download and
    mass_sum convert, the raw planes (planes(), set_planes(), step_into) hold the stored form.
    wg_waves (1, 2, 4; 0 = 4) and launch_order (0 XCD-contiguous, 1 plain) set the launch shape
    of the single-grid kernels, as seg_h sets the rows per wavefront; results do not depend on
    them."""
    d = ClawDesc()
    d.depth_ref = float(depth_ref)
    d.wg_waves, d.launch_order = int(wg_waves), int(launch_order)
    bathy = int(solver_id) == _lib.SOLVER_SHALLOW_BATHY_FWAVE
    if num_aux is None:
        num_aux = max(1 if bathy else 0, int(index_capa) + 1)
    d.num_aux = int(num_aux)
    d.mcapa = int(index_capa) + 1 if index_capa is not None and index_capa >= 0 else 0
    d.fwave = int(bool(bathy if fwave is None else fwave))
    d.dry_tolerance = float(dry_tolerance)
    abl = bc_lower if aux_bc_lower is None else aux_bc_lower
    abu = bc_upper if aux_bc_upper is None else aux_bc_upper
    for k in range(2):
        d.aux_bc_lower[k], d.aux_bc_upper[k] = int(abl[k]), int(abu[k])
    d.solver, d.dtype, d.mx, d.my = int(solver_id), int(dtype), int(mx), int(my)
    d.num_eqn, d.num_waves, d.num_ghost = int(num_eqn), int(num_waves), int(num_ghost)
    d.order, d.transverse_waves = int(order), int(transverse_waves)
    lim = list(mthlim)
    lim = lim + [lim[-1]] * (4 - len(lim))
    for k in range(4):
        d.mthlim[k] = int(lim[k])
    for k in range(2):
        d.bc_lower[k], d.bc_upper[k] = int(bc_lower[k]), int(bc_upper[k])
        d.reflect_index[k] = int(reflect_index[k])
    d.seg_h, d.dt_variable = int(seg_h), int(bool(dt_variable))
    d.dx, d.dy, d.grav, d.gamma = float(dx), float(dy), float(grav), float(gamma)
    d.cfl_max, d.cfl_desired, d.dt_max = float(cfl_max), float(cfl_desired), float(dt_max)
    return d


class DeviceGrid:
    """One grid on one GPU: q ping-pong pair, control block, mass-sum workspace, and the aux
    planes (num_aux, my, mx) when the solver reads aux.  split=True steps with dimensional
    splitting (step2ds) instead of the unsplit step2."""

    def __init__(self, desc, device="cuda", split=False):
        self.lib = _lib.load()
        check(self.lib.claw2d_validate(C.byref(desc)), "claw2d_validate")
        if not torch.cuda.is_available():
            raise RuntimeError("clawpack_amd needs a HIP device (torch.cuda.is_available() is False)")
        self.desc = desc
        self.device = torch.device(device)
        self.torch_dtype = torch.float64 if desc.dtype == _lib.FP64 else torch.float32
        shape = (desc.num_eqn, desc.my, desc.mx)
        self.q = [torch.zeros(shape, dtype=self.torch_dtype, device=self.device),
                  torch.zeros(shape, dtype=self.torch_dtype, device=self.device)]
        nb = (_CTRL_BYTES + 7) // 8
        self.ctrl_t = torch.zeros(nb, dtype=torch.float64, device=self.device)
        ws = self.lib.claw2d_mass_sum_workspace_bytes(C.byref(desc))
        self.mass_ws = torch.zeros((ws + 7) // 8, dtype=torch.float64, device=self.device)
        self.mass_out = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.cur = 0
        self.split = bool(split)
        self.aux = torch.zeros((desc.num_aux, desc.my, desc.mx), dtype=self.torch_dtype,
                               device=self.device) if desc.num_aux > 0 else None

    def _step_fn(self):
        """(q_in, q_out, dt, ctrl, stream) -> status, with the aux planes bound when present."""
        if self.aux is None:
            return self.lib.claw2d_step_ds if self.split else self.lib.claw2d_step
        fn = self.lib.claw2d_step_ds_aux if self.split else self.lib.claw2d_step_aux
        aux = C.c_void_p(self.aux.data_ptr())
        return lambda d, qi, qo, dt, ctrl, s: fn(d, qi, qo, aux, dt, ctrl, s)

    def _advance_fn(self):
        if self.aux is None:
            return self.lib.claw2d_advance_ds if self.split else self.lib.claw2d_advance
        fn = self.lib.claw2d_advance_ds_aux if self.split else self.lib.claw2d_advance_aux
        aux = C.c_void_p(self.aux.data_ptr())
        return lambda d, q0, q1, ctrl, n, s: fn(d, q0, q1, aux, ctrl, n, s)

    def upload_aux(self, aux_host):
        """aux_host: pyclaw layout (num_aux, mx, my) -> device planes (num_aux, my, mx)."""
        if self.aux is None:
            raise ValueError("this solver takes no aux arrays")
        a = torch.from_numpy(np.ascontiguousarray(np.asarray(aux_host, dtype=np.float64)))
        self.aux.copy_(a.permute(0, 2, 1).contiguous().to(self.torch_dtype))

    def set_aux_planes(self, planes):
        """planes: device or host tensor (num_aux, my, mx)."""
        self.aux.copy_(planes)

    # ---- control block ------------------------------------------------------------------
    @property
    def ctrl_ptr(self):
        return C.c_void_p(self.ctrl_t.data_ptr())

    def read_ctrl(self):
        raw = self.ctrl_t.cpu().numpy().tobytes()[:_CTRL_BYTES]
        return ClawCtrl.from_buffer_copy(raw)

    def write_ctrl(self, c):
        buf = np.frombuffer(bytes(c) + b"\0" * (self.ctrl_t.numel() * 8 - _CTRL_BYTES), dtype=np.float64)
        self.ctrl_t.copy_(torch.from_numpy(buf.copy()))

    def init_ctrl(self, t0, dt0, tend=None, target_steps=0, max_steps=0):
        c = ClawCtrl()
        run_to_tend = tend is not None
        check(self.lib.claw2d_ctrl_init(C.byref(self.desc), C.byref(c), float(t0), float(dt0),
                                        float(tend if run_to_tend else 0.0), int(run_to_tend),
                                        int(target_steps), int(max_steps)), "claw2d_ctrl_init")
        c.cur = self.cur
        self.write_ctrl(c)
        return c

    # ---- state transfer -------------------------------------------------------------------
    def upload(self, q_host):
        """q_host: pyclaw layout (num_eqn, mx, my) -> device planes (num_eqn, my, mx); with a
        depth_ref the depth is stored as h - depth_ref (formed in FP64, then rounded)."""
        q = torch.from_numpy(np.array(q_host, dtype=np.float64))
        if self.desc.depth_ref != 0.0:
            q[0] -= self.desc.depth_ref
        planes = q.permute(0, 2, 1).contiguous().to(self.torch_dtype)
        self.q[self.cur].copy_(planes)

    def set_planes(self, planes):
        """planes: device or host tensor (num_eqn, my, mx)."""
        self.q[self.cur].copy_(planes)

    def planes(self):
        return self.q[self.cur]

    def download(self):
        """Current state in pyclaw layout (num_eqn, mx, my), float64 numpy, Fortran order (a
        stored depth deviation is turned back into h in FP64)."""
        t = self.q[self.cur].to(torch.float64).cpu()
        if self.desc.depth_ref != 0.0:
            t[0] += self.desc.depth_ref
        return np.asfortranarray(t.permute(0, 2, 1).numpy())

    # ---- compute --------------------------------------------------------------------------
    def step(self, dt, stream=None):
        """One f2py-style step2 with fixed dt (q[cur] -> q[other]); returns the CFL number.
        The control block's smax accumulators must be zero (init_ctrl leaves them zero)."""
        src, dst = self.q[self.cur], self.q[1 - self.cur]
        check(self._step_fn()(C.byref(self.desc), C.c_void_p(src.data_ptr()),
                              C.c_void_p(dst.data_ptr()), float(dt), self.ctrl_ptr,
                              _stream_handle(stream)), "claw2d_step")
        self.cur = 1 - self.cur
        c = self.read_ctrl()
        return c.cfl, c

    def step_into(self, src, dst, dt, stream=None):
        """One fixed-dt step from the planes `src` into the planes `dst` (device tensors of this
        grid's shape and dtype; they may be any buffers, cur is not changed).  The CFL number and
        status bits land in the control block, as for step()."""
        for t in (src, dst):
            if tuple(t.shape) != tuple(self.q[0].shape) or t.dtype != self.torch_dtype or not t.is_contiguous():
                raise ValueError("step_into needs contiguous planes of shape %s" % (tuple(self.q[0].shape),))
        if src.data_ptr() == dst.data_ptr():
            raise ValueError("step_into: src and dst must differ")
        check(self._step_fn()(C.byref(self.desc), C.c_void_p(src.data_ptr()),
                              C.c_void_p(dst.data_ptr()), float(dt), self.ctrl_ptr,
                              _stream_handle(stream)), "claw2d_step")

    def step_padded(self, qbc, dst, dt, stream=None):
        """One unsplit fixed-dt step from the ghost-padded planes `qbc` (num_eqn, my+4, mx+4), whose
        ghost cells the caller has filled, into the interior planes `dst`; cur is not changed."""
        d = self.desc
        shape = (d.num_eqn, d.my + 4, d.mx + 4)
        if tuple(qbc.shape) != shape or qbc.dtype != self.torch_dtype or not qbc.is_contiguous():
            raise ValueError("step_padded needs contiguous padded planes of shape %s" % (shape,))
        if tuple(dst.shape) != tuple(self.q[0].shape) or dst.dtype != self.torch_dtype or not dst.is_contiguous():
            raise ValueError("step_padded: dst must be interior planes of shape %s" % (tuple(self.q[0].shape),))
        aux = C.c_void_p(self.aux.data_ptr()) if self.aux is not None else None
        check(self.lib.claw2d_step_padded(C.byref(d), C.c_void_p(qbc.data_ptr()), C.c_void_p(dst.data_ptr()),
                                          aux, float(dt), self.ctrl_ptr, _stream_handle(stream)),
              "claw2d_step_padded")

    def sweep_padded(self, qbc, dst, dt, ids, stream=None):
        """One sweep of the split step from the ghost-padded planes `qbc` (num_eqn, my+4, mx+4):
        ids 1 writes q* into the interior of the padded `dst` (its ghost cells are the caller's to
        fill), ids 2 writes the interior planes `dst`.  ids 1 leaves the CFL accumulation open and
        ids 2 closes it (ctrl.cfl = max over both sweeps); cur is not changed."""
        d = self.desc
        shape = (d.num_eqn, d.my + 4, d.mx + 4)
        if tuple(qbc.shape) != shape or qbc.dtype != self.torch_dtype or not qbc.is_contiguous():
            raise ValueError("sweep_padded needs contiguous padded planes of shape %s" % (shape,))
        out_shape = shape if ids == 1 else tuple(self.q[0].shape)
        if tuple(dst.shape) != out_shape or dst.dtype != self.torch_dtype or not dst.is_contiguous():
            raise ValueError("sweep_padded(ids=%d): dst must have shape %s" % (ids, out_shape))
        aux = C.c_void_p(self.aux.data_ptr()) if self.aux is not None else None
        check(self.lib.claw2d_sweep_padded(C.byref(d), C.c_void_p(qbc.data_ptr()), C.c_void_p(dst.data_ptr()),
                                           aux, float(dt), int(ids), self.ctrl_ptr, _stream_handle(stream)),
              "claw2d_sweep_padded")

    def enqueue_step(self, dt, stream=None):
        """Like step() but without reading the control block back (no host synchronisation)."""
        src, dst = self.q[self.cur], self.q[1 - self.cur]
        check(self._step_fn()(C.byref(self.desc), C.c_void_p(src.data_ptr()),
                              C.c_void_p(dst.data_ptr()), float(dt), self.ctrl_ptr,
                              _stream_handle(stream)), "claw2d_step")
        self.cur = 1 - self.cur

    def advance(self, attempts, stream=None):
        """Enqueue `attempts` evolve_to_time attempts (device-side accept/reject and dt)."""
        check(self._advance_fn()(C.byref(self.desc), C.c_void_p(self.q[0].data_ptr()),
                                 C.c_void_p(self.q[1].data_ptr()), self.ctrl_ptr,
                                 int(attempts), _stream_handle(stream)), "claw2d_advance")

    def sync_cur(self, c=None):
        c = self.read_ctrl() if c is None else c
        self.cur = int(c.cur)
        return c

    def mass_sum(self, m=0, stream=None):
        """Compensated FP64 sum of component m (deterministic order), returned as a float."""
        check(self.lib.claw2d_mass_sum(C.byref(self.desc), C.c_void_p(self.q[self.cur].data_ptr()),
                                       int(m), C.c_void_p(self.mass_out.data_ptr()),
                                       C.c_void_p(self.mass_ws.data_ptr()), _stream_handle(stream)),
              "claw2d_mass_sum")
        return float(self.mass_out.item())
