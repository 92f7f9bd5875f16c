"""Many independent grids advanced together on one GPU.

All grids share the method (solver, precision, limiters, BCs, physics) and differ in size and
spacing; each keeps its own time, its own CFL-chosen dt and its own accept/reject history, as if
each had its own pyclaw solver.  Per attempt the library launches one persistent kernel over the
(grid, strip, segment) work items of every grid plus one per-grid control kernel; a run of
attempts can be captured once as a HIP graph and replayed.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import ClawBatchGrid, ClawCtrl, check
from .device import _stream_handle

_CTRL_BYTES = C.sizeof(ClawCtrl)
_ALIGN = 64  # grid planes start on 64-element boundaries (aligned rows for the first strip)


class BatchedGrids:
    """grids: list of (q (num_eqn, mx, my) host array, dx, dy); desc: the shared method."""

    def __init__(self, desc, grids, seg_h=0, device="cuda"):
        self.lib = _lib.load()
        check(self.lib.claw2d_validate(C.byref(desc)), "claw2d_validate")
        if not torch.cuda.is_available():
            raise RuntimeError("clawpack_amd needs a HIP device (torch.cuda.is_available() is False)")
        self.desc = desc
        self.device = torch.device(device)
        self.torch_dtype = torch.float64 if desc.dtype == _lib.FP64 else torch.float32
        self.neq = desc.num_eqn
        self.n = len(grids)
        table = (ClawBatchGrid * self.n)()
        off = 0
        self.shapes = []
        for g, (q, dx, dy) in enumerate(grids):
            neq, mx, my = q.shape
            assert neq == self.neq
            table[g].offset, table[g].mx, table[g].my = off, mx, my
            table[g].dx, table[g].dy = float(dx), float(dy)
            self.shapes.append((mx, my))
            off += -(-neq * mx * my // _ALIGN) * _ALIGN
        total = C.c_int64(0)
        check(self.lib.claw2d_batch_prepare(C.byref(desc), table, self.n, int(seg_h), C.byref(total)),
              "claw2d_batch_prepare")
        self.table_host = table
        self.total_items = total.value
        self.offsets = [table[g].offset for g in range(self.n)]
        raw = np.frombuffer(bytes(table), dtype=np.uint8)
        self.table = torch.from_numpy(raw.copy()).to(self.device)
        self.q = [torch.zeros(off, dtype=self.torch_dtype, device=self.device),
                  torch.zeros(off, dtype=self.torch_dtype, device=self.device)]
        self.ctrls = torch.zeros(self.n * _CTRL_BYTES // 8, dtype=torch.float64, device=self.device)
        self.counter = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.graphs = []
        host = torch.empty(off, dtype=self.torch_dtype)
        for g, (q, dx, dy) in enumerate(grids):
            mx, my = self.shapes[g]
            qq = torch.from_numpy(np.array(q, dtype=np.float64))
            if desc.depth_ref != 0.0:
                qq[0] -= desc.depth_ref  # stored depth h - depth_ref (FP32 shallow water)
            planes = qq.permute(0, 2, 1)
            host[self.offsets[g]:self.offsets[g] + self.neq * mx * my] = \
                planes.contiguous().reshape(-1).to(self.torch_dtype)
        self.q[0].copy_(host)

    def __del__(self):
        for gexec in getattr(self, "graphs", []):
            self.lib.claw2d_graph_destroy(gexec)

    # ---- control blocks ---------------------------------------------------------------------
    def init(self, t0=0.0, dt0=0.1, tend=None, target_steps=0, max_steps=0):
        buf = bytearray()
        for g in range(self.n):
            c = ClawCtrl()
            d = self._desc_of(g)
            check(self.lib.claw2d_ctrl_init(C.byref(d), C.byref(c), float(t0), float(dt0),
                                            float(tend if tend is not None else 0.0), int(tend is not None),
                                            int(target_steps), int(max_steps)), "claw2d_ctrl_init")
            buf += bytes(c)
        self.ctrls.copy_(torch.from_numpy(np.frombuffer(bytes(buf), dtype=np.float64).copy()))
        self.counter.zero_()

    def read_ctrls(self):
        raw = self.ctrls.cpu().numpy().tobytes()
        return [ClawCtrl.from_buffer_copy(raw[g * _CTRL_BYTES:(g + 1) * _CTRL_BYTES]) for g in range(self.n)]

    def _desc_of(self, g):
        d = ClawBatchGrid  # noqa: F841 (type reference for readers)
        dd = type(self.desc).from_buffer_copy(bytes(self.desc))
        dd.mx, dd.my = self.shapes[g]
        dd.dx, dd.dy = self.table_host[g].dx, self.table_host[g].dy
        return dd

    # ---- advancing ------------------------------------------------------------------------------
    def _args(self):
        return (C.byref(self.desc), C.c_void_p(self.table.data_ptr()), self.n, self.total_items,
                C.c_void_p(self.q[0].data_ptr()), C.c_void_p(self.q[1].data_ptr()),
                C.c_void_p(self.ctrls.data_ptr()), C.c_void_p(self.counter.data_ptr()))

    def advance(self, attempts, stream=None):
        check(self.lib.claw2d_batch_advance(*self._args(), int(attempts), _stream_handle(stream)),
              "claw2d_batch_advance")

    def make_graph(self, attempts):
        """Capture `attempts` batched attempts as a HIP graph; returns a handle for replay()."""
        h = C.c_void_p()
        check(self.lib.claw2d_batch_graph_create(*self._args(), int(attempts), C.byref(h)),
              "claw2d_batch_graph_create")
        self.graphs.append(h)
        return h

    def replay(self, handle, stream=None):
        check(self.lib.claw2d_graph_launch(handle, _stream_handle(stream)), "claw2d_graph_launch")

    def advance_sequential(self, attempts, stream=None):
        """The baseline: each grid advanced by its own single-grid launches, one grid after another
        (per attempt: step kernel + control kernel per grid)."""
        es = self.q[0].element_size()
        base0, base1 = self.q[0].data_ptr(), self.q[1].data_ptr()
        cbase = self.ctrls.data_ptr()
        sh = _stream_handle(stream)
        descs = [self._desc_of(g) for g in range(self.n)]
        for _ in range(int(attempts)):
            for g in range(self.n):
                off = self.offsets[g] * es
                check(self.lib.claw2d_advance(C.byref(descs[g]), C.c_void_p(base0 + off),
                                              C.c_void_p(base1 + off),
                                              C.c_void_p(cbase + g * _CTRL_BYTES), 1, sh),
                      "claw2d_advance")

    # ---- results --------------------------------------------------------------------------------
    def download(self, g, ctrls=None):
        c = (ctrls or self.read_ctrls())[g]
        mx, my = self.shapes[g]
        buf = self.q[int(c.cur)]
        flat = buf[self.offsets[g]:self.offsets[g] + self.neq * mx * my]
        t = flat.reshape(self.neq, my, mx).to(torch.float64).cpu()
        if self.desc.depth_ref != 0.0:
            t[0] += self.desc.depth_ref
        return np.asfortranarray(t.permute(0, 2, 1).numpy())

    def cells(self):
        return sum(mx * my for mx, my in self.shapes)
