"""Adjoint (reverse-mode gradient) of the 2D shallow-water step on the device.

The forward map is q^0 -> q^1 -> ... -> q^N with a *frozen* dt sequence (dt_n fixed, taken from a
CFL-controlled run), so every step is a fixed function of its input and the chain rule applies
step by step.  Gradients come from the hand-written reverse step ``claw2d_adjoint_step`` (branches
of the entropy fix, limiter and transverse split taken as in the forward), not from a tape:
forward states are recomputed from checkpoints every ``checkpoint_every`` steps (sqrt(N) memory;
the last segment, which the backward reverses first, is kept as the forward computes it), and
lambda is propagated backwards in FP64 on the GPU.

    samples, qN = evolve_frozen(q0, desc, dts, gauges, sample_every=10, checkpoint_every=20)

is differentiable in q0 (a device tensor of shape (3, my, mx), SoA planes, float64); samples has
shape (N // sample_every, 3, len(gauges)) and holds q at the gauge cells (flat index y*mx + x)
after steps sample_every, 2*sample_every, ...
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check
from .device import DeviceGrid, _stream_handle


def adjoint_workspace(desc, device="cuda"):
    lib = _lib.load()
    nb = lib.claw2d_adjoint_workspace_bytes(C.byref(desc))
    if nb == 0:
        raise _lib.ClawLibError("claw2d_adjoint_workspace_bytes: " + lib.claw2d_last_error().decode())
    return torch.empty((nb + 7) // 8, dtype=torch.float64, device=device)


def adjoint_step(desc, q, lam_in, dt, lam_out=None, workspace=None, stream=None):
    """lam_out = (d step(q) / d q)^T lam_in for one fixed-dt step (device tensors (3, my, mx))."""
    lib = _lib.load()
    for t in (q, lam_in):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("adjoint_step needs contiguous float64 device tensors")
        if tuple(t.shape) != (3, desc.my, desc.mx):
            raise ValueError("expected shape (3, my, mx) = (3, %d, %d)" % (desc.my, desc.mx))
    if lam_out is None:
        lam_out = torch.empty_like(lam_in)
    if workspace is None:
        workspace = adjoint_workspace(desc, q.device)
    check(lib.claw2d_adjoint_step(C.byref(desc), C.c_void_p(q.data_ptr()), C.c_void_p(lam_in.data_ptr()),
                                  C.c_void_p(lam_out.data_ptr()), C.c_void_p(workspace.data_ptr()),
                                  float(dt), _stream_handle(stream)), "claw2d_adjoint_step")
    return lam_out


_STREAMS = {}


def _stream_pair(device):
    """Two streams of the default priority, made one after the other and kept for the process:
    the backward reverses on the first and recomputes checkpoint segments on the second.  The
    same pair on every call keeps the overlap of the two alike from call to call."""
    idx = torch.device(device).index
    idx = torch.cuda.current_device() if idx is None else idx
    if idx not in _STREAMS:
        _STREAMS[idx] = (torch.cuda.Stream(device=idx), torch.cuda.Stream(device=idx))
    return _STREAMS[idx]


class _Stepper:
    """Fixed-dt forward steps into caller-owned tensors (no host synchronisation)."""

    def __init__(self, desc, device):
        self.lib = _lib.load()
        self.desc = desc
        self.grid = DeviceGrid(desc, device)  # owns a zeroed control block for the CFL atomics

    def step(self, src, dst, dt, stream=None):
        check(self.lib.claw2d_step(C.byref(self.desc), C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()),
                                   float(dt), self.grid.ctrl_ptr, _stream_handle(stream)), "claw2d_step")


def cfl_dt_sequence(desc, q0_planes, nsteps, dt0=None):
    """The dt of each of the first `nsteps` accepted steps of Solver.evolve_to_time from q0 (one
    host read of the CFL number per attempt).  Rejected attempts are retried with the smaller dt,
    as pyclaw does."""
    g = DeviceGrid(desc, q0_planes.device)
    g.set_planes(q0_planes)
    g.init_ctrl(0.0, 1.0)  # zeroes the speed accumulators; its dt is not used by step()
    dt = float(dt0) if dt0 is not None else 0.1
    dts = []
    attempts = 0
    while len(dts) < nsteps:
        attempts += 1
        if attempts > 10 * nsteps + 100:
            raise RuntimeError("cfl_dt_sequence: too many rejected steps")
        cfl, _ = g.step(dt)
        if not np.isfinite(cfl):
            raise FloatingPointError("non-finite CFL number in the forward run")
        if cfl <= desc.cfl_max:
            dts.append(dt)
        else:
            g.cur = 1 - g.cur  # reject: q^n is still in the other buffer
        if cfl > 0.0:
            dt = min(desc.dt_max, dt * desc.cfl_desired / cfl)
    return np.array(dts)


class FrozenDtEvolve(torch.autograd.Function):
    """(samples, qN) = N fixed-dt steps from q0 with gauge samples; gradient by the adjoint."""

    @staticmethod
    def forward(ctx, q0, desc, dts, gauges, sample_every, checkpoint_every):
        if q0.dtype != torch.float64 or not q0.is_cuda:
            raise ValueError("q0 must be a float64 device tensor")
        if desc.solver != 0 or desc.dtype != _lib.FP64:
            raise ValueError("the adjoint supports shallow water in FP64")
        dts = [float(x) for x in dts]
        n = len(dts)
        stepper = _Stepper(desc, q0.device)
        a = q0.detach().contiguous().clone()
        b = torch.empty_like(a)
        ckpts = {}
        samples = []
        flat_idx = gauges.to(q0.device)
        # the states of the last segment are kept as the forward makes them: the backward reverses
        # that segment first, and would otherwise recompute it from its checkpoint at once
        last0 = ((n - 1) // checkpoint_every) * checkpoint_every if n else 0
        tail = []
        for k in range(n):
            if k % checkpoint_every == 0:
                ckpts[k] = a.clone()
                if k == last0:
                    tail.append(ckpts[k])
            elif k > last0:
                tail.append(a.clone())
            stepper.step(a, b, dts[k])
            a, b = b, a
            if (k + 1) % sample_every == 0:
                samples.append(a.view(3, -1)[:, flat_idx])
        out = torch.stack(samples) if samples else a.new_zeros((0, 3, flat_idx.numel()))
        ctx.desc, ctx.dts, ctx.sample_every, ctx.checkpoint_every = desc, dts, sample_every, checkpoint_every
        ctx.ckpts, ctx.flat_idx, ctx.stepper, ctx.tail = ckpts, flat_idx, stepper, tail
        return out, a.clone()

    @staticmethod
    def backward(ctx, g_samples, g_qN):
        desc, dts = ctx.desc, ctx.dts
        n = len(dts)
        K, S = ctx.checkpoint_every, ctx.sample_every
        stepper, flat_idx = ctx.stepper, ctx.flat_idx
        shape = (3, desc.my, desc.mx)
        dev = flat_idx.device
        lam = torch.zeros(shape, dtype=torch.float64, device=dev) if g_qN is None else \
            g_qN.to(torch.float64).contiguous().clone()
        ws = adjoint_workspace(desc, dev)
        gs = None if g_samples is None else g_samples.to(torch.float64)
        # the reverse steps call the library directly with the constant arguments bound once: the
        # tensors are this function's own (float64, contiguous, (3, my, mx)), so adjoint_step's
        # per-call checks would only add host time to a host-bound loop
        lib = _lib.load()
        rev = lib.claw2d_adjoint_step
        # Segments are reversed last to first.  The states of the next segment are recomputed from
        # its checkpoint on a second stream while this segment's reverse steps run on the first:
        # at grid sizes where one step fills a fraction of the device the two overlap.  Two
        # buffers of K - 1 states alternate between the segments; events order each buffer's
        # refill after the reverse steps that read it, and each segment's reverse steps after its
        # recompute.  The caller's stream waits for both at the end.
        main = torch.cuda.current_stream(dev)
        rs, side = _stream_pair(dev)
        rs.wait_stream(main)
        side.wait_stream(main)  # checkpoints and the forward's last steps
        dref, wsp, lamp = C.byref(desc), C.c_void_p(ws.data_ptr()), C.c_void_p(lam.data_ptr())
        sh = _stream_handle(rs)

        def add_direct(lam, step_index):
            # samples taken after step `step_index` (i.e. of q^{step_index})
            if gs is not None and step_index > 0 and step_index % S == 0 and step_index // S - 1 < gs.shape[0]:
                lam.view(3, -1).index_add_(1, flat_idx, gs[step_index // S - 1])

        order = list(reversed(sorted(ctx.ckpts.keys())))
        have_tail = bool(ctx.tail)
        first = 1 if have_tail else 0  # the first segment to recompute (in reverse order)
        # one buffer per segment to recompute, at most two, taken in turn
        bufs = [[torch.empty(shape, dtype=torch.float64, device=dev) for _ in range(K - 1)]
                for _ in range(min(2, len(order) - first))]
        ready = [None] * len(order)
        free = [None, None]

        def span(i):
            return order[i], min(order[i] + K, n)

        def buf(i):  # buffer index of segment i >= first
            return (i - first) % 2

        def states_of(i):  # q^{s0} .. q^{s1-1} of segment i (in reverse order)
            s0, s1 = span(i)
            if i < first:
                return ctx.tail
            return [ctx.ckpts[s0]] + bufs[buf(i)][:s1 - s0 - 1]

        def recompute(i):  # enqueues segment i's recompute on `side`, one step per next()
            s0, s1 = span(i)
            st = states_of(i)
            if free[buf(i)] is not None:
                side.wait_event(free[buf(i)])
            for k in range(s0, s1 - 1):
                stepper.step(st[k - s0], st[k - s0 + 1], dts[k], stream=side)
                yield
            ready[i] = torch.cuda.Event()
            ready[i].record(side)

        with torch.cuda.stream(rs):
            add_direct(lam, n)
            if order and first == 0:
                for _ in recompute(0):
                    pass
            for i in range(len(order)):
                s0, s1 = span(i)
                nxt = recompute(i + 1) if i + 1 < len(order) else iter(())
                if ready[i] is not None:
                    rs.wait_event(ready[i])
                st = states_of(i)
                for k in range(s1 - 1, s0 - 1, -1):
                    rc = rev(dref, C.c_void_p(st[k - s0].data_ptr()), lamp, lamp, wsp, dts[k], sh)
                    if rc:
                        check(rc, "claw2d_adjoint_step")
                    add_direct(lam, k)
                    next(nxt, None)
                for _ in nxt:
                    pass
                if i >= first:
                    free[buf(i)] = torch.cuda.Event()
                    free[buf(i)].record(rs)
        main.wait_stream(rs)
        main.wait_stream(side)
        ctx.tail = None
        return lam, None, None, None, None, None


def evolve_frozen(q0, desc, dts, gauges, sample_every=10, checkpoint_every=20):
    """Differentiable frozen-dt run (see the module docstring)."""
    g = torch.as_tensor(gauges, dtype=torch.long)
    return FrozenDtEvolve.apply(q0, desc, list(dts), g, int(sample_every), int(checkpoint_every))
