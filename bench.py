"""Throughput of the fused 2D unsplit step on one MI355X.

Headline workload: 2D shallow water, Roe + entropy fix, unsplit, transverse_waves=2,
MC limiter, CFL 0.9, periodic BCs, 4096x4096 cells on [-2.5,2.5]^2, initial h = 1 + 16 seeded
Gaussian humps, hu = hv = 0.  One "step" = one accepted time step of the whole grid (fused step
kernel + device-side CFL/dt controller), state resident in HBM.

Prints one JSON line: cell-updates/s (FP64), roofline of the step kernel, and the other lines
(FP32, split step, batched grids, Euler, bathymetry, FP32-vs-FP64 error, adjoint cost, CPU
baseline), each of which can be switched off with its --no-* flag.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


# MI355X peak engine clock: a wave64 VALU instruction occupies one of a CU's 4 SIMDs for 4 cycles
VALU_CLOCK_GHZ = 2.4


def profiled_value(n, key):
    """A `key: value` line of the committed rocprofv3 summary of the FP64 step kernel at this size
    (profiles/step2_fused_swe_fp64_NxN.md), or None."""
    path = os.path.join(ROOT, "profiles", "step2_fused_swe_fp64_%dx%d.md" % (n, n))
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def profiled_traffic(n):
    """HBM bytes per launch of the FP64 step kernel from the committed rocprofv3 summary (FETCH_SIZE
    and WRITE_SIZE scaled by a copy-kernel calibration), if one exists for this size."""
    return profiled_value(n, "traffic_bytes_per_launch")


def valu_ceiling(n, cells, per_attempt_s):
    """The VALU issue ceiling of the FP64 step kernel beside the HBM roofline: the kernel's wave-level
    VALU instructions per launch (SQ_INSTS_VALU of the committed profile) issued at one per 4 cycles
    on every SIMD at the peak clock, against the measured launch time."""
    import torch
    wave_instr = profiled_value(n, "valu_wave_instructions_per_launch")
    if wave_instr is None:
        return None
    simds = 4 * torch.cuda.get_device_properties(0).multi_processor_count
    ceiling_s = wave_instr * 4.0 / simds / (VALU_CLOCK_GHZ * 1e9)
    return {"wave_instructions_per_launch": wave_instr, "lane_instructions_per_cell": wave_instr * 64.0 / cells,
            "issue_ceiling_ms": ceiling_s * 1e3, "frac_of_issue_ceiling": ceiling_s / per_attempt_s,
            "clock_ghz": VALU_CLOCK_GHZ, "simds": simds}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=500)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--prewarm", type=float, default=0.5,
                   help="seconds of untimed steps of each kernel before its counted warm-up")
    p.add_argument("--n", type=int, default=4096, help="cells per side")
    p.add_argument("--no-tune", action="store_true",
                   help="skip the tile-shape autotuner (profiling runs: its candidates would share "
                        "the step kernel's name in a trace); the shape is --shape or the default")
    p.add_argument("--shape", default=None, metavar="SEG_H,WG_WAVES,ORDER",
                   help="tile shape of the headline FP64 kernel when not tuned")
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--no-fp32", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-head-steps", type=int, default=2,
                   help="oracle steps at the headline size on one thread (several s each)")
    p.add_argument("--cpu-head-mt-steps", type=int, default=10,
                   help="oracle steps at the headline size on all threads")
    p.add_argument("--cpu-n", type=int, default=640)
    p.add_argument("--cpu-steps", type=int, default=100)
    p.add_argument("--cpu-mt-n", type=int, default=1024)
    p.add_argument("--cpu-mt-steps", type=int, default=200)
    p.add_argument("--no-batch", action="store_true")
    p.add_argument("--batch-grids", type=int, default=512)
    p.add_argument("--batch-steps", type=int, default=200)
    p.add_argument("--seq-steps", type=int, default=10)
    p.add_argument("--no-split", action="store_true")
    p.add_argument("--no-large", action="store_true", help="skip the 8192^2 FP64/FP32 line")
    p.add_argument("--large-n", type=int, default=8192)
    p.add_argument("--large-steps", type=int, default=100)
    p.add_argument("--no-euler", action="store_true")
    p.add_argument("--euler-n", type=int, default=2048)
    p.add_argument("--euler-steps", type=int, default=500)
    p.add_argument("--no-bathy", action="store_true")
    p.add_argument("--bathy-steps", type=int, default=500)
    p.add_argument("--no-precision", action="store_true")
    p.add_argument("--precision-sizes", default="4096,8192")
    p.add_argument("--precision-steps", type=int, default=1000)
    p.add_argument("--no-adjoint", action="store_true")
    p.add_argument("--adjoint-n", type=int, default=512)
    p.add_argument("--adjoint-steps", type=int, default=400)
    p.add_argument("--adjoint-fd", type=int, default=16, help="finite-difference directions checked")
    p.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                   help="only run a plain device copy of BYTES bytes 5 times (calibrates the "
                        "FETCH_SIZE/WRITE_SIZE counters) and exit")
    return p.parse_args()


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def copy_probe(nbytes):
    import torch
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "repeats": 5}), flush=True)


def measured_copy_gbs(nbytes=1 << 31, reps=5):
    """Device-to-device copy bandwidth of this GPU (read + write bytes per second, median of reps
    copies of nbytes), the measured second denominator of the HBM roofline."""
    import numpy as np
    import torch
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    s = torch.cuda.current_stream()
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        dst.copy_(src)
        e1.record(s)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    del src, dst
    return 2.0 * n * 8 / (float(np.median(times)) * 1e-3) / 1e9


def prewarm(g, q, dt0, seconds, aux=None):
    """Runs the grid's own step for at least `seconds` of wall time (clocks and caches settle on
    the kernel about to be timed), then restores the initial state, so the counted warm-up and the
    timed steps start from q as if nothing had run.  Returns the seconds spent."""
    import torch
    t0 = time.perf_counter()
    g.init_ctrl(0.0, dt0, target_steps=10 ** 9)
    while time.perf_counter() - t0 < seconds:
        g.advance(50)
        torch.cuda.synchronize()
    el = time.perf_counter() - t0
    g.upload(q)
    if aux is not None:
        g.upload_aux(aux)
    torch.cuda.synchronize()
    return el


def run_gpu(n, dtype_id, steps, warmup, seed, split=False, prewarm_s=0.0, tune=True, shape=None):
    """Times `steps` evolve_to_time attempts of the headline workload.  tune: the tile shape
    (rows per wavefront, wavefronts per workgroup, launch order) is chosen by the autotuner,
    measured here on this workload (the library default is one of the candidates)."""
    import torch
    from clawpack_amd import autotune, problems
    from clawpack_amd.device import DeviceGrid, make_desc

    q, dx, dy = problems.gaussian_humps(n, n, seed)
    # the FP32 state stores the depth as h - 1 (the humps' still-water level), as the precision
    # leg measures it
    d = make_desc(0, dtype_id, n, n, dx, dy, 3, 3, order=2, transverse_waves=2, mthlim=(4, 4, 4),
                  bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0, depth_ref=1.0 if dtype_id == 1 else 0.0)
    g = DeviceGrid(d, "cuda:0", split=split)
    g.upload(q)
    warmed = prewarm(g, q, 1e-4, prewarm_s) if prewarm_s > 0 else 0.0
    shape_ms = None
    if tune:  # measured after the pre-warm, with the clocks settled on this kernel
        shape, shape_ms = autotune.tune_grid(d, q, 1e-4, steps=30, rounds=3, use_cache=False, split=split)
        torch.cuda.empty_cache()
    if shape is None:
        shape = {"seg_h": 0, "wg_waves": 4, "launch_order": 0}  # the library's default shape
    g.desc = autotune.apply_shape(type(d).from_buffer_copy(bytes(d)), shape)
    del q
    # dt from the initial wave speed, as after pyclaw's first accepted step
    g.init_ctrl(0.0, 1e-4, target_steps=warmup + steps + 10**9)
    g.advance(warmup)
    torch.cuda.synchronize()
    c0 = g.read_ctrl()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(s)
    g.advance(steps)
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    c1 = g.read_ctrl()
    ms = e0.elapsed_time(e1)
    accepted = int(c1.numsteps - c0.numsteps)
    attempts = int(c1.attempts - c0.attempts)
    assert c1.status == 0, "device status %d" % c1.status
    return dict(ms=ms, wall=wall, accepted=accepted, attempts=attempts, cells=n * n,
                t=c1.t, dt=c1.dt, cfl=c1.cfl, prewarm_s=warmed, tile_shape=shape, tile_shape_ms=shape_ms)


def run_batch(ngrids, dtype_id, steps, seq_steps, seed):
    """Batched workload: ngrids independent periodic hump grids (sides U{64..512}), each with its
    own CFL-0.9 dt; `steps` attempts as one replayed HIP graph vs one-grid-at-a-time launches.
    The rows per batched work item come from the autotuner, measured here on this workload (64,
    the library default, is one of the candidates)."""
    import torch
    from clawpack_amd import autotune, problems
    from clawpack_amd.batch import BatchedGrids
    from clawpack_amd.device import make_desc

    grids = problems.batch_workload(ngrids, seed)
    d = make_desc(0, dtype_id, 8, 8, 1.0, 1.0, 3, 3, order=2, transverse_waves=2, mthlim=(4, 4, 4),
                  bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0)
    seg, tune_ms = autotune.tune_batch(d, grids, candidates=(16, 32, 48, 64, 96, 128, 192),
                                       attempts=5, use_cache=False)
    out = {"autotune": {"seg_h": seg, "ms_per_attempt": tune_ms}}
    s = torch.cuda.current_stream()
    for mode in ("graph", "sequential"):
        b = BatchedGrids(d, grids, seg_h=seg)
        b.init(0.0, 1e-3, target_steps=10 ** 9)
        b.advance(3)  # warm-up (first attempts settle each grid's dt)
        n_att = steps if mode == "graph" else seq_steps
        h = b.make_graph(n_att) if mode == "graph" else None
        torch.cuda.synchronize()
        c0 = b.read_ctrls()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        if mode == "graph":
            b.replay(h)
        else:
            b.advance_sequential(n_att)
        e1.record(s)
        torch.cuda.synchronize()
        c1 = b.read_ctrls()
        ms = e0.elapsed_time(e1)
        upd = sum(b.shapes[g][0] * b.shapes[g][1] * (c1[g].numsteps - c0[g].numsteps)
                  for g in range(b.n))
        assert all(c.status == 0 for c in c1)
        out[mode] = {"ms": ms, "attempts": n_att, "cell_updates_per_s": upd / (ms * 1e-3)}
        del b
    out["cells"] = sum(q.shape[1] * q.shape[2] for q, _, _ in grids)
    out["speedup_vs_sequential"] = out["graph"]["cell_updates_per_s"] / out["sequential"]["cell_updates_per_s"]
    return out


def fp32_vs_fp64(make, q, dt, steps, comps, aux=None, split=False):
    """Error of the FP32 path against the FP64 path after `steps` fixed-dt steps from the same
    state q (make(dtype_id) -> descriptor): per component, L1 = sum|dq_m| / sum|q_m| and
    Linf = max|dq_m| / max|q_m|; the worst of each over the components leads the record."""
    import torch
    from clawpack_amd.device import DeviceGrid
    planes = {}
    for dtype_id in (0, 1):
        g = DeviceGrid(make(dtype_id), "cuda:0", split=split)
        g.upload(q)
        if aux is not None:
            g.upload_aux(aux)
        g.init_ctrl(0.0, dt, target_steps=steps)
        for _ in range(steps):
            g.enqueue_step(dt)
        torch.cuda.synchronize()
        c = g.read_ctrl()
        assert c.status == 0 and c.cfl <= 1.0, "status %d cfl %g" % (c.status, c.cfl)
        planes[dtype_id] = g.planes().to(torch.float64).clone()
        planes[dtype_id][0] += g.desc.depth_ref  # a stored depth deviation back to h
        del g
    a, b = planes[0], planes[1]
    rec = {}
    for m, comp in enumerate(comps):
        diff = (b[m] - a[m]).abs()
        rec[comp] = {"L1_rel": float(diff.sum() / a[m].abs().sum()),
                     "Linf_rel": float(diff.max() / a[m].abs().max())}
    out = {"max_component_Linf_rel": max(r["Linf_rel"] for r in rec.values()),
           "max_component_L1_rel": max(r["L1_rel"] for r in rec.values()),
           "per_component": rec, "steps": steps, "dt": dt,
           "Linf_rel_to_max_q": float((b - a).abs().max() / a.abs().max())}
    return out


def run_euler(n, steps, warmup, seed, prewarm_s=0.0, err_steps=500):
    """Euler gas dynamics: four-quadrant Riemann problem (seeded), euler_4wave with entropy fix,
    unsplit, MC, extrapolation BCs, n x n on [0,1]^2, CFL 0.9 device-controlled dt."""
    import torch
    from clawpack_amd import problems
    from clawpack_amd.device import DeviceGrid, make_desc

    out = {}
    q, dx, dy = problems.euler_quadrants(n, n, seed)
    for dtype_id, name in ((0, "fp64"), (1, "fp32")):
        d = make_desc(1, dtype_id, n, n, dx, dy, 4, 4, order=2, transverse_waves=2, mthlim=(4, 4, 4, 4),
                      bc_lower=(1, 1), bc_upper=(1, 1), gamma=1.4)
        g = DeviceGrid(d, "cuda:0")
        g.upload(q)
        if prewarm_s > 0:
            prewarm(g, q, 1e-5, prewarm_s)
        g.init_ctrl(0.0, 1e-5, target_steps=10 ** 9)
        g.advance(warmup)
        torch.cuda.synchronize()
        c0 = g.read_ctrl()
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        g.advance(steps)
        e1.record(s)
        torch.cuda.synchronize()
        c1 = g.read_ctrl()
        assert c1.status == 0, "device status %d" % c1.status
        acc = int(c1.numsteps - c0.numsteps)
        att = int(c1.attempts - c0.attempts)
        ms = e0.elapsed_time(e1)
        bpc = 64.0 if dtype_id == 0 else 32.0
        out[name] = {"cell_updates_per_s": n * n * acc / (ms * 1e-3), "ms_per_step": ms / acc,
                     "accepted": acc, "attempts": att, "t": c1.t,
                     "roofline_frac": bpc * n * n * att / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if dtype_id == 0:
            dt_err = 0.8 * c0.dt  # the controller's dt after the warm-up steps, with CFL margin
        del g
    out["fp32"]["ratio_vs_fp64"] = out["fp32"]["cell_updates_per_s"] / out["fp64"]["cell_updates_per_s"]
    out["fp32"]["error_vs_fp64"] = fp32_vs_fp64(
        lambda t: make_desc(1, t, n, n, dx, dy, 4, 4, order=2, transverse_waves=2,
                            mthlim=(4, 4, 4, 4), bc_lower=(1, 1), bc_upper=(1, 1), gamma=1.4),
        q, dt_err, err_steps, ("rho", "rho_u", "rho_v", "E"))
    out["workload"] = "euler_4wave_efix_unsplit_mc_extrap_quadrants_seed%d_%dx%d" % (seed, n, n)
    return out


def run_bathy(n, steps, warmup, prewarm_s=0.0, err_steps=500):
    """Shallow water over bathymetry (f-wave solver, aux plane b): the 2D sill problem on
    [-1,1]^2 at n x n, g = 9.81, MC, transverse_waves 2, x extrap / y wall, CFL 0.9 device dt;
    unsplit and split, FP64 and FP32.  Algorithmic bytes per cell-update: q read + q write + the
    aux plane read, 7 values (56 B FP64, 28 B FP32)."""
    import torch
    from clawpack_amd import problems
    from clawpack_amd.device import DeviceGrid, make_desc

    q, aux, dx, dy = problems.sill(n, n)
    out = {}
    for split in (False, True):
        for dtype_id, name, bpc in ((0, "fp64", 56.0), (1, "fp32", 28.0)):
            d = make_desc(2, dtype_id, n, n, dx, dy, 3, 3, order=2, transverse_waves=2,
                          mthlim=(4, 4, 4), bc_lower=(1, 3), bc_upper=(1, 3), grav=9.81,
                          dry_tolerance=1e-3)
            g = DeviceGrid(d, "cuda:0", split=split)
            g.upload(q)
            g.upload_aux(aux)
            if prewarm_s > 0:
                prewarm(g, q, 1e-5, prewarm_s, aux=aux)
            g.init_ctrl(0.0, 1e-5, target_steps=10 ** 9)
            g.advance(warmup)
            torch.cuda.synchronize()
            c0 = g.read_ctrl()
            s = torch.cuda.current_stream()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            g.advance(steps)
            e1.record(s)
            torch.cuda.synchronize()
            c1 = g.read_ctrl()
            assert c1.status == 0, "device status %d" % c1.status
            acc = int(c1.numsteps - c0.numsteps)
            att = int(c1.attempts - c0.attempts)
            ms = e0.elapsed_time(e1)
            out[("split_" if split else "unsplit_") + name] = {
                "cell_updates_per_s": n * n * acc / (ms * 1e-3), "ms_per_step": ms / acc,
                "accepted": acc, "attempts": att, "t": c1.t,
                "roofline_frac": bpc * n * n * att / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
            if dtype_id == 0:
                dt_err = 0.8 * c0.dt
            del g
        out[("split_" if split else "unsplit_") + "fp32"]["error_vs_fp64"] = fp32_vs_fp64(
            lambda t: make_desc(2, t, n, n, dx, dy, 3, 3, order=2, transverse_waves=2,
                                mthlim=(4, 4, 4), bc_lower=(1, 3), bc_upper=(1, 3), grav=9.81,
                                dry_tolerance=1e-3),
            q, dt_err, err_steps, ("h", "hu", "hv"), aux=aux, split=split)
    out["workload"] = "swe_bathymetry_fwave_sill_mc_extrap_wall_%dx%d" % (n, n)
    return out


def run_precision(sizes, steps, seed):
    """FP32 beside FP64 on the periodic 16-hump shallow-water workload, same fixed dt sequence
    (CFL about 0.8 of the initial wave speed), compensated FP64 mass sum after every step.
    Reports L1 / Linf of the FP32 state against the FP64 state per component (the FP64 HIP path,
    parity-proven against the oracle, is the reference at these sizes) and the mass drift."""
    import math
    import torch
    from clawpack_amd import problems
    from clawpack_amd.device import DeviceGrid, make_desc

    out = {}
    for n in sizes:
        q, dx, dy = problems.gaussian_humps(n, n, seed)
        dt = 0.8 * min(dx, dy) / (math.sqrt(float(q[0].max())) + 0.5)
        grids, rec = {}, {}
        # fp32: the FP32 path as offered, depth stored as h - 1 (the still-water level of the
        # humps); fp32_stored_h: the same kernels storing h itself, for comparison
        for dtype_id, name, ref in ((0, "fp64", 0.0), (1, "fp32", 1.0), (1, "fp32_stored_h", 0.0)):
            d = make_desc(0, dtype_id, n, n, dx, dy, 3, 3, order=2, transverse_waves=2,
                          mthlim=(4, 4, 4), bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0, dt_variable=False,
                          depth_ref=ref)
            g = DeviceGrid(d, "cuda:0")
            g.upload(q)
            g.init_ctrl(0.0, dt, target_steps=steps)
            m0 = g.mass_sum(0)
            drift = 0.0
            for k in range(steps):
                g.advance(1)
                g.sync_cur()
                drift = max(drift, abs(g.mass_sum(0) - m0) / m0)
            c = g.read_ctrl()
            assert c.status == 0 and c.numsteps == steps, "status %d steps %d" % (c.status, c.numsteps)
            grids[name] = g
            rec[name] = {"mass_rel_drift_max": drift, "depth_ref": ref}
        a = grids["fp64"].planes()
        for name, key in (("fp32", "fp32_vs_fp64"), ("fp32_stored_h", "fp32_stored_h_vs_fp64")):
            b = grids[name].planes().to(torch.float64)
            b[0] += grids[name].desc.depth_ref  # stored h - depth_ref -> h
            err = {}
            qmax = float(a.abs().max())
            dall = (b - a).abs()
            for m, comp in enumerate(("h", "hu", "hv")):
                diff = dall[m]
                err[comp] = {"L1_rel": float(diff.sum() / a[m].abs().sum()),
                             "Linf_rel": float(diff.max() / a[m].abs().max())}
            # the state as a whole: max |dq| / max |q| and sum |dq| / sum |q| over all components
            err["q"] = {"Linf_rel_to_max_q": float(dall.max()) / qmax,
                        "L1_rel": float(dall.sum() / a.abs().sum())}
            rec[key] = {
                "max_component_Linf_rel": max(err[c]["Linf_rel"] for c in ("h", "hu", "hv")),
                "max_component_L1_rel": max(err[c]["L1_rel"] for c in ("h", "hu", "hv")),
                "per_component": {c: err[c] for c in ("h", "hu", "hv")}, "whole_state": err["q"]}
            del b
        rec["steps"] = steps
        rec["dt"] = dt
        out["%dx%d" % (n, n)] = rec
        del grids, a
        torch.cuda.empty_cache()
    return out


def run_adjoint(n, nsteps, seed, fd_dirs):
    """Inversion workload: SWE n x n FP64 periodic, nsteps frozen-dt steps from 1 + eta0, h sampled
    at 64 gauges every 10 steps, checkpoints every 20.  Times the forward (steps + checkpoint
    stores + samples) and the backward (segment recompute + reverse steps) and checks the
    gradient against central differences (eps 1e-6) along fd_dirs seeded unit directions, for the
    MC-limited method and for the unlimited second-order method (no limiter kinks, so the loss is
    smooth and the finite differences converge to the directional derivative itself)."""
    import numpy as np
    import torch
    from clawpack_amd import problems
    from clawpack_amd.adjoint import cfl_dt_sequence, evolve_frozen
    from clawpack_amd.device import make_desc

    eta_true, gauges, dx, dy = problems.inversion_workload(n, n, seed=4)
    g = torch.tensor(gauges)
    rng = np.random.default_rng(seed)
    eta0 = 0.5 * eta_true + 1e-3 * rng.standard_normal(eta_true.shape)
    dirs = []
    for _ in range(fd_dirs):
        v = rng.standard_normal(eta0.shape)
        dirs.append(v / np.linalg.norm(v))
    out = {"workload": "swe_inversion_periodic_mc_%dx%d_%dsteps_64gauges_fp64" % (n, n, nsteps),
           "checkpoint_every": 20}
    for lim, name in (((4, 4, 4), "mc"), ((0, 0, 0), "unlimited")):
        d = make_desc(0, 0, n, n, dx, dy, 3, 3, order=2, transverse_waves=2, mthlim=lim,
                      bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0)
        q_true = torch.tensor(np.stack([1.0 + eta_true, 0 * eta_true, 0 * eta_true]), device="cuda")
        dts = cfl_dt_sequence(d, q_true, nsteps)
        with torch.no_grad():
            obs = evolve_frozen(q_true, d, dts, g, 10, 20)[0][:, 0, :]

        def loss_of(eta):
            z = torch.zeros_like(eta)
            samples, _ = evolve_frozen(torch.stack([1.0 + eta, z, z]), d, dts, g, 10, 20)
            return ((samples[:, 0, :] - obs) ** 2).sum()

        fwd, bwd = [], []
        # the first two runs are warm-up (first launches, the caching allocator's growth for the
        # backward's two segment buffers); the median of the next six is reported
        for rep in range(8 if name == "mc" else 1):
            eta = torch.tensor(eta0, device="cuda", requires_grad=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            loss = loss_of(eta)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if rep >= 2:
                fwd.append(t1 - t0)
                bwd.append(t2 - t1)
        grad = eta.grad.cpu().numpy()
        gnorm = float(np.linalg.norm(grad))
        rel_dir, rel_norm = [], []
        eps = 1e-6
        for v in dirs:
            with torch.no_grad():
                lp = float(loss_of(torch.tensor(eta0 + eps * v, device="cuda")))
                lm = float(loss_of(torch.tensor(eta0 - eps * v, device="cuda")))
            fd = (lp - lm) / (2 * eps)
            ad = float((grad * v).sum())
            rel_dir.append(abs(fd - ad) / abs(ad))
            rel_norm.append(abs(fd - ad) / gnorm)
        rec = {"loss": float(loss.detach()),
               "fd": {"eps": eps, "directions": fd_dirs,
                      "median_rel_err_vs_directional": float(np.median(rel_dir)) if rel_dir else None,
                      "max_rel_err_vs_directional": max(rel_dir) if rel_dir else None,
                      "max_abs_err_over_grad_norm": max(rel_norm) if rel_norm else None}}
        if name == "unlimited" and dirs:
            # Richardson-extrapolated central differences, (4 D(e/2) - D(e)) / 3 with e = 1e-3:
            # truncation O(e^4) and a rounding error of about |L| 1e-16 / e, both far below those
            # of the plain e = 1e-6 difference on this smooth loss
            er, rel_r = 1e-3, []
            for v in dirs:
                with torch.no_grad():
                    def D(e):
                        return (float(loss_of(torch.tensor(eta0 + e * v, device="cuda")))
                                - float(loss_of(torch.tensor(eta0 - e * v, device="cuda")))) / (2 * e)
                    fd = (4.0 * D(er / 2) - D(er)) / 3.0
                ad = float((grad * v).sum())
                rel_r.append(abs(fd - ad) / abs(ad))
            rec["fd_richardson"] = {"eps": er, "directions": len(dirs),
                                    "median_rel_err_vs_directional": float(np.median(rel_r)),
                                    "max_rel_err_vs_directional": max(rel_r)}
        if fwd:
            f, b = float(np.median(fwd)), float(np.median(bwd))
            rec.update(forward_ms=f * 1e3, backward_ms=b * 1e3, backward_over_forward=b / f,
                       total_over_forward=(f + b) / f)
        if name == "mc":
            out.update(rec)
        else:
            out["unlimited"] = rec
    return out


def cpu_baseline(n, steps, seed, threads=1):
    """The C oracle (restatement of step2/flux2/rpn2/rpt2) on the host: same workload at n x n,
    `steps` accepted steps, `threads` OpenMP threads (slices of a sweep in three colour classes)."""
    from clawpack_amd import problems
    from oracle import corc
    q, dx, dy = problems.gaussian_humps(n, n, seed)
    d = corc.make_desc("swe", n, n, dx, dy, grav=1.0)
    c = corc.make_ctrl(bc_lower=(2, 2), bc_upper=(2, 2))
    corc.set_threads(threads)
    try:
        corc.evolve(d, c, q, nsteps=1, dt=1e-4)  # warm the library and page tables
        t0 = time.perf_counter()
        _, st = corc.evolve(d, c, q, nsteps=steps, dt=1e-4)
        el = time.perf_counter() - t0
    finally:
        corc.set_threads(1)
    return n * n * st["numsteps"] / el, el


def host_threads():
    """CPU share for the multi-core baseline: OMP_NUM_THREADS if set, else the core count, at most 16."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return max(1, min(16, n if n > 0 else (os.cpu_count() or 1)))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: a multi-GPU domain decomposition is not implemented; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.init()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    forced = None
    if a.shape:
        h, w, o = (int(x) for x in a.shape.split(","))
        forced = {"seg_h": h, "wg_waves": w, "launch_order": o}
    r64 = run_gpu(a.n, 0, a.steps, a.warmup, a.seed, prewarm_s=a.prewarm, tune=not a.no_tune,
                  shape=forced)
    progress("fp64 step: %.4f ms/attempt" % (r64["ms"] / r64["attempts"]))
    ms_step = r64["ms"] / r64["accepted"]
    value = r64["cells"] * r64["accepted"] / (r64["ms"] * 1e-3)
    bytes_per_cell = 48.0
    per_attempt_s = r64["ms"] * 1e-3 / r64["attempts"]
    achieved = bytes_per_cell * r64["cells"] / per_attempt_s / 1e9
    out = {
        "metric": "2D SWE cell-updates/s FP64 & FP32, % HBM roofline; adjoint/forward cost",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "prewarm_s": r64["prewarm_s"],
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian humps)",
        "config": {"workload": "swe_roe_efix_unsplit_mc_periodic_humps16_%dx%d_fp64" % (a.n, a.n),
                   "cells": r64["cells"], "accepted_steps": r64["accepted"],
                   "attempts": r64["attempts"], "parallelism": "single GPU",
                   "tile_shape": r64["tile_shape"]},
        "autotune": {"fp64": r64["tile_shape_ms"]} if r64["tile_shape_ms"] else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(a.n),
                     "kernel": "step2_fused<double,SweP<double>,4,false,2> (one launch per attempt)",
                     "bytes_per_cell": bytes_per_cell},
    }
    out["roofline"]["valu"] = valu_ceiling(a.n, r64["cells"], per_attempt_s)
    copy_gbs = measured_copy_gbs()
    out["roofline"]["measured_copy_gbs"] = copy_gbs
    out["roofline"]["frac_of_measured_copy"] = achieved / copy_gbs
    if not a.no_fp32:
        r32 = run_gpu(a.n, 1, a.steps, a.warmup, a.seed, prewarm_s=a.prewarm, tune=not a.no_tune)
        v32 = r32["cells"] * r32["accepted"] / (r32["ms"] * 1e-3)
        progress("fp32 step: %.4f ms/attempt" % (r32["ms"] / r32["attempts"]))
        out["fp32"] = {"value": v32, "ms_per_step": r32["ms"] / r32["accepted"],
                       "ratio_vs_fp64": v32 / value,
                       "roofline_frac": 24.0 * v32 / 1e9 / HBM_PEAK_GBS,
                       "tile_shape": r32["tile_shape"], "depth_ref": 1.0}
        if out["autotune"] is not None:
            out["autotune"]["fp32"] = r32["tile_shape_ms"]
    if not a.no_large:
        lg = {"workload": "the headline workload at %dx%d (config 2's larger size)" % (a.large_n, a.large_n)}
        for dt_id, name, bpc in ((0, "fp64", 48.0), (1, "fp32", 24.0)):
            r = run_gpu(a.large_n, dt_id, a.large_steps, 5, a.seed, prewarm_s=a.prewarm, tune=False)
            lg[name] = {"cell_updates_per_s": r["cells"] * r["accepted"] / (r["ms"] * 1e-3),
                        "ms_per_step": r["ms"] / r["accepted"],
                        "roofline_frac": bpc * r["cells"] * r["attempts"] / (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
        lg["fp32"]["ratio_vs_fp64"] = lg["fp32"]["cell_updates_per_s"] / lg["fp64"]["cell_updates_per_s"]
        out["large"] = lg
        progress("8192 done")
    if not a.no_split:
        sp = {}
        for dt_id, name, bpc in ((0, "fp64", 48.0), (1, "fp32", 24.0)):
            r = run_gpu(a.n, dt_id, a.steps, a.warmup, a.seed, split=True, prewarm_s=a.prewarm,
                        tune=not a.no_tune)
            sp[name] = {"cell_updates_per_s": r["cells"] * r["accepted"] / (r["ms"] * 1e-3),
                        "ms_per_step": r["ms"] / r["accepted"], "tile_shape": r["tile_shape"],
                        "roofline_frac": bpc * r["cells"] * r["attempts"] / (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
        sp["workload"] = "same as the headline line with dimensional_split=True (step2ds, one pass per step)"
        out["split"] = sp
        progress("split done")
    if not a.no_batch:
        out["batch"] = {"grids": a.batch_grids, "fp64": run_batch(a.batch_grids, 0, a.batch_steps, a.seq_steps, a.seed),
                        "fp32": run_batch(a.batch_grids, 1, a.batch_steps, a.seq_steps, a.seed)}
        progress("batch done")
    if not a.no_euler:
        out["euler"] = run_euler(a.euler_n, a.euler_steps, 10, 1, prewarm_s=a.prewarm)
        progress("euler done")
    if not a.no_bathy:
        out["bathymetry"] = run_bathy(a.n, a.bathy_steps, a.warmup, prewarm_s=a.prewarm)
        progress("bathymetry done")
    if not a.no_precision:
        out["precision"] = run_precision([int(x) for x in a.precision_sizes.split(",")],
                                         a.precision_steps, a.seed)
        progress("precision done")
    if not a.no_adjoint:
        out["adjoint"] = run_adjoint(a.adjoint_n, a.adjoint_steps, a.seed, a.adjoint_fd)
        progress("adjoint done")
    if not a.no_cpu:
        nth = host_threads()
        rate_mt, el_mt = cpu_baseline(a.n, a.cpu_head_mt_steps, a.seed, nth)
        progress("cpu %d threads at the headline size done" % nth)
        rate1, el1 = cpu_baseline(a.n, a.cpu_head_steps, a.seed, 1)
        progress("cpu 1 thread at the headline size done")
        rate1s, el1s = cpu_baseline(a.cpu_n, a.cpu_steps, a.seed, 1)
        rate_mts, el_mts = cpu_baseline(a.cpu_mt_n, a.cpu_mt_steps, a.seed, nth)
        progress("cpu smaller samples done")
        out["cpu_baseline"] = {
            "value": rate_mt, "unit": "cell-updates/s", "cores": nth, "kind": "port",
            "sample": "C oracle (CPU restatement of step2/flux2/rpn2/rpt2, not the reference binary), "
                      "the headline workload (%dx%d, FP64) for %d steps on %d OpenMP threads "
                      "(%.1f s wall)" % (a.n, a.n, a.cpu_head_mt_steps, nth, el_mt),
            "single_core": {"value": rate1, "cores": 1,
                            "sample": "the headline workload (%dx%d) for %d steps (%.1f s)"
                                      % (a.n, a.n, a.cpu_head_steps, el1)},
            "smaller_samples": {
                "single_core": {"value": rate1s, "sample": "%dx%d for %d steps (%.1f s)"
                                % (a.cpu_n, a.cpu_n, a.cpu_steps, el1s)},
                "threads": {"value": rate_mts, "cores": nth, "sample": "%dx%d for %d steps (%.1f s)"
                            % (a.cpu_mt_n, a.cpu_mt_n, a.cpu_mt_steps, el_mts)}}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
