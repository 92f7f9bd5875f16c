"""Autotuner for the marching-strip decomposition (rows per work item, `seg_h`).

The step kernel's cost per wavefront is proportional to its rows; seg_h trades halo rows
(4 per segment) and per-segment warm-up against the number of wavefronts in flight and the
balance of the last round.  The library's default picks seg_h from the kernel's occupancy so
every wavefront is resident at once; this module measures candidates on the device with HIP
events and caches the winner in JSON keyed by device name, HIP runtime version, solver,
precision and size class (sides rounded up to powers of two).

Cache file: $CLAWPACK_AMD_TUNE_CACHE or ~/.cache/clawpack_amd/autotune.json.
"""
import json
import math
import os

import numpy as np
import torch

from . import _lib
from .device import DeviceGrid

DEFAULT_CANDIDATES = (0, 16, 24, 32, 48, 64, 96, 128, 192, 256)


def cache_path():
    p = os.environ.get("CLAWPACK_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "clawpack_amd", "autotune.json")


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table):
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _size_class(n):
    return 1 << max(4, int(math.ceil(math.log2(max(int(n), 1)))))


def key(desc, kind="grid", split=False):
    """Cache key: everything that selects a different kernel or launch shape (device, runtime,
    solver, precision, size class, split or unsplit step, order, transverse mode, the limiter of
    every wave, aux planes)."""
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
    hip = getattr(torch.version, "hip", None) or "unknown"
    return "|".join([kind, dev, "hip" + str(hip), "solver%d" % desc.solver,
                     "fp64" if desc.dtype == _lib.FP64 else "fp32",
                     "%dx%d" % (_size_class(desc.mx), _size_class(desc.my)),
                     "split" if split else "unsplit",
                     "order%d" % desc.order, "tw%d" % desc.transverse_waves,
                     "lim" + "".join(str(int(desc.mthlim[k])) for k in range(desc.num_waves)),
                     "aux%d" % desc.num_aux])


def time_seg_h(desc, q, seg_h, dt, steps=10, warmup=2, split=False):
    """ms per fixed-dt step of the single-grid kernel (split: the step2ds kernel) with
    desc.seg_h = seg_h."""
    d = type(desc).from_buffer_copy(bytes(desc))
    d.seg_h = int(seg_h)
    g = DeviceGrid(d, split=split)
    g.upload(q)
    g.init_ctrl(0.0, dt)
    for _ in range(warmup):
        g.enqueue_step(dt)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        g.enqueue_step(dt)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def tune_grid(desc, q, dt, candidates=DEFAULT_CANDIDATES, steps=10, use_cache=True, split=False):
    """Best seg_h for this descriptor (0 = the occupancy-derived default) and step form.  Returns
    (seg_h, table of ms per step by candidate)."""
    k = key(desc, split=split)
    table = _load() if use_cache else {}
    if use_cache and k in table:
        return int(table[k]["seg_h"]), table[k]["ms"]
    ms = {}
    for c in candidates:
        if c and c > desc.my:
            continue
        ms[str(c)] = time_seg_h(desc, q, c, dt, steps=steps, split=split)
    best = int(min(ms, key=ms.get))
    if use_cache:
        table = _load()
        table[k] = {"seg_h": best, "ms": ms}
        _store(table)
    return best, ms


def tune_batch(desc, grids, candidates=(16, 32, 48, 64, 128), attempts=5, use_cache=True):
    """Best rows per work item for a batch of grids (BatchedGrids); returns (seg_h, ms table)."""
    from .batch import BatchedGrids
    total_cells = sum(q.shape[1] * q.shape[2] for q, _, _ in grids)
    k = key(desc, kind="batch") + "|n%d|cells%d" % (len(grids), _size_class(total_cells))
    table = _load() if use_cache else {}
    if use_cache and k in table:
        return int(table[k]["seg_h"]), table[k]["ms"]
    ms = {}
    s = torch.cuda.current_stream()
    for c in candidates:
        b = BatchedGrids(desc, grids, seg_h=c)
        b.init(0.0, 1e-3, target_steps=10 ** 9)
        b.advance(2)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        b.advance(attempts)
        e1.record(s)
        torch.cuda.synchronize()
        ms[str(c)] = e0.elapsed_time(e1) / attempts
        del b
    best = int(min(ms, key=ms.get))
    if use_cache:
        table = _load()
        table[k] = {"seg_h": best, "ms": ms}
        _store(table)
    return best, ms


def tuned_seg_h(desc, q=None, dt=None, split=False):
    """Cached seg_h for desc and step form (tuning on a synthetic state of the right shape if
    needed)."""
    k = key(desc, split=split)
    table = _load()
    if k in table:
        return int(table[k]["seg_h"])
    if q is None:
        neq = desc.num_eqn
        q = np.ones((neq, desc.mx, desc.my), order="F")
        if neq == 4:
            q[3] = 2.5
    if dt is None:
        dt = 0.1 * min(desc.dx, desc.dy)
    return tune_grid(desc, q, dt, split=split)[0]
