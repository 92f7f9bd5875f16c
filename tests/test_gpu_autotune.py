"""Autotuner: measures seg_h candidates, caches the winner, and the tuned decomposition gives the
same result bit for bit (the work split never changes the arithmetic)."""
import json
import os

import numpy as np
import pytest

from clawpack_amd import problems

pytestmark = pytest.mark.gpu


def test_tune_and_cache(tmp_path, monkeypatch):
    from clawpack_amd import autotune
    from clawpack_amd.device import DeviceGrid, make_desc
    monkeypatch.setenv("CLAWPACK_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    n = 512
    q, dx, dy = problems.gaussian_humps(n, n, 1)
    d = make_desc(0, 0, n, n, dx, dy, 3, 3, bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0)
    best, ms = autotune.tune_grid(d, q, 1e-3, candidates=(0, 16, 64, 128), steps=5)
    print("tuned seg_h", best, ms)
    assert str(best) in ms and ms[str(best)] == min(ms.values())
    table = json.load(open(os.environ["CLAWPACK_AMD_TUNE_CACHE"]))
    assert list(table.values())[0]["seg_h"] == best
    assert autotune.tuned_seg_h(d) == best  # served from the cache
    outs = []
    for seg in (0, best, 16):
        dd = type(d).from_buffer_copy(bytes(d))
        dd.seg_h = seg
        g = DeviceGrid(dd)
        g.upload(q)
        g.init_ctrl(0.0, 1e-3)
        for _ in range(3):
            g.enqueue_step(1e-3)
        outs.append(g.download())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_tune_batch(tmp_path, monkeypatch):
    from clawpack_amd import autotune
    from clawpack_amd.device import make_desc
    monkeypatch.setenv("CLAWPACK_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    grids = problems.batch_workload(32, 4, nmin=64, nmax=256)
    d = make_desc(0, 1, 8, 8, 1.0, 1.0, 3, 3, bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0)
    best, ms = autotune.tune_batch(d, grids, candidates=(16, 64), attempts=3)
    print("batch seg_h", best, ms)
    assert best in (16, 64)
