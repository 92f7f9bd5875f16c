"""Autotuner: measures tile shapes (rows per wavefront, wavefronts per workgroup, launch order),
caches the winner, and every shape gives the same result bit for bit (the work split and launch
shape never change the arithmetic)."""
import collections
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
    shape, ms = autotune.tune_grid(d, q, 1e-3, candidates=(0, 16, 64, 128), steps=5)
    print("tuned shape", shape, ms)
    name = autotune.shape_name(shape["seg_h"], shape["wg_waves"], shape["launch_order"])
    dflt = autotune.shape_name(0, 4, 0)
    # every packing and order was measured at one height (the best of the first stage), and the
    # default shape beside them
    heights = collections.Counter(n.split("/")[0] for n in ms)
    h2, count = heights.most_common(1)[0]
    assert count >= len(autotune.SHAPES) and dflt in ms
    # the default stays unless the fastest of the second stage beats it by more than 1%
    stage2 = {n: t for n, t in ms.items() if n.startswith(h2 + "/") or n == dflt}
    assert name == dflt or (ms[name] == min(stage2.values()) and ms[name] < 0.99 * ms[dflt])
    table = json.load(open(os.environ["CLAWPACK_AMD_TUNE_CACHE"]))
    assert list(table.values())[0]["seg_h"] == shape["seg_h"]
    assert autotune.tuned_shape(d) == shape  # served from the cache
    outs = []
    for seg, wpb, order in ((0, 0, 0), (shape["seg_h"], shape["wg_waves"], shape["launch_order"]),
                            (16, 1, 1), (24, 2, 0)):
        g = DeviceGrid(autotune._with_shape(d, seg, wpb, order))
        g.upload(q)
        g.init_ctrl(0.0, 1e-3)
        for _ in range(3):
            g.enqueue_step(1e-3)
        outs.append(g.download())
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


def test_tune_batch(tmp_path, monkeypatch):
    from clawpack_amd import autotune
    from clawpack_amd.device import make_desc
    monkeypatch.setenv("CLAWPACK_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    grids = problems.batch_workload(32, 4, nmin=64, nmax=256)
    d = make_desc(0, 1, 8, 8, 1.0, 1.0, 3, 3, bc_lower=(2, 2), bc_upper=(2, 2), grav=1.0)
    best, ms = autotune.tune_batch(d, grids, candidates=(16, 64), attempts=3)
    print("batch seg_h", best, ms)
    assert best in (16, 64)
