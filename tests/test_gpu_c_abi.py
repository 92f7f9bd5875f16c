"""The C ABI driven from plain C (tests/c_abi/drive_advance.c, built by __graft_entry__.build()):
no Python and no PyTorch in the process that calls the library.  Its runs of config 0's radial
dam break (unsplit and split) and of the Euler quadrants match the oracle's evolve_to_time."""
import os
import struct
import subprocess

import numpy as np
import pytest

from clawpack_amd import problems
from oracle import corc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "c_abi", "drive_advance")


def run_driver(tmp_path, solver_id, q, dx, dy, tend, bc=1, split=0, grav=1.0, gamma=1.4, dt0=0.1):
    assert os.path.exists(DRIVER), "tests/c_abi/drive_advance is not built: run __graft_entry__.build()"
    neq, mx, my = q.shape
    planes = np.ascontiguousarray(q.transpose(0, 2, 1))  # (num_eqn, my, mx), x fastest
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(inp, "wb") as f:
        f.write(struct.pack("<8i", solver_id, mx, my, 2, 2, 4, bc, split))
        f.write(struct.pack("<6d", dx, dy, grav, gamma, dt0, tend))
        f.write(planes.tobytes())
    r = subprocess.run([DRIVER, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    status, done = struct.unpack_from("<2i", raw, 0)
    numsteps, rejected, attempts = struct.unpack_from("<3q", raw, 8)
    t, dt, cflmax = struct.unpack_from("<3d", raw, 32)
    qn = np.frombuffer(raw, dtype=np.float64, offset=56).reshape(neq, my, mx).transpose(0, 2, 1)
    return dict(status=status, done=done, numsteps=numsteps, rejected=rejected, t=t, dt=dt), qn


@pytest.mark.parametrize("split", [0, 1])
def test_c_driver_dam_break(tmp_path, split):
    print("case c driver dam break split", split, flush=True)
    q, dx, dy = problems.radial_dam_break(200, 200)
    st, qn = run_driver(tmp_path, 0, q, dx, dy, tend=0.5, split=split)
    d = corc.make_desc("swe", 200, 200, dx, dy, grav=1.0)
    qo, so = corc.evolve(d, corc.make_ctrl(dimensional_split=bool(split)), q, tend=0.5)
    assert st["status"] == 0 and st["done"] == 1
    assert st["numsteps"] == so["numsteps"] and st["rejected"] == so["rejected"]
    assert abs(st["t"] - 0.5) <= 1e-14
    assert np.abs(qn - qo).max() <= 1e-12 * np.abs(q).max()


def test_c_driver_euler(tmp_path):
    print("case c driver euler", flush=True)
    q, dx, dy = problems.euler_quadrants(96, 96, 1)
    st, qn = run_driver(tmp_path, 1, q, dx, dy, tend=0.05, dt0=1e-3)
    d = corc.make_desc("euler", 96, 96, dx, dy, gamma=1.4)
    qo, so = corc.evolve(d, corc.make_ctrl(dt_initial=1e-3), q, tend=0.05)
    assert st["status"] == 0 and st["numsteps"] == so["numsteps"]
    assert np.abs(qn - qo).max() <= 1e-12 * np.abs(q).max()
