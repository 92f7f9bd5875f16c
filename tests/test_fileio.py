"""Frame files (pyclaw fileio.ascii and fileio.binary): layout, round trip, and the Controller's output loop.
The format follows the published Clawpack ascii layout; no reference frame files exist in the
reference tree, so byte-level parity with upstream is unpinned (the checks are structural)."""
import os

import numpy as np
import pytest

from clawpack_amd import pyclaw
from clawpack_amd.pyclaw.fileio import ascii


def make_solution(mx=7, my=5, num_eqn=3, num_aux=0, t=0.25):
    domain = pyclaw.Domain([pyclaw.Dimension(-1.0, 2.0, mx, name="x"),
                            pyclaw.Dimension(0.5, 1.5, my, name="y")])
    state = pyclaw.State(domain, num_eqn, num_aux)
    rng = np.random.default_rng(3)
    state.q[...] = rng.standard_normal(state.q.shape) * 10.0 ** rng.integers(-3, 3, size=state.q.shape)
    if num_aux:
        state.aux[...] = rng.standard_normal(state.aux.shape)
    state.t = t
    return pyclaw.Solution(state, domain)


def test_frame_layout(tmp_path):
    sol = make_solution()
    sol.write(3, str(tmp_path))
    assert sorted(os.listdir(tmp_path)) == ["fort.q0003", "fort.t0003"]
    t_lines = open(tmp_path / "fort.t0003").read().splitlines()
    assert [l.split()[1] for l in t_lines] == ["time", "num_eqn", "nstates", "num_aux", "num_dim", "num_ghost"]
    assert float(t_lines[0].split()[0]) == 0.25 and int(t_lines[1].split()[0]) == 3
    q_lines = open(tmp_path / "fort.q0003").read().split("\n")
    names = [l.split()[1] for l in q_lines[:8]]
    assert names == ["patch_number", "AMR_level", "mx", "my", "xlow", "ylow", "dx", "dy"]
    assert float(q_lines[4].split()[0]) == -1.0 and abs(float(q_lines[6].split()[0]) - 3.0 / 7) < 1e-8
    assert q_lines[8] == ""
    data = q_lines[9:]
    # 5 rows of 7 cells, each row closed by a blank line; x fastest
    assert data[7] == "" and len(data[0].split()) == 3 and len(data[0]) == 3 * 18
    first = [float(v) for v in data[0].split()]
    second = [float(v) for v in data[1].split()]
    np.testing.assert_allclose(first, sol.state.q[:, 0, 0], rtol=1e-8)
    np.testing.assert_allclose(second, sol.state.q[:, 1, 0], rtol=1e-8)


@pytest.mark.parametrize("num_eqn,num_aux", [(3, 0), (4, 2)])
def test_round_trip(tmp_path, num_eqn, num_aux):
    sol = make_solution(num_eqn=num_eqn, num_aux=num_aux)
    sol.write(12, str(tmp_path), write_aux=bool(num_aux), file_prefix="claw")
    back = pyclaw.Solution(12, path=str(tmp_path), file_prefix="claw")
    assert back.t == pytest.approx(0.25, rel=1e-8)
    assert back.state.q.shape == sol.state.q.shape
    np.testing.assert_allclose(back.state.q, sol.state.q, rtol=1e-8, atol=0)
    d0, d1 = back.domain.patch.dimensions
    assert (d0.num_cells, d1.num_cells) == (7, 5)
    assert d0.lower == pytest.approx(-1.0) and d0.upper == pytest.approx(2.0, rel=1e-8)
    assert d1.delta == pytest.approx(0.2, rel=1e-8)
    if num_aux:
        np.testing.assert_allclose(back.state.aux, sol.state.aux, rtol=1e-8, atol=0)


def test_unsupported_formats(tmp_path):
    sol = make_solution()
    with pytest.raises(NotImplementedError):
        sol.write(0, str(tmp_path), file_format="hdf5")
    with pytest.raises(NotImplementedError):
        pyclaw.Solution(0, path=str(tmp_path), file_format="netcdf")


class _HostSolver:
    """Stand-in solver for the Controller loop (advances t only; no compute)."""

    def __init__(self):
        self.calls = []

    def setup(self, solution):
        pass

    def evolve_to_time(self, solution, tend):
        self.calls.append(tend)
        solution.state.q[0] += 1.0
        solution.t = tend
        return {"cflmax": 0.5, "dtmin": 0.01, "dtmax": 0.02, "numsteps": 3}


def test_controller_writes_frames(tmp_path):
    sol = make_solution(t=0.0)
    claw = pyclaw.Controller()
    assert claw.output_format == "ascii" and claw.outdir == "./_output"
    claw.solution, claw.solver = sol, _HostSolver()
    claw.tfinal, claw.num_output_times = 1.0, 4
    claw.outdir = str(tmp_path)
    claw.keep_copy = True
    q0 = sol.state.q.copy()
    status = claw.run()
    files = sorted(os.listdir(tmp_path))
    assert files == ["fort.q%04d" % k for k in range(5)] + ["fort.t%04d" % k for k in range(5)]
    assert status["numsteps"] == 12
    for k in range(5):
        back = pyclaw.Solution(k, path=str(tmp_path))
        assert back.t == pytest.approx(0.25 * k, abs=1e-12)
        np.testing.assert_allclose(back.state.q[0], q0[0] + k, rtol=1e-8)
        np.testing.assert_allclose(back.state.q, claw.frames[k].state.q, rtol=1e-8)


# ---- binary frames (pyclaw fileio.binary; Clawpack's binary64 layout) ----
from clawpack_amd.pyclaw.fileio import binary  # noqa: E402


@pytest.mark.parametrize("num_eqn,num_aux", [(3, 0), (4, 2)])
def test_binary_round_trip_bitwise(tmp_path, num_eqn, num_aux):
    sol = make_solution(num_eqn=num_eqn, num_aux=num_aux, t=0.1 + 1e-12)
    sol.write(7, str(tmp_path), file_format="binary", write_aux=bool(num_aux))
    back = pyclaw.Solution(7, path=str(tmp_path), file_format="binary")
    assert back.state.q.tobytes() == np.asfortranarray(sol.state.q).tobytes()
    assert back.t == pytest.approx(sol.t, rel=1e-8)
    if num_aux:
        assert np.array_equal(back.state.aux, sol.state.aux)


def test_binary_layout(tmp_path):
    """The data file is float64 little-endian, component fastest, then x, then y; the q file holds
    only the ascii patch header; the t file names the format on its seventh line."""
    sol = make_solution(mx=7, my=5, num_eqn=3)
    sol.write(2, str(tmp_path), file_format="binary")
    assert sorted(os.listdir(tmp_path)) == ["fort.b0002", "fort.q0002", "fort.t0002"]
    t_lines = open(tmp_path / "fort.t0002").read().splitlines()
    assert [l.split()[1] for l in t_lines] == ["time", "num_eqn", "nstates", "num_aux", "num_dim",
                                               "num_ghost", "file_format"]
    assert t_lines[6].split()[0] == "binary64" and int(t_lines[5].split()[0]) == 0
    q_lines = [l for l in open(tmp_path / "fort.q0002").read().split("\n") if l.strip()]
    assert [l.split()[1] for l in q_lines] == ["patch_number", "AMR_level", "mx", "my", "xlow", "ylow",
                                               "dx", "dy"]
    raw = np.fromfile(tmp_path / "fort.b0002", dtype="<f8")
    assert raw.size == 3 * 7 * 5
    q = sol.state.q
    # element (m, i, j) sits at m + 3*(i + 7*j)
    for (m, i, j) in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (2, 6, 0), (0, 0, 1), (2, 6, 4)]:
        assert raw[m + 3 * (i + 7 * j)] == q[m, i, j]


def test_binary_reads_ghost_cells(tmp_path):
    """A frame whose data carries num_ghost ghost layers (as Fortran codes write it) is read back
    without them."""
    sol = make_solution(mx=6, my=4, num_eqn=3)
    sol.write(0, str(tmp_path), file_format="binary")
    g = 2
    q = sol.state.q
    padded = np.full((3, 6 + 2 * g, 4 + 2 * g), -7.0)
    padded[:, g:-g, g:-g] = q
    with open(tmp_path / "fort.b0000", "wb") as f:
        f.write(np.asfortranarray(padded).tobytes(order="F"))
    t_path = tmp_path / "fort.t0000"
    lines = open(t_path).read().splitlines()
    lines[5] = "%5i                  num_ghost" % g
    open(t_path, "w").write("\n".join(lines) + "\n")
    back = pyclaw.Solution(0, path=str(tmp_path), file_format="binary")
    assert np.array_equal(back.state.q, q)


def test_binary_errors(tmp_path):
    sol = make_solution()
    sol.write(1, str(tmp_path))  # ascii
    with pytest.raises(IOError):
        binary.read(None, 1, str(tmp_path))
    sol.write(2, str(tmp_path), file_format="binary")
    with open(tmp_path / "fort.b0002", "ab") as f:
        f.write(b"\0" * 8)
    with pytest.raises(IOError):
        pyclaw.Solution(2, path=str(tmp_path), file_format="binary")


def test_unknown_format_writes_nothing(tmp_path):
    sol = make_solution()
    with pytest.raises(NotImplementedError):
        sol.write(6, str(tmp_path), file_format=["binary", "hdf5"])
    assert not os.listdir(tmp_path)
    sol.write(4, str(tmp_path / "a"), file_format="ascii")
    sol.write(4, str(tmp_path / "b"), file_format="binary")
    a = pyclaw.Solution(4, path=str(tmp_path / "a"), file_format="ascii")
    b = pyclaw.Solution(4, path=str(tmp_path / "b"), file_format="binary")
    np.testing.assert_allclose(a.state.q, b.state.q, rtol=1e-8)


def test_controller_binary_frames(tmp_path):
    sol = make_solution(t=0.0)
    claw = pyclaw.Controller()
    claw.solution, claw.solver = sol, _HostSolver()
    claw.tfinal, claw.num_output_times = 1.0, 2
    claw.outdir = str(tmp_path)
    claw.output_format = "binary"
    q0 = sol.state.q.copy()
    claw.run()
    for k in range(3):
        back = pyclaw.Solution(k, path=str(tmp_path), file_format="binary")
        assert np.array_equal(back.state.q[0], q0[0] + k)
        assert back.t == pytest.approx(k * 0.5)
