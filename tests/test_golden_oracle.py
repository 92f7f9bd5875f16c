"""The oracle reproduces its committed golden fixtures bit for bit (CPU regression pin)."""
import glob
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_reproduces_fixture(path):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixtures", os.path.join(HERE, "golden", "make_fixtures.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    name = os.path.basename(path)[:-4]
    ref = np.load(path)
    q0, qn, st, dx, dy, aux = mk.run_case(name)
    assert np.array_equal(q0, ref["q0"])
    if aux is not None:
        assert np.array_equal(aux, ref["aux"])
    assert np.array_equal(qn, ref["q"])
    assert st["t"] == float(ref["t"]) and st["numsteps"] == int(ref["numsteps"])
    assert st["rejected"] == int(ref["rejected"]) and st["dt"] == float(ref["dt"])


def test_fixture_set_present():
    assert len(FIXTURES) >= 6
    assert any("bathy" in p for p in FIXTURES)
