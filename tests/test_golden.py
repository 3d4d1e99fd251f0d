"""Regression fixtures (tests/golden/*.npz, written by make_golden.py from the
oracle): the oracle must still reproduce them, and the GPU engine must match
them to the FP64 parity bar."""
import os

import numpy as np
import pytest

from golden.make_golden import CASES
from oracle_lib import Oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(name):
    d = load(name)
    pf = CASES[name][0]
    o = Oracle(pf(), int(d["Nx"]), int(d["Ny"]), float(d["dx"]), float(d["dy"]))
    o.set_fields(d["fields_in"])
    o.steps(0, int(d["steps"]))
    fave, nave = o.evaluate()
    out = o.get_fields()
    assert np.max(np.abs(out - d["fields_out"])) <= 1e-13 * np.max(np.abs(d["fields_out"]))
    assert abs(fave - float(d["fave"])) <= 1e-13 * abs(float(d["fave"]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_matches_fixture(gpu, name):
    from pfc_amd import FP64
    from pfc_amd.engine import Engine
    d = load(name)
    pf = CASES[name][0]
    g = Engine(pf(), int(d["Nx"]), int(d["Ny"]), float(d["dx"]), float(d["dy"]), FP64)
    g.set_fields(d["fields_in"])
    g.step(0, int(d["steps"]))
    fave, nave = g.evaluate()
    out = g.get_fields()
    ref = d["fields_out"]
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-12
    assert abs(fave - float(d["fave"])) <= 1e-12 * abs(float(d["fave"]))
    np.testing.assert_allclose(nave, d["nave"], rtol=1e-12, atol=1e-15)
