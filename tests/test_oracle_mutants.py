"""The closed-form pins of tests/pins.py must reject an oracle with a
plausible mistake.  Each case builds oracle/pfc_oracle.c with one
ORACLE_MUTANT value (a planted mistake, see the top of that file) into a
temporary directory and runs the same check that the shipped oracle passes
in tests/test_oracle.py.  CPU only."""
import os
import subprocess

import pytest

import pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_mutant(tmp_path, k):
    so = str(tmp_path / f"liboracle_mutant{k}.so")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-std=c99", "-shared",
                    f"-DORACLE_MUTANT={k}", "-o", so, os.path.join(ROOT, "oracle", "pfc_oracle.c"),
                    "-lm"], check=True)
    return so


def test_pair_gamma_pin_rejects_swapped_mc_v(tmp_path):
    """m_u and m_v swapped in the gamma_uv term of mc[v]."""
    assert pins.pair_gamma_error(build_mutant(tmp_path, 1)) > 1e-4


def test_cadence_pin_rejects_step_before_print(tmp_path):
    r = pins.cadence_errors(build_mutant(tmp_path, 2))
    assert r["log"] > 1e-6


def test_cadence_pin_rejects_print_before_box_search(tmp_path):
    r = pins.cadence_errors(build_mutant(tmp_path, 3))
    assert r["box"] > 1e-4


def test_cadence_pin_rejects_snapshot_after_step(tmp_path):
    r = pins.cadence_errors(build_mutant(tmp_path, 4))
    assert r["snap"] > 1e-4
