"""The drop-in boundary on the CPU: the C ABI library loads and exports every
entry point of include/pfc_relax.h, and the pfc-relax binary keeps the
reference's argv contract, input grammar errors and exit codes
(src/pfc-relax.c:242-411 fail(), :414-725 read_input, :1402-1412 main).

No test here launches GPU work: every case fails in the parser or in the
data-file reader, which run before the device is touched.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from pfc_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfc_relax.h")
BIN = engine.BIN_PATH


def header_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pfc_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_header_symbol():
    syms = header_symbols()
    assert len(syms) >= 10
    lib = ctypes.CDLL(engine.LIB_PATH)
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(engine.SYMBOLS)


def test_strerror_and_create_without_device():
    lib = engine.load()
    assert "grid size not supported by this build" in engine.strerror(-4)
    from pfc_amd import graphene_params
    p = graphene_params()
    h = ctypes.c_void_p()
    # invalid arguments are rejected before any device call: a prime length
    # whose Bluestein transform (M = 32768) exceeds one block's LDS, and a
    # one-point axis
    assert lib.pfc_create(ctypes.byref(h), ctypes.byref(p), 9973, 64, 0) == -4
    assert lib.pfc_create(ctypes.byref(h), ctypes.byref(p), 64, 1, 0) == -4
    assert lib.pfc_create(ctypes.byref(h), ctypes.byref(p), 64, 64, 7) == -1
    bad = graphene_params()
    bad.U = 9
    assert lib.pfc_create(ctypes.byref(h), ctypes.byref(bad), 64, 64, 0) == -1


# ------------------------------------------------------------------ CLI

def run(args, cwd, env=None):
    e = dict(os.environ)
    e.pop("PMI_RANK", None)
    e.pop("OMPI_COMM_WORLD_RANK", None)
    e.pop("PMIX_RANK", None)
    e.update(env or {})
    r = subprocess.run([BIN, *args], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, env=e, timeout=60)
    return r.returncode, r.stdout, r.stderr


def write_density(path, a, dx=0.7, dy=0.7):
    Ny, Nx = a.shape
    with open(path, "w") as f:
        f.write(f"{Nx} {Ny} {dx:.6e} {dy:.6e}\n")
        for v in a.ravel():
            f.write(f"{v:.6e}\n")


HONEY = """# an input file in the shape of demos/honey/honey.in
I\t{inp}
\t{out}\t{fout}
\t{wline}

M\t1
l\t0.0
a\t-0.15
b\t1.0
g\t0.8
d\t1.0
e\t
s\t0.2
n\t1.0
o\t1.0
p\t0.0
m\t
R\t0\t{tend}\t0.7\t0.7\t0.5\t0\t1\t1
"""


def honey_input(tmp_path, wline="1000\t[\t0\t100\t1000\t10000\t]\t12\t6", tend=10, **over):
    a = 0.06 + 0.01 * np.random.default_rng(0).standard_normal((64, 64))
    write_density(tmp_path / "honey.n", a)
    txt = HONEY.format(inp="honey.n", out="honey-#.n", fout="honey-#.f", wline=wline, tend=tend)
    for k, v in over.items():
        txt = txt.replace(k, v)
    (tmp_path / "honey.in").write_text(txt)
    return "honey.in"


def test_usage_and_syntax(tmp_path):
    rc, out, _ = run([], tmp_path)
    assert rc == 0 and out.startswith("This is a program for relaxing PFC systems.")
    rc, out, _ = run(["a", "b"], tmp_path)
    assert rc == 41 and out == "Error: Invalid syntax.\n"


def test_missing_input_file(tmp_path):
    rc, out, _ = run(["nope.in"], tmp_path)
    assert rc == 0 and out == "Error: Input file not found.\n"


def test_nonzero_rank_exits_quietly(tmp_path):
    rc, out, _ = run(["nope.in"], tmp_path, env={"PMI_RANK": "2"})
    assert rc == 0 and out == ""
    rc, out, _ = run([], tmp_path, env={"OMPI_COMM_WORLD_RANK": "1"})
    assert rc == 0 and out == ""


@pytest.mark.parametrize("change,code,msg", [
    ({"# an input": "XY an input"}, 1, "Error: Invalid multicharacter label found."),
    ({"I\thoney.n": "I\tmissing.n"}, 3, "Error: Input data file not found."),
    ({"\thoney-#.n\thoney-#.f": "\t"}, 6, "Error: Invalid number of output data files specified."),
    ({"M\t1": "M\tx"}, 11, "Error: Failed to read the number of density fields."),
    ({"M\t1": "M\t2"}, 12, "Error: Mismatching numbers of input data files and density fields specified."),
    ({"l\t0.0": "l\tzero"}, 14, "Error: Invalid syntax for parameters lambda."),
    ({"a\t-0.15": "a\tminus"}, 16, "Error: Invalid syntax for parameters alpha."),
    ({"b\t1.0": "b\tone"}, 18, "Error: Invalid syntax for parameters beta."),
    ({"g\t0.8": "g\tx"}, 20, "Error: Invalid syntax for parameters gamma."),
    ({"d\t1.0": "d\tx"}, 22, "Error: Invalid syntax for parameters delta."),
    ({"s\t0.2": "s\tx"}, 26, "Error: Invalid syntax for parameters sigma."),
    ({"n\t1.0": "n\tx"}, 28, "Error: Invalid syntax for parameters 1/nu."),
    ({"o\t1.0": "o\tx"}, 30, "Error: Invalid syntax for parameters omega."),
    ({"p\t0.0": "p\tx"}, 32, "Error: Invalid syntax for parameters phi."),
    ({"R\t0\t10\t0.7\t0.7\t0.5\t0\t1\t1": "R\t0\t10\t0.7"}, 35, "Error: Invalid parameters for relaxation."),
    ({"R\t0": "Q\t0"}, 36, "Error: Invalid label found."),
])
def test_grammar_errors(tmp_path, change, code, msg):
    inp = honey_input(tmp_path, **change)
    rc, out, _ = run([inp], tmp_path)
    assert (rc, out) == (code, msg + "\n")


@pytest.mark.parametrize("label,code", [("l", 13), ("a", 15), ("b", 17), ("g", 19), ("d", 21),
                                        ("e", 23), ("s", 25), ("n", 27), ("o", 29), ("p", 31),
                                        ("m", 33)])
def test_missing_model_labels(tmp_path, label, code):
    """A label that never appears: the search for it runs into the end of
    the file and fails with the label's code (src/pfc-relax.c:549-698)."""
    inp = honey_input(tmp_path)
    txt = (tmp_path / inp).read_text()
    cut = txt.index("\n" + label + "\t") + 1
    # drop that label and everything after it
    (tmp_path / inp).write_text(txt[:cut])
    rc, out, _ = run([inp], tmp_path)
    assert rc == code, (rc, out)


@pytest.mark.parametrize("wline,code", [
    ("1000\t1000\t12", 9),
    ("1000\t(\t0\t100\t)\t12\t6", 9),
    ("1000\t[\t0\t100\t50\t]\t12\t6", 10),
])
def test_write_parameter_errors(tmp_path, wline, code):
    inp = honey_input(tmp_path, wline=wline)
    rc, out, _ = run([inp], tmp_path)
    assert rc == code, out


def test_data_file_errors(tmp_path):
    inp = honey_input(tmp_path)
    # bad header
    (tmp_path / "honey.n").write_text("x y\n")
    rc, out, _ = run([inp], tmp_path)
    assert (rc, out) == (4, "Error: Invalid input data file header.\n")
    # too few values
    write_density(tmp_path / "honey.n", np.zeros((64, 64)))
    lines = (tmp_path / "honey.n").read_text().splitlines()
    (tmp_path / "honey.n").write_text("\n".join(lines[:100]) + "\n")
    rc, out, _ = run([inp], tmp_path)
    assert (rc, out) == (38, "Error: Invalid data in data file.\n")
    # mismatching dimensions of two input files
    hetero = HONEY.replace("I\t{inp}", "I\t{inp}\tb.n").replace("\t{out}\t{fout}", "\t{out}\tb-#.n")
    write_density(tmp_path / "honey.n", np.zeros((64, 64)))
    write_density(tmp_path / "b.n", np.zeros((64, 128)))
    (tmp_path / "two.in").write_text(hetero.format(inp="honey.n", out="honey-#.n", fout="",
                                                   wline="10 10 12 6", tend=10))
    rc, out, _ = run(["two.in"], tmp_path)
    assert (rc, out) == (5, "Error: Invalid mismatching dimensions of input data files.\n")


def test_comment_at_end_of_file(tmp_path):
    inp = honey_input(tmp_path)
    (tmp_path / inp).write_text((tmp_path / inp).read_text() + "#")
    rc, out, _ = run([inp], tmp_path)
    assert (rc, out) == (40, "Error: An error occurred while reading the input file.\n")


def test_valid_input_reaches_the_device(tmp_path):
    """The demo-shaped input parses completely: without a GPU the binary
    stops at device creation (exit 42) after reading the data file."""
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("GPU present: the end-to-end runs are in test_gpu_cli.py")
    inp = honey_input(tmp_path)
    rc, out, err = run([inp], tmp_path)
    assert rc == 42 and "pfc_create failed" in err and out == ""
