"""Builds libpfcgpu.so (HIP, gfx950) and the pfc-relax host binary in-tree.

    python -m pfc_amd.build            # everything
    python -m pfc_amd.build --lib      # only the HIP library

Outputs: pfc_amd/lib/libpfcgpu.so and pfc_amd/bin/pfc-relax (rpath to ../lib).
The HIP translation units (FP64 kernels, FP32 kernels, engine, fields, batch)
compile in parallel.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
OBJDIR = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PFC_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-pass-failed",
]
HIP_UNITS = ["pfc_launch_f64.hip", "pfc_launch_f32.hip", "pfc_engine.hip", "pfc_fields.hip",
             "pfc_batch_f32_row.hip", "pfc_batch_f32_col.hip", "pfc_batch_f64_row.hip",
             "pfc_batch_f64_col.hip", "pfc_batch.hip", "pfc_noalloc.hip", "pfc_generic.hip"]
HOST_SRCS = ["pfc_relax_main.c", "pfc_io.c"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _stale(obj, src):
    """True unless obj is newer than src, every header under csrc/ and the
    public header (the units include each other's headers freely)."""
    if not os.path.exists(obj):
        return True
    deps = [src, os.path.join(ROOT, "include", "pfc_relax.h")]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build_lib(verbose=False, libdir=None, objdir=None, defines=(), incremental=False):
    libdir = libdir or LIBDIR
    objdir = objdir or OBJDIR
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(libdir, exist_ok=True)
    objs = []
    cmds = []
    for u in HIP_UNITS:
        obj = os.path.join(objdir, u.replace(".hip", ".o"))
        objs.append(obj)
        if incremental and not defines and not _stale(obj, os.path.join(CSRC, u)):
            continue
        cmds.append([HIPCC, *HIP_FLAGS, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, u),
                     "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, len(cmds))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    lib = os.path.join(libdir, "libpfcgpu.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


def build_host():
    os.makedirs(BINDIR, exist_ok=True)
    exe = os.path.join(BINDIR, "pfc-relax")
    srcs = [os.path.join(HOST, s) for s in HOST_SRCS]
    _run(["gcc", "-O2", "-std=gnu99", "-Wall", "-Wextra", "-pthread", "-I", os.path.join(ROOT, "include"),
          *srcs, "-o", exe + ".tmp", "-L", LIBDIR, "-lpfcgpu", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    os.replace(exe + ".tmp", exe)
    sm = os.path.join(BINDIR, "pfc-smooth")
    _run(["gcc", "-O2", "-std=gnu99", "-Wall", "-Wextra", "-pthread", "-I", os.path.join(ROOT, "include"),
          os.path.join(HOST, "pfc_smooth_main.c"), os.path.join(HOST, "pfc_io.c"), "-o", sm + ".tmp",
          "-L", LIBDIR, "-lpfcgpu", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    os.replace(sm + ".tmp", sm)
    fl = os.path.join(BINDIR, "pfc-fields")
    _run(["gcc", "-O2", "-std=gnu99", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
          os.path.join(HOST, "pfc_fields_main.c"), "-o", fl + ".tmp",
          "-L", LIBDIR, "-lpfcgpu", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    os.replace(fl + ".tmp", fl)
    build_coordinator()
    build_init()
    return exe


def build_coordinator():
    """pfc-coordinator (coordinator.jar's command) and libpfccoord.so (the same
    conversion for in-memory fields); plain C, no GPU.  -ffp-contract=off keeps
    the arithmetic unfused, as Java evaluates it."""
    os.makedirs(BINDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    flags = ["-O2", "-std=gnu99", "-Wall", "-Wextra", "-ffp-contract=off"]
    core = os.path.join(HOST, "pfc_coord.c")
    exe = os.path.join(BINDIR, "pfc-coordinator")
    _run(["gcc", *flags, os.path.join(HOST, "pfc_coordinator_main.c"), core, "-o", exe + ".tmp", "-lm"])
    os.replace(exe + ".tmp", exe)
    lib = os.path.join(LIBDIR, "libpfccoord.so")
    _run(["gcc", *flags, "-shared", "-fPIC", core, "-o", lib + ".tmp", "-lm"])
    os.replace(lib + ".tmp", lib)
    return exe


def build_init():
    """pfc-init (pfc-init.jar's command): plain C, no GPU, unfused arithmetic
    so that grain values round as Java's do."""
    os.makedirs(BINDIR, exist_ok=True)
    exe = os.path.join(BINDIR, "pfc-init")
    _run(["gcc", "-O2", "-std=gnu99", "-Wall", "-Wextra", "-ffp-contract=off", "-pthread",
          "-I", os.path.join(ROOT, "include"), os.path.join(HOST, "pfc_init_main.c"), os.path.join(HOST, "pfc_jrand.c"),
          os.path.join(HOST, "pfc_io.c"), "-o", exe + ".tmp", "-lm"])
    os.replace(exe + ".tmp", exe)
    return exe


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="only the HIP library")
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only the HIP units older than their sources or headers")
    a = ap.parse_args(argv)
    lib = build_lib(a.verbose, incremental=a.incremental)
    print("built", lib)
    if not a.lib:
        print("built", build_host())


if __name__ == "__main__":
    sys.exit(main())
