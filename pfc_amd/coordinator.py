"""Density field to atom coordinates (the reference's coordinator,
src/Coordinator.java:29-318) through libpfccoord.so, and the comparison of two
atom sets used to state the mixed form's error in lattice constants.

The command-line tool with coordinator.jar's argv contract is
pfc_amd/bin/pfc-coordinator.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libpfccoord.so")
BIN_PATH = os.path.join(PKG, "bin", "pfc-coordinator")

# flags of pfc_coord_flags (include in pfc_amd/host/pfc_coord.h)
BAD_BONDS, BAD_RINGS_PER_ATOM, BAD_RING_SIZE, RING_MISMATCH, SMALL_SYSTEM = 1, 2, 4, 8, 16

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m pfc_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        L.pfc_coord_run.restype = vp
        L.pfc_coord_run.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.pfc_coord_atoms.argtypes = [vp]
        L.pfc_coord_rings.argtypes = [vp]
        L.pfc_coord_positions.argtypes = [vp, dp]
        L.pfc_coord_box.argtypes = [vp, dp, dp]
        L.pfc_coord_flags.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.pfc_coord_free.argtypes = [vp]
        L.pfc_java_double.argtypes = [ctypes.c_double, ctypes.c_char_p]
        _lib = L
    return _lib


def coordinate(psi, dx, dy, lpfc, la):
    """Atoms of the density psi (Ny x Nx): returns (positions (n, 2) in units
    of la, (width, height), number of rings, flags)."""
    L = _load()
    a = np.array(psi, dtype=np.float64, order="C", copy=True)   # modified in place by the C code
    Ny, Nx = a.shape
    h = L.pfc_coord_run(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), Nx, Ny, dx, dy,
                        lpfc, la)
    if not h:
        raise RuntimeError("coordinate conversion failed")
    try:
        n = L.pfc_coord_atoms(h)
        xy = np.zeros((n, 2))
        L.pfc_coord_positions(h, xy.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        w, hh = ctypes.c_double(), ctypes.c_double()
        L.pfc_coord_box(h, ctypes.byref(w), ctypes.byref(hh))
        bad = ctypes.c_int()
        flags = L.pfc_coord_flags(h, ctypes.byref(bad))
        return xy, (w.value, hh.value), L.pfc_coord_rings(h), flags
    finally:
        L.pfc_coord_free(h)


def java_double(d):
    """Java's Double.toString as the coordinate files print numbers."""
    buf = ctypes.create_string_buffer(32)
    _load().pfc_java_double(float(d), buf)
    return buf.value.decode()


def compare_atoms(xy_a, xy_b, box):
    """Nearest-neighbour displacement (periodic) from each atom of set b to
    set a: median, 99th percentile and maximum, in the units of the
    positions, and the atom counts."""
    from scipy.spatial import cKDTree
    w, h = box
    pa = np.mod(xy_a, [w, h])
    pb = np.mod(xy_b, [w, h])
    # the box is the same for both sets; guard against a coordinate equal to w
    pa = np.where(pa >= [w, h], 0.0, pa)
    pb = np.where(pb >= [w, h], 0.0, pb)
    d, _ = cKDTree(pa, boxsize=[w, h]).query(pb)
    return {"atoms_a": int(len(xy_a)), "atoms_b": int(len(xy_b)),
            "median": float(np.median(d)), "p99": float(np.percentile(d, 99)),
            "max": float(d.max())}
