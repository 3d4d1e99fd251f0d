"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: only tests/, the smoke check and bench.py's
CPU-baseline leg load it, always as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from pfc_amd.params import PfcParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_libs = {}


def lib(path=None):
    """The oracle library at path (default: oracle/liboracle.so, built on
    first use).  Other paths are builds of the same source, such as the
    seeded-mistake builds of tests/test_oracle_mutants.py."""
    path = path or ORACLE_SO
    if path not in _libs:
        if path == ORACLE_SO and not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.oracle_fft1d.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int]
        L.oracle_r2c2d.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_c2r2d.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_create.argtypes = [ctypes.POINTER(PfcParams), ctypes.c_int, ctypes.c_int]
        L.oracle_create.restype = vp
        L.oracle_destroy.argtypes = [vp]
        L.oracle_set_fields.argtypes = [vp, dp]
        L.oracle_get_fields.argtypes = [vp, dp]
        L.oracle_get_energy.argtypes = [vp, dp]
        L.oracle_get_spectral.argtypes = [vp, ctypes.c_int, dp]
        L.oracle_get_AB.argtypes = [vp, ctypes.c_int, dp, dp]
        L.oracle_set_box.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.oracle_step.argtypes = [vp, ctypes.c_int]
        L.oracle_evaluate.argtypes = [vp, dp, dp]
        L.oracle_optimize.argtypes = [vp, dp, ctypes.c_int, ctypes.c_int]
        L.oracle_get_box.argtypes = [vp, dp, dp]
        L.ENERGY_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_double, vp)
        L.oracle_optimize_with.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, ctypes.c_int,
                                           ctypes.c_int, L.ENERGY_FN, vp]
        L.oracle_relax.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                   ctypes.c_int, dp, dp]
        _libs[path] = L
    return _libs[path]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    assert lib().oracle_fft1d(x.size, _dp(x), _dp(out), sign) == 0
    return out


def r2c2d(a):
    """Real (Ny, Nx) -> transposed half spectrum (Nx//2+1, Ny)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    Ny, Nx = a.shape
    out = np.empty((Nx // 2 + 1, Ny), dtype=np.complex128)
    assert lib().oracle_r2c2d(Nx, Ny, _dp(a), _dp(out)) == 0
    return out


def c2r2d(X, Nx):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    Ny = X.shape[1]
    out = np.empty((Ny, Nx), dtype=np.float64)
    assert lib().oracle_c2r2d(Nx, Ny, _dp(X), _dp(out)) == 0
    return out


class Oracle:
    """One oracle solver instance (the reference's struct S on one rank)."""

    def __init__(self, params: PfcParams, Nx: int, Ny: int, dx: float, dy: float, so=None):
        self.L = lib(so)
        self.p = params
        self.U = params.U
        self.Nx, self.Ny = Nx, Ny
        self._h = self.L.oracle_create(ctypes.byref(params), Nx, Ny)
        if not self._h:
            raise ValueError("oracle_create failed")
        self.L.oracle_set_box(self._h, dx, dy)

    def close(self):
        if self._h:
            self.L.oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_fields(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.U, self.Ny, self.Nx)
        self.L.oracle_set_fields(self._h, _dp(f))

    def get_fields(self):
        out = np.empty((self.U, self.Ny, self.Nx))
        self.L.oracle_get_fields(self._h, _dp(out))
        return out

    def get_energy(self):
        out = np.empty((self.Ny, self.Nx))
        self.L.oracle_get_energy(self._h, _dp(out))
        return out

    def spectral(self, u):
        out = np.empty((self.Nx // 2 + 1, self.Ny), dtype=np.complex128)
        self.L.oracle_get_spectral(self._h, u, _dp(out))
        return out

    def AB(self, u):
        nk = (self.Nx // 2 + 1) * self.Ny
        A = np.empty(nk)
        B = np.empty(nk)
        self.L.oracle_get_AB(self._h, u, _dp(A), _dp(B))
        return A.reshape(self.Nx // 2 + 1, self.Ny), B.reshape(self.Nx // 2 + 1, self.Ny)

    def set_box(self, dx, dy):
        self.L.oracle_set_box(self._h, dx, dy)

    def box(self):
        dx = ctypes.c_double()
        dy = ctypes.c_double()
        self.L.oracle_get_box(self._h, ctypes.byref(dx), ctypes.byref(dy))
        return dx.value, dy.value

    def step(self, t):
        return self.L.oracle_step(self._h, t)

    def steps(self, t0, n):
        for t in range(t0, t0 + n):
            rc = self.step(t)
            if rc:
                return rc
        return 0

    def evaluate(self):
        fave = ctypes.c_double()
        nave = np.zeros(self.U)
        self.L.oracle_evaluate(self._h, ctypes.byref(fave), _dp(nave))
        return fave.value, nave

    def optimize(self, d, xopt=1, yopt=1):
        dd = ctypes.c_double(d)
        self.L.oracle_optimize(self._h, ctypes.byref(dd), xopt, yopt)
        return dd.value

    def relax(self, tend, Tprint, Topt=0, xopt=1, yopt=1, snap_steps=(), energy=False):
        """(log, field snapshots), and with energy=True also the energy
        densities that write_energy would write at the snapshot steps."""
        rows = tend // max(Tprint, 1) + 2
        log = np.zeros((rows, 4 + self.U))
        snaps = np.zeros((max(len(snap_steps), 1), self.U, self.Ny, self.Nx))
        fsnaps = np.zeros((max(len(snap_steps), 1), self.Ny, self.Nx)) if energy else None
        ss = (ctypes.c_int * max(len(snap_steps), 1))(*snap_steps)
        n = self.L.oracle_relax(self._h, tend, Tprint, Topt, xopt, yopt, _dp(log), rows, ss,
                               len(snap_steps), _dp(snaps),
                               _dp(fsnaps) if energy else None)
        if n < 0:
            raise FloatingPointError(f"oracle relax failed with code {-n}")
        if energy:
            return log[:n], snaps[:len(snap_steps)], fsnaps[:len(snap_steps)]
        return log[:n], snaps[:len(snap_steps)]


def optimize_with(Nx, Ny, dx, dy, d, xopt, yopt, fn):
    """Runs the oracle's optimisation arithmetic against a Python energy."""
    L = lib()
    cb = L.ENERGY_FN(lambda a, b, _u: float(fn(a, b)))
    cdx, cdy, cd = ctypes.c_double(dx), ctypes.c_double(dy), ctypes.c_double(d)
    L.oracle_optimize_with(Nx, Ny, ctypes.byref(cdx), ctypes.byref(cdy), ctypes.byref(cd),
                           xopt, yopt, cb, None)
    return cdx.value, cdy.value, cd.value
