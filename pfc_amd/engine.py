"""ctypes binding of libpfcgpu (include/pfc_relax.h).

This is the host-side mirror used by the tests and bench.py; the pfc-relax
binary binds the same C ABI directly.  Loading fails loudly when the built
library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import FP64, MIXED, PfcParams

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PFC_LIB") or os.path.join(PKG, "lib", "libpfcgpu.so")
BIN_PATH = os.path.join(PKG, "bin", "pfc-relax")
SMOOTH_PATH = os.path.join(PKG, "bin", "pfc-smooth")
FIELDS_PATH = os.path.join(PKG, "bin", "pfc-fields")

# exported symbols of include/pfc_relax.h
SYMBOLS = (
    "pfc_create", "pfc_set_fields", "pfc_set_box", "pfc_step", "pfc_evaluate", "pfc_evaluate_sums",
    "pfc_get_fields", "pfc_get_energy", "pfc_nan_step", "pfc_timing", "pfc_timing_read",
    "pfc_destroy", "pfc_strerror", "pfc_smooth", "pfc_fields",
    "pfc_batch_create", "pfc_batch_step", "pfc_batch_shape", "pfc_batch_destroy",
    "pfc_smooth_noalloc", "pfc_perturb", "pfc_step_evaluate",
)

BATCH_AUTOTUNE = 1
BATCH_NO_GRAPH = 2

_lib = None


class PfcError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what}: {strerror(code)} ({code})")


def load():
    """Loads libpfcgpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m pfc_amd.build` first")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    L.pfc_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(PfcParams), ctypes.c_int,
                             ctypes.c_int, ctypes.c_int]
    L.pfc_set_fields.argtypes = [vp, dp]
    L.pfc_set_box.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.pfc_step.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.pfc_evaluate.argtypes = [vp, dp, dp]
    L.pfc_evaluate_sums.argtypes = [vp, dp, dp]
    L.pfc_step_evaluate.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp]
    L.pfc_get_fields.argtypes = [vp, dp]
    L.pfc_get_energy.argtypes = [vp, dp]
    L.pfc_nan_step.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.pfc_timing.argtypes = [vp, ctypes.c_int]
    L.pfc_perturb.argtypes = [vp, ctypes.c_double, ctypes.c_ulonglong]
    L.pfc_timing_read.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int)]
    L.pfc_destroy.argtypes = [vp]
    L.pfc_smooth.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, ctypes.c_int]
    L.pfc_smooth_noalloc.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int, vp]
    L.pfc_fields.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                             ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp]
    L.pfc_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int]
    L.pfc_batch_step.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.pfc_batch_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int)]
    L.pfc_batch_destroy.argtypes = [vp]
    L.pfc_strerror.argtypes = [ctypes.c_int]
    L.pfc_strerror.restype = ctypes.c_char_p
    for s in SYMBOLS:
        if s != "pfc_strerror":
            getattr(L, s).restype = ctypes.c_int
    _lib = L
    return L


def strerror(code):
    return load().pfc_strerror(code).decode()


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Engine:
    """One device context: U fields on an Nx x Ny grid, resident on the GPU."""

    def __init__(self, params: PfcParams, Nx: int, Ny: int, dx: float, dy: float,
                 precision: int = FP64):
        L = load()
        self.U = params.U
        self.Nx, self.Ny = Nx, Ny
        self.params = params
        self.precision = precision
        self._h = ctypes.c_void_p()
        self._check(L.pfc_create(ctypes.byref(self._h), ctypes.byref(params), Nx, Ny, precision),
                    "pfc_create")
        self.set_box(dx, dy)

    def _check(self, rc, what):
        if rc != 0:
            raise PfcError(rc, what)

    def close(self):
        if self._h:
            load().pfc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_box(self, dx, dy):
        self.dx, self.dy = dx, dy
        self._check(load().pfc_set_box(self._h, dx, dy), "pfc_set_box")

    def set_fields(self, fields):
        f = np.ascontiguousarray(fields, dtype=np.float64).reshape(self.U, self.Ny, self.Nx)
        self._check(load().pfc_set_fields(self._h, _dp(f)), "pfc_set_fields")

    def get_fields(self):
        out = np.empty((self.U, self.Ny, self.Nx))
        self._check(load().pfc_get_fields(self._h, _dp(out)), "pfc_get_fields")
        return out

    def get_energy(self):
        out = np.empty((self.Ny, self.Nx))
        self._check(load().pfc_get_energy(self._h, _dp(out)), "pfc_get_energy")
        return out

    def step(self, t0, n=1):
        self._check(load().pfc_step(self._h, t0, n), "pfc_step")

    def evaluate(self, density=True):
        """(fave, nave).  density=False takes pfc_evaluate_sums: the same sums
        without the energy density field that get_energy downloads."""
        fave = ctypes.c_double()
        nave = np.zeros(self.U)
        fn = "pfc_evaluate" if density else "pfc_evaluate_sums"
        self._check(getattr(load(), fn)(self._h, ctypes.byref(fave), _dp(nave)), fn)
        return fave.value, nave

    def step_evaluate(self, t0, n):
        """(fave, nave) of the state at t0 as evaluate(density=False) gives
        them, then n >= 1 steps from t0 (pfc_step_evaluate: step t0's passes
        take the sums)."""
        fave = ctypes.c_double()
        nave = np.zeros(self.U)
        self._check(load().pfc_step_evaluate(self._h, t0, n, ctypes.byref(fave), _dp(nave)),
                    "pfc_step_evaluate")
        return fave.value, nave

    def perturb(self, rel, seed):
        """Multiplies the step state by 1 + rel * (unit-variance noise) (pfc_perturb)."""
        self._check(load().pfc_perturb(self._h, rel, seed), "pfc_perturb")

    def timing(self, enable=True):
        self._check(load().pfc_timing(self._h, 1 if enable else 0), "pfc_timing")

    def timing_read(self):
        """Summed (row, column) pass milliseconds and step count."""
        ms = np.zeros(2)
        n = ctypes.c_int()
        self._check(load().pfc_timing_read(self._h, _dp(ms), ctypes.byref(n)), "pfc_timing_read")
        return ms, n.value

    def nan_step(self):
        s = ctypes.c_int()
        self._check(load().pfc_nan_step(self._h, ctypes.byref(s)), "pfc_nan_step")
        return s.value


__all__ = ["Engine", "Batch", "PfcError", "load", "strerror", "smooth", "smooth_noalloc", "fields", "FP64", "MIXED",
           "LIB_PATH", "BIN_PATH", "SMOOTH_PATH", "FIELDS_PATH", "SYMBOLS", "BATCH_AUTOTUNE",
           "BATCH_NO_GRAPH"]


def smooth(field, dx, dy, sigma, precision=FP64):
    """Gaussian smoothing of one Ny x Nx field on the GPU (pfc_smooth)."""
    L = load()
    a = np.ascontiguousarray(field, dtype=np.float64)
    Ny, Nx = a.shape
    out = np.empty_like(a)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.pfc_smooth(a.ctypes.data_as(dp), out.ctypes.data_as(dp), Nx, Ny, dx, dy, sigma,
                      precision)
    if rc != 0:
        raise PfcError(rc, "pfc_smooth")
    return out


def smooth_noalloc(in_ptr, out_ptr, Nx, Ny, dx, dy, sigma, precision=FP64, stream=None):
    """pfc_smooth_noalloc on device buffers given by address (in_ptr, out_ptr:
    Ny x Nx doubles); queued on stream (a hipStream_t handle, None for the
    null stream).  Returns the status code."""
    return load().pfc_smooth_noalloc(ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr), Nx, Ny, dx, dy,
                                     sigma, precision, ctypes.c_void_p(stream or 0))


def fields(density, dx, dy, fold, sk, sb, power=None):
    """Orientation field phi (complex, max |phi| = 1) and, when power is
    given, deformation field chi of one H x W density on the GPU (pfc_fields,
    the computation of src/fields.c:104-359).  Returns (phi, chi or None)."""
    L = load()
    a = np.ascontiguousarray(density, dtype=np.float64)
    H, W = a.shape
    phi = np.empty((H, W), dtype=np.complex128)
    chi = np.empty((H, W)) if power is not None else None
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.pfc_fields(a.ctypes.data_as(dp), W, H, dx, dy, fold, sk, sb,
                      0.0 if power is None else power, phi.ctypes.data_as(dp),
                      chi.ctypes.data_as(dp) if chi is not None else None)
    if rc != 0:
        raise PfcError(rc, "pfc_fields")
    return phi, chi


class Batch:
    """Many Engines (same precision, U = 1 or 3, sides in {256, 384, 512, 768,
    1024, 1536, 2048}) stepped together by pfc_batch_step; the engines stay
    usable on their own (fields, evaluate) and must outlive the batch."""

    def __init__(self, engines, flags=0):
        L = load()
        self.engines = list(engines)
        arr = (ctypes.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        self._h = ctypes.c_void_p()
        rc = L.pfc_batch_create(ctypes.byref(self._h), arr, len(self.engines), flags)
        if rc != 0:
            raise PfcError(rc, "pfc_batch_create")

    def step(self, t0, nsteps):
        rc = load().pfc_batch_step(self._h, t0, nsteps)
        if rc != 0:
            raise PfcError(rc, "pfc_batch_step")

    def shape(self, pass_, n, U):
        """points per thread chosen for pass 0 (rows) or 1 (columns) of (n, U)"""
        p = ctypes.c_int()
        rc = load().pfc_batch_shape(self._h, pass_, n, U, ctypes.byref(p))
        if rc != 0:
            raise PfcError(rc, "pfc_batch_shape")
        return p.value

    def close(self):
        if self._h:
            load().pfc_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
