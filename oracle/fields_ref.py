"""fields_ref.py -- numpy FP64 restatement of the reference's fields tool.

TEST INFRASTRUCTURE ONLY: the checker for pfc_fields (include/pfc_relax.h);
nothing in the product imports it.

Correspondence with src/fields.c:
  orientation()  <- write_phi  :104-226  (min shift, complex kernel
                    exp(a (|k|-1)^2) e^{i fold atan2(ky, kx)}, product with the
                    shifted density, Gaussian smoothing, division by max |z|)
  deformation()  <- write_chi  :229-359  (gradient of the smoothed field from
                    the kept spectrum, |grad|^power, multiscale Gaussian sum)
  waves()        <- the kx, ky of :136-147 (index i < n/2 -> i dk, else (i-n) dk)
numpy's FFTs are unnormalised forward and 1/n backward; FFTW's backward
transform (the reference's) is unnormalised, so every inverse here is scaled
back by W*H.  Pinned by closed forms in tests/test_fields.py (a plane wave of
wave number 1 at two angles, and power 2, whose deformation field is a finite
cosine series).
"""
import numpy as np

PI = 3.14159265358979323846264338327   # fields.c:15


def waves(n, d):
    i = np.arange(n)
    return np.where(i < n // 2, i, i - n) * (2.0 * PI / (d * n))


def _ifft2(Z):
    H, W = Z.shape
    return np.fft.ifft2(Z) * (W * H)


def orientation(p, dx, dy, fold, sk, sb):
    """phi normalised to max |phi| = 1, and the kept smoothed spectrum
    (fields.c:168-180 stores it in zx for write_chi)."""
    H, W = p.shape
    kx = waves(W, dx)[None, :]
    ky = waves(H, dy)[:, None]
    divA = 1.0 / W / H
    shifted = p - p.min()
    Z = np.fft.fft2(shifted.astype(np.complex128))
    q = np.sqrt(kx * kx + ky * ky) - 1.0
    Z = Z * (np.exp(-0.5 / (sk * sk) * q * q) * np.exp(1j * fold * np.arctan2(ky, kx)) * divA)
    z = _ifft2(Z) * shifted
    Z = np.fft.fft2(z) * (np.exp(-0.5 / (sb * sb) * (kx * kx + ky * ky)) * divA)
    kept = Z.copy()
    z = _ifft2(Z)
    return z * (1.0 / np.abs(z).max()), kept


def scales(W, H, dx, dy):
    """The scales n = 2, 4, ... of fields.c:307-309."""
    l0 = 4.0 * PI / np.sqrt(3.0)
    L = min(W * dx, H * dy)
    out, n = [], 2
    while l0 * n < 0.5 * L:
        out.append(n)
        n *= 2
    return out


def deformation(kept, dx, dy, power):
    H, W = kept.shape
    kx = waves(W, dx)[None, :]
    ky = waves(H, dy)[:, None]
    divA = 1.0 / W / H
    zx = _ifft2(1j * kx * kept)
    zy = _ifft2(1j * ky * kept)
    G = np.fft.fft2((np.abs(zx) ** 2 + np.abs(zy) ** 2) ** (0.5 * power))
    chi = np.zeros((H, W))
    for n in scales(W, H, dx, dy):
        s = 1.0 / n
        lev = _ifft2(np.exp(-0.5 / (s * s) * (kx * kx + ky * ky)) * divA * G).real
        chi += lev * (1.0 / lev.max())
    return chi
