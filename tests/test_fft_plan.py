"""CPU emulation of the device FFT plan (BlockFFT in pfc_amd/csrc/fft_device.hpp):
the stage radices, the Stockham write and gather indices, the LDS index
swizzle and its composition from a per-lane base and per-access constants,
for LS transforms interleaved in one buffer.  Each emulated transform must
equal numpy's FFT, so a wrong stage index, a swizzle that is not a bijection
or a base/offset pair that is not bit-disjoint fails here.  The device turns
the swizzle on for FP64 with at most two interleaved transforms (E = 16,
LS <= 2) only; the
emulation also runs it on every other power-of-two shape, so the index
algebra is checked wherever it could be turned on."""
import math

import numpy as np
import pytest


def ilog(b, n):
    return 0 if n < b else 1 + ilog(b, n // b)


class Plan:
    """Index arithmetic of BlockFFT<T, L, TPB, DIR, TWS, LS> (E = sizeof(cpx<T>))."""

    def __init__(self, L, TPB, LS, E, swizzle=None, padc=0):
        self.L, self.TPB, self.LS, self.E = L, TPB, LS, E
        self.P = L // TPB
        self.pow2 = L & (L - 1) == 0
        if swizzle is None:   # the device's rule
            swizzle = E == 16 and LS <= 2 and padc == 0
        lsb = E * LS
        self.SWB = 0 if (not swizzle or not self.pow2 or lsb >= 256) else ilog(2, 256 // lsb)
        self.PC = padc if (self.pow2 and self.SWB == 0) else 0
        self.PSH = ilog(2, 256 // E)
        self.lds_len = L + self.PC * ((L - 1) >> self.PSH)

    def pd(self, i):
        return i + self.PC * (i >> self.PSH)

    def sw(self, i):
        if self.SWB == 0:
            return i
        c = i >> self.SWB
        return i ^ ((c ^ (c << 3)) & ((1 << self.SWB) - 1))

    def radices(self):
        L, P = self.L, self.P
        if self.pow2:
            S = ilog(P, L)
            R0 = L // P**S
            return ([R0] if R0 > 1 else []) + [P] * S
        out, rem = [], L
        while rem > 1:
            g = math.gcd(P, rem)
            out.append(g)
            rem //= g
        return out

    def run(self, x, DIR):
        """x: (LS, L) complex; returns the LS transforms, DIR = -1 forward."""
        L, TPB, P, LS = self.L, self.TPB, self.P, self.LS
        t = np.arange(TPB)
        # thread t holds elements t + q TPB (natural order)
        v = np.stack([x[:, t + q * TPB] for q in range(P)], axis=2)   # (LS, TPB, P)
        lds = np.full(self.lds_len * LS, np.nan, dtype=complex)
        rr = np.arange(LS)[:, None]
        rads = self.radices()
        Ns = 1
        for s, R in enumerate(rads):
            B = P // R
            last = s == len(rads) - 1
            nv = v.copy()
            for b in range(B):
                j = t + b * TPB
                k = j & (Ns - 1) if self.pow2 else j % Ns
                xs = [v[:, :, b + r * B] for r in range(R)]
                if Ns > 1:
                    step = L // (Ns * R)
                    xs = [xs[r] * np.exp(DIR * 2j * np.pi * (k * r * step) / L) for r in range(R)]
                X = [sum(xs[r] * np.exp(DIR * 2j * np.pi * r * kk / R) for r in range(R))
                     for kk in range(R)]
                if last:
                    for r in range(R):
                        nv[:, :, b + r * B] = X[r]
                else:
                    logical = (j - k) * R + k
                    if self.PC:
                        base = self.pd(logical)
                    else:
                        base = np.array([self.sw(int(a)) for a in logical])
                    for r in range(R):
                        if self.SWB:
                            assert not np.any(logical & (r * Ns))
                            idx = base ^ self.sw(r * Ns)
                        else:
                            idx = base + self.pd(r * Ns)
                            assert np.array_equal(idx, self.pd(logical + r * Ns))
                        lds[idx[None, :] * LS + rr] = X[r]
            if last:
                v = nv
                break
            Ns *= R
            Rn = rads[s + 1]
            Bn = P // Rn
            for b in range(Bn):
                j = t + b * TPB
                jb = self.pd(j) if self.PC else np.array([self.sw(int(a)) for a in j])
                for r in range(Rn):
                    if self.SWB:
                        assert not np.any(j & (r * (L // Rn)))
                        idx = jb ^ self.sw(r * (L // Rn))
                    else:
                        idx = jb + self.pd(r * (L // Rn))
                        assert np.array_equal(idx, self.pd(j + r * (L // Rn)))
                    v[:, :, b + r * Bn] = lds[idx[None, :] * LS + rr]
        out = np.empty_like(x)
        for q in range(P):
            out[:, t + q * TPB] = v[:, :, q]
        return out


def check(L, TPB, LS, E, rng, swizzle=None, padc=0):
    p = Plan(L, TPB, LS, E, swizzle, padc)
    x = rng.standard_normal((LS, L)) + 1j * rng.standard_normal((LS, L))
    ref = np.fft.fft(x, axis=1)
    scale = np.abs(ref).max()
    assert np.abs(p.run(x, -1) - ref).max() < 1e-10 * scale
    assert np.abs(p.run(x, +1) - np.fft.ifft(x, axis=1) * L).max() < 1e-10 * scale
    return p


# (L, TPB, LS, E): the row transforms (L = Nx/2, LS = rows per block), the
# column transforms (L = Ny, LS = columns per block) and the batch shapes
SHAPES = [
    (2048, 256, 4, 8), (2048, 256, 2, 16), (4096, 512, 1, 8), (4096, 512, 1, 16),
    (1024, 128, 8, 8), (1024, 128, 4, 16), (2048, 256, 1, 8), (2048, 256, 1, 16),
    (4096, 1024, 2, 8), (4096, 512, 1, 16), (8192, 1024, 1, 8), (8192, 1024, 1, 16),
    (512, 64, 4, 8), (512, 64, 2, 16), (1024, 128, 2, 8), (1024, 128, 2, 16),
    (256, 32, 8, 8), (256, 32, 8, 16), (128, 16, 16, 8), (128, 16, 16, 16),
    (64, 8, 32, 8), (256, 16, 16, 8), (512, 32, 8, 16), (1024, 64, 4, 8),
    (2048, 128, 8, 8), (2048, 128, 4, 16), (1024, 256, 4, 8), (2048, 512, 2, 8),
    (128, 64, 16, 8), (64, 16, 32, 16),
]


@pytest.mark.parametrize("swizzle", [None, True])
@pytest.mark.parametrize("L,TPB,LS,E", SHAPES)
def test_power_of_two_plans(L, TPB, LS, E, swizzle):
    p = check(L, TPB, LS, E, np.random.default_rng(L + TPB + LS + E), swizzle)
    on = swizzle or (E == 16 and LS <= 2)
    assert p.SWB == (0 if (not on or E * LS >= 256) else ilog(2, 256 // (E * LS)))


@pytest.mark.parametrize("L", [12, 24, 48, 96, 192, 384, 768, 1536])
@pytest.mark.parametrize("P", [6, 12])
def test_three_times_power_of_two_plans(L, P):
    if L % P:
        pytest.skip("P does not divide L")
    p = check(L, L // P, 1, 16, np.random.default_rng(L * P), swizzle=True)
    assert p.SWB == 0   # never for 3 * 2^k
    assert p.radices()[0] == P and all(r in (2, 4) for r in p.radices()[1:])


def test_swizzle_is_a_bijection_within_bank_rows():
    for E, LS in ((8, 1), (16, 1), (8, 4), (16, 2), (8, 2)):
        p = Plan(4096, 512, LS, E, swizzle=True)
        idx = np.arange(4096)
        s = np.array([p.sw(int(i)) for i in idx])
        assert sorted(s) == list(idx)
        assert np.all(s >> p.SWB == idx >> p.SWB)   # only the low bits move


def stage_table(L, P, NW):
    """The compact stage table as the device lays it out (BlockFFT::soff,
    kTab, tab_src): per stage with Ns > 1, Ns entries w1 and, with NW = 2,
    Ns entries wh; F is the first such Ns."""
    S = ilog(P, L)
    R0 = L // P**S
    F = R0 if R0 > 1 else P
    soff = lambda Ns: NW * (Ns - F) // (P - 1)
    kTab = NW * (L - F) // (P - 1)
    nstg = ilog(P, L // F)

    def tab_src(i):
        m, ns = 0, F
        for _ in range(nstg):
            step = L // (ns * P)
            m = i * step if 0 <= i < ns else m
            i -= ns
            if NW == 2:
                m = i * (P // 2) * step if 0 <= i < ns else m
                i -= ns
            ns *= P
        return m
    return soff, kTab, np.exp(-2j * np.pi * np.array([tab_src(i) for i in range(kTab)]) / L)


@pytest.mark.parametrize("NW", [1, 2])
@pytest.mark.parametrize("L,TPB", [(2048, 256), (2048, 128), (1024, 128), (1024, 64), (512, 64),
                                   (512, 32), (256, 32), (256, 16), (128, 16), (64, 8), (32, 8),
                                   (4096, 512), (8192, 1024), (4096, 256)])
def test_compact_stage_table(L, TPB, NW):
    """Every stage that multiplies twiddles finds w1[k] = e^{-2 pi i k step/L}
    (and wh[k] = w1[k]^(R/2)) at its block of the table, blocks do not
    overlap, and the table has exactly kTab entries."""
    p = Plan(L, TPB, 1, 8 * NW)
    soff, kTab, tab = stage_table(L, p.P, NW)
    Ns, used = 1, set()
    for R in p.radices():
        if Ns > 1:
            assert R == p.P
            step = L // (Ns * R)
            k = np.arange(Ns)
            o = soff(Ns)
            assert np.allclose(tab[o + k], np.exp(-2j * np.pi * k * step / L), atol=1e-15)
            idx = set(range(o, o + NW * Ns))
            if NW == 2:
                assert np.allclose(tab[o + Ns + k], np.exp(-2j * np.pi * k * (R // 2) * step / L),
                                   atol=1e-15)
            assert not (idx & used)
            used |= idx
        Ns *= R
    assert used == set(range(kTab))


@pytest.mark.parametrize("H,TPB", [(2048, 256), (2048, 128), (1024, 64), (192, 16), (384, 32),
                                   (768, 64), (96, 8), (32, 8), (4096, 512)])
def test_row_partner_index(H, TPB):
    """row_partner (pfc_kernels.hpp): the Hermitian partner H - k of
    k = t + q TPB, 0 for k = 0, for power-of-two and 3 * 2^k half lengths."""
    def row_partner(t, q):
        return (0 if t == 0 else H - t) if q == 0 else H - t - q * TPB
    for q in range(H // TPB):
        for t in range(TPB):
            k = t + q * TPB
            assert row_partner(t, q) == (H - k) % H


@pytest.mark.parametrize("L,TPB,LS", [(L, T, LS) for (L, T, LS, E) in SHAPES if E == 8] +
                         [(2048, 256, 2), (4096, 1024, 2), (1024, 64, 4), (256, 16, 16)])
def test_padded_fp32_plans(L, TPB, LS):
    """The FP32 step transforms' LDS padding (index i at i + 3 (i >> 5)): every
    access's base + immediate form equals the padded index (asserted inside),
    the transform is still numpy's, and the buffer holds lds_len * LS."""
    p = check(L, TPB, LS, 8, np.random.default_rng(L * LS + 1), padc=3)
    assert p.PC == 3 and p.lds_len == L + 3 * ((L - 1) >> 5)


@pytest.mark.parametrize("L,TPB,LS", [(L, T, LS) for (L, T, LS, E) in SHAPES if E == 16])
def test_padded_fp64_plans(L, TPB, LS):
    """The FP64 alternative to the swizzle (index i at i + 2 (i >> 4))."""
    p = check(L, TPB, LS, 16, np.random.default_rng(L * LS + 2), padc=2)
    assert p.SWB == 0 and p.PC == 2 and p.lds_len == L + 2 * ((L - 1) >> 4)
