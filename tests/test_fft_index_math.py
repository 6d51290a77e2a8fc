"""CPU checks of the index arithmetic behind the runtime-length in-LDS FFTs
(wave-vortex-model_amd/csrc/wv_device.hpp), which the GPU parity tests
exercise only at the lengths they run:

- FastDiv: q = trunc(float32(a) * float32(rcp(d) * (1 + 2^-16))) equals
  a // d for every a + d < 2^15, 1 <= d <= 4096, with the FP32 reciprocal
  rounded down, to nearest or up (the hardware reciprocal's 1-ulp error);
- the prime-factor maps of the composite radices 6, 10, 12, 14 and 15: input
  n = (P2 n1 + P1 n2) mod R, output k = (e1 k1 + e2 k2) mod R, a DFT of P1
  points along n1 then of P2 along n2, equal the R-point DFT in both
  directions;
- the Stockham schedule of the mixed-radix passes (odd factors first, the
  composite passes, radix 16 and the 8/4/2 remainder; twiddle
  exp(-2 pi i r k / (Ns R)) = tw[r k N / (Ns R)]) equals numpy's FFT.
"""
import numpy as np
import pytest


def test_fast_division_is_exact_over_the_index_range():
    a = np.arange(1 << 15, dtype=np.int64)
    af = a.astype(np.float32)
    bump = np.float32(1.0 + 2.0 ** -16)
    for d in range(1, 4097):
        r = np.float32(1.0) / np.float32(d)
        for rcp in (np.nextafter(r, np.float32(0)), r, np.nextafter(r, np.float32(1))):
            inv = np.float32(rcp * bump)
            q = (af * inv).astype(np.int64)  # float32 product, truncation
            lim = a + d < (1 << 15)
            assert np.array_equal(q[lim], a[lim] // d), d


def _dft(x, sign):
    n = np.arange(len(x))
    return np.exp(sign * 2j * np.pi * np.outer(n, n) / len(x)) @ x


def _pfa(v, p1, p2, sign):
    r = p1 * p2
    e1 = p2 * pow(p2, -1, p1) if p1 > 1 else 0
    e2 = p1 * pow(p1, -1, p2)
    assert e1 % p1 == 1 and e1 % p2 == 0 and e2 % p1 == 0 and e2 % p2 == 1
    a = np.array([[v[(p2 * n1 + p1 * n2) % r] for n2 in range(p2)] for n1 in range(p1)])
    a = np.stack([_dft(a[:, n2], sign) for n2 in range(p2)], axis=1)
    a = np.stack([_dft(a[k1, :], sign) for k1 in range(p1)], axis=0)
    out = np.zeros(r, complex)
    for k1 in range(p1):
        for k2 in range(p2):
            out[(e1 * k1 + e2 * k2) % r] = a[k1, k2]
    return out


@pytest.mark.parametrize("p1,p2", [(3, 2), (5, 2), (3, 4), (7, 2), (3, 5), (9, 2), (5, 4), (3, 8)])
def test_prime_factor_composite_radices(p1, p2):
    rng = np.random.default_rng(p1 * p2)
    x = rng.standard_normal(p1 * p2) + 1j * rng.standard_normal(p1 * p2)
    for sign in (-1, 1):
        assert np.abs(_pfa(x, p1, p2, sign) - _dft(x, sign)).max() < 1e-12


def _radices(n, nb):
    """The pass sequence of mixedRadixFft for length n and nb transforms."""
    rest, p2, e = n, 0, {3: 0, 5: 0, 7: 0, 11: 0, 13: 0}
    while rest % 2 == 0:
        rest //= 2
        p2 += 1
    for p in e:
        while rest % p == 0:
            rest //= p
            e[p] += 1
    assert rest == 1
    single = lambda r: nb * (n // r) <= 256
    out = []
    while e[3] and e[5] and single(15):
        out.append(15)
        e[3] -= 1
        e[5] -= 1
    # the power of two's remainder beyond its radix-16 passes with odd factors
    for rem, p, cnt, r in ((1, 3, 2, 18), (1, 3, 1, 6), (2, 3, 1, 12), (1, 5, 1, 10), (1, 7, 1, 14)):
        if p2 % 4 == rem and e[p] >= cnt and single(r):
            out.append(r)
            e[p] -= cnt
            p2 -= rem
            break
    while e[3] >= 2 and single(9):
        out.append(9)
        e[3] -= 2
    for p in (3, 5, 7, 11, 13):
        out += [p] * e[p]
    out += [16] * (p2 // 4)
    if p2 % 4:
        out.append(1 << (p2 % 4))
    return out


def _stockham(x, radices, sign):
    n = len(x)
    tw = np.exp(-2j * np.pi * np.arange(n) / n)
    buf, ns = x.astype(complex), 1
    for r in radices:
        nr, step, out = n // r, n // (ns * r), np.zeros(n, complex)
        for jj in range(nr):
            k = jj % ns
            v = np.array([buf[jj + q * nr] * (tw[q * k * step] if sign < 0 else np.conj(tw[q * k * step]))
                          for q in range(r)])
            v = _dft(v, sign)
            d = (jj // ns) * ns * r + k
            for q in range(r):
                out[d + q * ns] = v[q]
        buf, ns = out, ns * r
    return buf


def test_radix9_cooley_tukey():
    """dftCt9: n = 3 n1 + n2, k = k1 + 3 k2, DFT3 along n1, twiddles
    W9^(n2 k1), DFT3 along n2."""
    rng = np.random.default_rng(9)
    x = rng.standard_normal(9) + 1j * rng.standard_normal(9)
    for sign in (-1, 1):
        y = np.zeros((3, 3), complex)
        for n2 in range(3):
            t = _dft(np.array([x[n2], x[3 + n2], x[6 + n2]]), sign)
            for k1 in range(3):
                y[k1, n2] = t[k1] * np.exp(sign * 2j * np.pi * n2 * k1 / 9)
        out = np.zeros(9, complex)
        for k1 in range(3):
            t = _dft(y[k1], sign)
            for k2 in range(3):
                out[k1 + 3 * k2] = t[k2]
        assert np.abs(out - _dft(x, sign)).max() < 1e-12


@pytest.mark.parametrize("n,nb", [(240, 4), (240, 16), (96, 8), (192, 4), (200, 8), (288, 8), (60, 2),
                                  (120, 8), (45, 4), (39, 8), (26, 2), (22, 2), (768, 8), (144, 8), (18, 4),
                                  (36, 8), (576, 4), (90, 8), (320, 8), (384, 8), (20, 4), (24, 8)])
def test_mixed_radix_schedule_matches_numpy(n, nb):
    radices = _radices(n, nb)
    assert np.prod(radices) == n
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(_stockham(x, radices, -1) - np.fft.fft(x)).max() < 1e-10 * np.abs(np.fft.fft(x)).max()
    assert np.abs(_stockham(x, radices, +1) - np.fft.ifft(x) * n).max() < 1e-10 * np.abs(np.fft.fft(x)).max()
