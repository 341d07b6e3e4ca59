"""CPU checks of the algebra behind two kernel forms (no GPU):

* the Rader stage of p = 113 with its FFT-112 in the Good-Thomas form, using the
  index maps and the (k1, k2) order of G that mm_get_rader113 builds
  (csrc/fft_plan.hip, csrc/fft_rader.h), against numpy.fft;
* the length-80 z pass (csrc/zfft80.h, k_tile_zfft): reflection padding with
  zpad = sz - 1 as Toeplitz + Hankel circular convolutions, against the z-kernel
  matrix built from the padding copies as k_tile_zconv builds it.
"""
import numpy as np
import pytest

P, L = 113, 112


def _rader_tables():
    gam = next(c for c in range(2, P) if len({pow(c, e, P) for e in range(1, P)}) == P - 1)
    ginv = pow(gam, P - 2, P)
    pw = [pow(gam, a, P) for a in range(L)]
    slot_in = [0] * P
    for a in range(L):
        slot_in[pow(ginv, a, P)] = (4 * a) % 7 + 7 * ((7 * a) % 16)
    slot_in[0] = L
    perm_out = [0] * L
    for m1 in range(16):
        for m2 in range(7):
            perm_out[m2 + 7 * m1] = pw[(7 * m1 + 16 * m2) % L]
    return pw, slot_in, perm_out


def _ghat(pw, sign):
    G = np.fft.fft(np.exp(sign * 2j * np.pi * np.array(pw) / P)) / L
    h = np.zeros(L, complex)
    for k1 in range(16):
        for k2 in range(7):
            h[k1 + 16 * k2] = G[(49 * k1 + 64 * k2) % L]
    return h


def _rader(x, inv):
    """The three LDS stages of rd_stages on one transform (slots of 114 values)."""
    pw, slot_in, perm_out = _rader_tables()
    gh = _ghat(pw, 1.0 if inv else -1.0)
    dft = np.fft.fft
    idft = lambda v: np.fft.ifft(v) * len(v)
    A = np.zeros(114, complex)
    for r in range(P):
        A[slot_in[r]] = x[r]
    B = np.zeros(114, complex)
    for q in range(7):  # DFT-16 over n1, written at k1 * 7 + n2
        v = dft(A[[q + 7 * r for r in range(16)]])
        B[[7 * k1 + q for k1 in range(16)]] = v
    A2 = A.copy()
    for q in range(16):  # DFT-7 over n2, times G, inverse DFT-7
        v = dft(B[[q * 7 + r for r in range(7)]])
        if q == 0:
            A2[113] = A[112] + v[0]
        A2[[q * 7 + r for r in range(7)]] = idft(v * gh[[q + 16 * r for r in range(7)]])
    Y = np.zeros(P, complex)
    for q in range(7):  # inverse DFT-16 over k1 -> C_b at perm_out
        v = idft(A2[[q + 7 * r for r in range(16)]])
        Y[[perm_out[q + 7 * r] for r in range(16)]] = A2[112] + v
    Y[0] = A2[113]
    return Y


@pytest.mark.parametrize("inv", [False, True])
def test_rader_good_thomas_tables(inv):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(P) + 1j * rng.standard_normal(P)
    want = np.fft.ifft(x) * P if inv else np.fft.fft(x)
    got = _rader(x, inv)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def _zmap(sz, zpad):
    """Padded plane -> data plane of the sequential reflections (deconvolver.py:131-144)."""
    Zp = sz + 2 * zpad
    m = [-1] * Zp
    for i in range(sz):
        m[zpad + i] = i
    for j in range(zpad):
        s = 2 * zpad - j
        if s < Zp:
            m[j] = m[s] if s < zpad + sz else -1
    top = zpad + sz - 1
    for j in range(top + 1, Zp):
        s = 2 * top - j
        if s >= 0:
            m[j] = m[s]
    return m


def _zpass_matrix(w, sz, zpad):
    Zp = sz + 2 * zpad
    K = np.zeros((sz, sz), complex)
    for p, s in enumerate(_zmap(sz, zpad)):
        if s >= 0:
            for t in range(sz):
                K[t, s] += w[(zpad + t - p) % Zp]
    return K


def _zpass_fft80(w, x, sz):
    Zp = 3 * sz - 2
    a = np.zeros(80, complex)
    for m in range(-(sz - 1), sz):
        a[m % 80] = w[m % Zp]
    g = np.zeros(80, complex)
    for m in range(2 * sz - 1):
        g[m] = w[m] + w[(m - 2 * (sz - 1)) % Zp]
    A, G = np.fft.fft(a) / 80, np.fft.fft(g) / 80
    X = np.fft.fft(np.concatenate([x, np.zeros(80 - sz)]))
    y = np.fft.ifft(A * X + G * X[(-np.arange(80)) % 80]) * 80
    c0 = w[:sz]
    c1 = np.array([w[(t - sz + 1) % Zp] for t in range(sz)])
    return y[:sz] - c0 * x[0] - c1 * x[sz - 1]


@pytest.mark.parametrize("sz", [2, 3, 7, 17, 39, 40])
def test_zpass_toeplitz_hankel_fft80(sz):
    rng = np.random.default_rng(sz)
    Zp = 3 * sz - 2
    w = rng.standard_normal(Zp) + 1j * rng.standard_normal(Zp)
    x = rng.standard_normal(sz) + 1j * rng.standard_normal(sz)
    want = _zpass_matrix(w, sz, sz - 1) @ x
    got = _zpass_fft80(w, x, sz)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()

