"""CPU checks of the algebra behind two kernel forms (no GPU):

* the Rader stage of p = 113 (FFT-112 as 16 x 7) and p = 239 (FFT-238 as 17 x 14)
  in the Good-Thomas form, using the index maps and the (k1, k2) order of G that
  mm_get_rader builds (csrc/fft_plan.hip, csrc/fft_rader.h), against numpy.fft,
  and the DFT-14 of the p = 239 stages as 2 x 7;
* the length-80 z pass (csrc/zfft80.h, k_tile_zfft): reflection padding with
  zpad = sz - 1 as Toeplitz + Hankel circular convolutions, against the z-kernel
  matrix built from the padding copies as k_tile_zconv builds it.
"""
import numpy as np
import pytest

SPECS = {113: (16, 7), 239: (17, 14)}  # p: (A, B), p - 1 = A * B


def _rader_tables(P):
    A, B = SPECS[P]
    L = P - 1
    gam = next(c for c in range(2, P) if len({pow(c, e, P) for e in range(1, P)}) == P - 1)
    ginv = pow(gam, P - 2, P)
    pw = [pow(gam, a, P) for a in range(L)]
    iBA, iAB = pow(B, -1, A), pow(A, -1, B)
    slot_in = [0] * P
    for a in range(L):
        slot_in[pow(ginv, a, P)] = (a * iAB) % B + B * ((a * iBA) % A)
    slot_in[0] = L
    perm_out = [0] * L
    for m1 in range(A):
        for m2 in range(B):
            perm_out[m2 + B * m1] = pw[(B * m1 + A * m2) % L]
    return pw, slot_in, perm_out


def _ghat(P, pw, sign):
    A, B = SPECS[P]
    L = P - 1
    cA, cB = B * pow(B, -1, A) % L, A * pow(A, -1, B) % L
    G = np.fft.fft(np.exp(sign * 2j * np.pi * np.array(pw) / P)) / L
    h = np.zeros(L, complex)
    for k1 in range(A):
        for k2 in range(B):
            h[k1 + A * k2] = G[(cA * k1 + cB * k2) % L]
    return h


def _rader(P, x, inv):
    """The three in-place LDS stages of rd_stages on one transform (slots of L + 2 values)."""
    A, B = SPECS[P]
    L = P - 1
    pw, slot_in, perm_out = _rader_tables(P)
    gh = _ghat(P, pw, 1.0 if inv else -1.0)
    dft = np.fft.fft
    idft = lambda v: np.fft.ifft(v) * len(v)
    S = np.zeros(L + 2, complex)
    for r in range(P):
        S[slot_in[r]] = x[r]
    for q in range(B):  # DFT-A over n1 at slots q + B n1, written back in place as k1
        S[[q + B * r for r in range(A)]] = dft(S[[q + B * r for r in range(A)]])
    for q in range(A):  # DFT-B over n2, times G, inverse DFT-B
        v = dft(S[[q * B + r for r in range(B)]])
        if q == 0:
            S[L + 1] = S[L] + v[0]
        S[[q * B + r for r in range(B)]] = idft(v * gh[[q + A * r for r in range(B)]])
    Y = np.zeros(P, complex)
    for q in range(B):  # inverse DFT-A over k1 -> C_b at perm_out
        v = idft(S[[q + B * r for r in range(A)]])
        Y[[perm_out[q + B * r] for r in range(A)]] = S[L] + v
    Y[0] = S[L + 1]
    return Y


@pytest.mark.parametrize("P", [113, 239])
@pytest.mark.parametrize("inv", [False, True])
def test_rader_good_thomas_tables(P, inv):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(P) + 1j * rng.standard_normal(P)
    want = np.fft.ifft(x) * P if inv else np.fft.fft(x)
    got = _rader(P, x, inv)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_dft14_good_thomas():
    """bfly14: n = 7 a + 2 b, k = 7 k1 + 8 k2 (mod 14); DFT-2 over a, then DFT-7 over b."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal(14) + 1j * rng.standard_normal(14)
    u = np.zeros((2, 7), complex)
    for b in range(7):
        u[:, b] = np.fft.fft([x[(2 * b) % 14], x[(7 + 2 * b) % 14]])
    u = np.fft.fft(u, axis=1)
    got = np.zeros(14, complex)
    for k1 in range(2):
        for k2 in range(7):
            got[(7 * k1 + 8 * k2) % 14] = u[k1, k2]
    np.testing.assert_allclose(got, np.fft.fft(x), atol=1e-12)


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


def _zpass_fft80_ends(w, x, sz):
    """The same operator with the end-plane corrections in the frequency domain (k_tile_zfft_own): the Toeplitz
    term acts on x without its planes 0 and sz - 1, Y = A (X - x0 - x[sz-1] e^{-2 pi i (sz-1) k / 80}) + G X[-k]."""
    Zp = 3 * sz - 2
    a = np.zeros(80, complex)
    for m in range(-(sz - 1), sz):
        a[m % 80] = w[m % Zp]
    g = np.zeros(80, complex)
    for m in range(2 * sz - 1):
        g[m] = w[m] + w[(m - 2 * (sz - 1)) % Zp]
    A, G = np.fft.fft(a) / 80, np.fft.fft(g) / 80
    X = np.fft.fft(np.concatenate([x, np.zeros(80 - sz)]))
    k = np.arange(80)
    Xt = X - x[0] - x[sz - 1] * np.exp(-2j * np.pi * ((sz - 1) * k % 80) / 80)
    y = np.fft.ifft(A * Xt + G * X[(-k) % 80]) * 80
    return y[:sz]


@pytest.mark.parametrize("sz", [1, 2, 3, 7, 17, 39, 40])
def test_zpass_fft80_end_planes_in_frequency_domain(sz):
    rng = np.random.default_rng(100 + sz)
    Zp = 3 * sz - 2
    w = rng.standard_normal(Zp) + 1j * rng.standard_normal(Zp)
    x = rng.standard_normal(sz) + 1j * rng.standard_normal(sz)
    want = _zpass_matrix(w, sz, sz - 1) @ x
    got = _zpass_fft80_ends(w, x, sz)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()

