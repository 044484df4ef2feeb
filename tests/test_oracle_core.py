"""Pins the CPU oracle: RNG known answers, FFT against numpy golden vectors,
k-grid conventions, spline weights against closed forms.  CPU only."""
import os
from fractions import Fraction

import numpy as np
import pytest

import orc

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


# --------------------------------------------------------------- Philox ----
# Known-answer vectors published with Random123 (kat_vectors, philox4x32 R=10).
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_kat(ctr, key, want):
    assert list(orc.philox(ctr, key)) == want


def test_philox_numpy_matches_c():
    c0 = np.arange(64, dtype=np.uint64)
    r = orc.philox_np(c0, 7, orc.TAG_TILD, 3, (11 << 32) | 5)
    for i in range(0, 64, 9):
        assert [int(v[i]) for v in r] == list(orc.philox([i, 7, orc.TAG_TILD, 3], [5, 11]))


def test_normals_moments():
    z = orc.normal4_np(np.arange(200000, dtype=np.uint64), 1, orc.TAG_FTS, 0, 42).reshape(-1)
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.01
    # fourth moment of a standard normal is 3
    assert abs(np.mean(z ** 4) - 3.0) < 0.05


# ------------------------------------------------------------------ FFT ----
@pytest.mark.parametrize("n", [8, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 7, 11, 25])
def test_fft1d_golden(n):
    x = GOLD[f"x1_{n}"]
    a = x.copy()
    orc.lib().orc_fft1d(orc._ptr(a), n, 1, -1)
    np.testing.assert_allclose(a, GOLD[f"fwd1_{n}"], rtol=0, atol=1e-12 * n)
    b = x.copy()
    orc.lib().orc_fft1d(orc._ptr(b), n, 1, +1)
    np.testing.assert_allclose(b, GOLD[f"inv1_{n}"], rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("shp", ["4x6x8", "8x12x16", "16x8x24"])
def test_fft3d_golden(shp):
    x = GOLD[f"x3_{shp}"]
    np.testing.assert_allclose(orc.fft3d(x, -1), GOLD[f"fwd3_{shp}"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(orc.fft3d(x, +1), GOLD[f"inv3_{shp}"], rtol=0, atol=1e-11)


def test_box_fft_convention_and_parseval():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(8, 12, 16)) + 0j
    Nx = (16, 12, 8)
    f = orc.box_fft(x, Nx, 1)
    # forward scaled by 1/M: Parseval sum|x|^2 = M sum|f|^2
    assert np.isclose(np.sum(np.abs(x) ** 2), x.size * np.sum(np.abs(f) ** 2), rtol=1e-12)
    back = orc.box_fft(f, Nx, -1)
    np.testing.assert_allclose(back, x, atol=1e-12)


def test_fp32_oracle_fft_close():
    x = GOLD["x3_8x12x16"]
    a = orc.fft3d(x, -1, bits=32)
    assert np.max(np.abs(a - GOLD["fwd3_8x12x16"])) < 1e-5 * np.max(np.abs(GOLD["fwd3_8x12x16"]))


# --------------------------------------------------------------- k grid ----
def test_kvec_nyquist_negative():
    Nx, L = (4, 6, 8), (2.0, 3.0, 5.0)
    # index with n = (2, 3, 4): every component at Nyquist -> negative frequency
    i = 2 + 4 * (3 + 6 * 4)
    k2, k = orc.kvec_f(i, Nx, L)
    PI2 = np.float32(6.2831853071)
    want = [np.float32(PI2 * np.float32(n - N)) / np.float32(l) for n, N, l in zip((2, 3, 4), Nx, L)]
    assert list(k) == want
    assert all(v < 0 for v in k)
    # n = 1 is positive, n = N-1 is -1
    k2, k = orc.kvec_f(1 + 4 * (5 + 6 * 0), Nx, L)
    assert k[0] > 0 and k[1] < 0 and k[2] == 0
    k2d, kd = orc.kvec_d(1 + 4 * (5 + 6 * 0), Nx, L)
    assert np.allclose(kd[:3], k[:3].astype(np.float64), rtol=0, atol=0)


# --------------------------------------------------------------- splines ---
def bspline(t, order):
    """Cardinal B-spline of the given order (support order) centred at 0,
    by the recursion M_n(t) = ((n/2 + t) M_{n-1}(t + 1/2) + (n/2 - t) M_{n-1}(t - 1/2)) / (n - 1)."""
    if order == 1:
        return 1.0 if -0.5 <= t < 0.5 else 0.0
    n = order
    return ((n / 2 + t) * bspline(t + 0.5, n - 1) + (n / 2 - t) * bspline(t - 0.5, n - 1)) / (n - 1)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_spline_weights_are_cardinal_bsplines(p):
    # ps_grid_utils.cu: point i of the stencil sits at offset (i - p/2) - s0 from
    # the particle, where s0 = s for even p and s + 1/2 for odd p
    for s in np.linspace(-0.4999, 0.4999, 23):
        W = orc.spline_weights(s, p)
        for i in range(p + 1):
            if p % 2 == 0:
                t = (i - p // 2) - s
            else:
                t = (i - p // 2) - (s + 0.5)
            assert abs(W[i] - bspline(t, p + 1)) < 1e-13, (p, s, i)


def test_spline_p3_at_node():
    W = orc.spline_weights(-0.5, 3)
    np.testing.assert_allclose(W, [1 / 6, 2 / 3, 1 / 6, 0.0], atol=1e-15)


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4])
def test_spline_partition_of_unity_and_first_moment(p):
    for s in np.linspace(-0.5, 0.5, 11):
        W = orc.spline_weights(s, p)
        assert abs(W.sum() - 1.0) < 1e-13
        if p > 0:
            centre = np.arange(p + 1) - p / 2.0
            assert abs(np.dot(centre, W) - s) < 1e-13


def test_spline_p4_reference_bug_known_answer():
    # ps_grid_utils.cu:57 uses s^4 for s^3: at s = 1/2, W0 = 1/192 (exact
    # quartic B-spline gives 0) and the third moment is 11/16 instead of 3/4
    s = Fraction(1, 2)
    s2 = s * s
    s3 = s2 * s2
    s4 = s2 * s2
    W0 = (1 - 8 * s + 24 * s2 - 32 * s3 + 16 * s4) / 384
    assert W0 == Fraction(1, 192)
    W = orc.spline_weights(0.5, 4)
    assert abs(W[0] - 1 / 192) < 1e-15
    centre = np.arange(5) - 2.0
    assert abs(np.dot(centre ** 3, W) - 11 / 16) < 1e-13
    assert abs(bspline(-2.5, 5)) < 1e-15  # the true quartic spline vanishes there
