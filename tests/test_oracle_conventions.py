"""Pins the oracle's pmesh conventions (paint windows, interlacing, FFT
normalisation, axes) with closed forms and invariants."""
import numpy as np
import pytest
from scipy.interpolate import BSpline

from oracle.cnpc_ref import conventions as C

ORDERS = {"ngp": 1, "cic": 2, "tsc": 3, "pcs": 4}


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_windows_are_cardinal_bsplines(order):
    """W_p(s) equals the centred cardinal B-spline of order p (the p-fold
    convolution of the unit box), whose Fourier transform is sinc^p -- the
    shape the reference's compensations divide out."""
    rng = np.random.default_rng(order)
    u = rng.uniform(-50, 50, 2000)
    anchor, ws = C.window_weights(u, order)
    bs = BSpline.basis_element(np.arange(order + 1) - order / 2.0, extrapolate=False)
    for j, w in enumerate(ws):
        s = u - (anchor + j)
        expect = np.nan_to_num(bs(s))
        if order == 1:
            expect = (np.abs(s) <= 0.5).astype(float)
        np.testing.assert_allclose(w, expect, atol=1e-13)
    np.testing.assert_allclose(sum(ws), 1.0, atol=1e-13)  # partition of unity


def test_single_particle_known_answers():
    n, L = 8, 8.0
    m = C.paint([[2.25, 0.0, 0.0]], [1.0], n, L, "cic").real
    assert m[2, 0, 0] == pytest.approx(0.75) and m[3, 0, 0] == pytest.approx(0.25)
    m = C.paint([[3.0, 3.0, 3.0]], [2.0], n, L, "tsc").real
    assert m[3, 3, 3] == pytest.approx(2 * 0.75 ** 3)
    assert m[2, 3, 3] == pytest.approx(2 * 0.125 * 0.75 ** 2)
    m = C.paint([[0.0, 0.0, 0.0]], [1.0], n, L, "pcs").real  # wraps to the far side
    assert m[0, 0, 0] == pytest.approx((2 / 3) ** 3)
    assert m[7, 0, 0] == pytest.approx((1 / 6) * (2 / 3) ** 2)
    assert m[2, 0, 0] == 0.0
    m = C.paint([[7.6, 0.2, 0.0]], [1.0], n, L, "ngp").real
    assert m[0, 0, 0] == 1.0  # nearest point of 7.6 is 8 == 0 (periodic)
    # the interlacing copy sits half a cell up in every axis
    m = C.paint([[2.0, 0.0, 0.0]], [1.0], n, L, "cic", shift=0.5).real
    for ix in (2, 3):
        for iy in (0, 1):
            for iz in (0, 1):
                assert m[ix, iy, iz] == pytest.approx(0.125)


@pytest.mark.parametrize("sampler", list(ORDERS))
def test_mass_conservation(sampler):
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, 100.0, (500, 3))
    w = rng.uniform(0.5, 1.5, 500)
    for shift in (0.0, 0.5):
        m = C.paint(pos, w, 16, 100.0, sampler, shift=shift)
        assert m.real.sum() == pytest.approx(w.sum(), rel=1e-13)


def test_fft_conventions():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((8, 8, 8))
    c = C.r2c(x)
    assert c[0, 0, 0] == pytest.approx(x.mean())  # forward normalised by 1/N^3
    np.testing.assert_allclose(C.c2r(c).real, x, atol=1e-13)
    # a plane wave e^{+i k.x} maps to one unit coefficient at +k
    n, L = 8, 8.0
    k = C.k_axes(n, L)
    xx = C.x_axes(n, L)
    kk = 2 * np.pi / L * np.array([1, -2, 3])
    field = np.exp(1j * (kk[0] * np.arange(n)[:, None, None] + kk[1] * np.arange(n)[None, :, None]
                         + kk[2] * np.arange(n)[None, None, :]) * (L / n))
    c = C.r2c(field)
    assert abs(c[1, -2, 3] - 1) < 1e-13 and np.sum(np.abs(c)) == pytest.approx(1.0)
    assert k[0][n // 2, 0, 0] == pytest.approx(-np.pi * n / L)  # Nyquist is negative
    assert xx[0][n - 1, 0, 0] == pytest.approx(-L / n)  # x wrapped to [-L/2, L/2)


def test_interlacing_keeps_even_alias_images():
    """Interlaced spectrum of one PCS particle equals the alias sum over
    image vectors m with even sum(m) -- pins the +i kH/2 phase sign."""
    n, L = 16, 16.0
    H = L / n
    x0 = np.array([3.3, 7.85, 11.1])
    c = C.r2c(C.pm_painter([x0], [1.0], n, L, "pcs", True))
    M = 25
    kf = 2 * np.pi / L
    nn = C.int_axis(n)
    mm = np.arange(-M, M + 1)
    terms_all, terms_alt = [], []
    for d in range(3):
        kd = (nn[:, None] + mm[None, :] * n) * kf  # (n, 2M+1)
        w = np.sinc(kd * H / (2 * np.pi)) ** 4 * np.exp(-1j * kd * x0[d])
        terms_all.append(w.sum(1))
        terms_alt.append((w * (-1.0) ** mm[None, :]).sum(1))
    a = terms_all[0][:, None, None] * terms_all[1][None, :, None] * terms_all[2][None, None, :]
    b = terms_alt[0][:, None, None] * terms_alt[1][None, :, None] * terms_alt[2][None, None, :]
    expect = 0.5 * (a + b) / n ** 3
    sel = (np.abs(nn) < n // 2)
    err = np.abs(c - expect)[np.ix_(sel, sel, sel)].max()
    assert err < 1e-6 * np.abs(expect).max()
    # the opposite phase sign would keep odd images too
    c_wrong = C.r2c(C.paint([x0], [1.0], n, L, "pcs")) * 0.5 + 0.5 * C.r2c(
        C.paint([x0], [1.0], n, L, "pcs", shift=0.5)) * np.exp(-0.5j * sum(kk * H for kk in C.k_axes(n, L)))
    assert np.abs(c_wrong - expect)[np.ix_(sel, sel, sel)].max() > 1e-3 * np.abs(expect).max()
