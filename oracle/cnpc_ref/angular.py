"""Angular and binning helpers of ``math_evaluator.py`` restated in numpy
(oracle; test infrastructure only).

Spherical harmonics, Legendre polynomials, Wigner 3j symbols and associated
Legendre functions are built with the sympy library exactly as the reference
builds them (same symbolic construction), then evaluated with numpy instead of
numexpr.  They are pinned against scipy's spherical harmonics, the addition
theorem, the Racah formula and closed forms in ``tests/test_oracle_angular.py``.
"""
import functools

import numpy as np
import sympy as sp


@functools.lru_cache(maxsize=None)
def _ylm_expr(l, m, racah):
    """Symbolic Y_lm in (xhat, yhat, zhat): ``get_Ylm`` (``math_evaluator.py:35-131``)."""
    l = int(l)
    m = int(m)
    if abs(m) > l:
        raise ValueError("abs(m) must be <= l")
    x, y, z = sp.symbols("x y z ", real=True)
    r = sp.symbols("r", real=True, positive=True)
    xhat, yhat, zhat = sp.symbols("xhat yhat zhat", real=True)
    phi, theta = sp.symbols("phi theta")
    defs = [
        (sp.sin(phi), y / sp.sqrt(x ** 2 + y ** 2)),
        (sp.cos(phi), x / sp.sqrt(x ** 2 + y ** 2)),
        (sp.cos(theta), z / sp.sqrt(x ** 2 + y ** 2 + z ** 2)),
    ]
    amp = sp.sqrt((2 * l + 1) / (4 * np.pi) * sp.factorial(l - abs(m)) / sp.factorial(l + abs(m)))
    expr = sp.assoc_legendre(l, abs(m), sp.cos(theta))
    expr *= (sp.cos(phi) + sp.I * sp.sin(phi)) ** abs(m)
    expr = expr.subs(defs)
    expr = sp.together(expr).subs(x ** 2 + y ** 2 + z ** 2, r ** 2)
    expr = amp * expr.expand().subs([(x / r, xhat), (y / r, yhat), (z / r, zhat)])
    if racah:
        expr = expr * sp.sqrt(4 * np.pi / (2 * l + 1))
    if m < 0:
        expr = (-1) ** m * sp.conjugate(expr)
    expr = sp.simplify(expr)
    return expr, (xhat, yhat, zhat)


def get_Ylm(l, m, Racah_normalized=False):
    """Callable complex Y_lm(xhat, yhat, zhat) (``math_evaluator.py:35-131``)."""
    expr, syms = _ylm_expr(int(l), int(m), bool(Racah_normalized))
    f = sp.lambdify(syms, expr, "numpy")

    def ylm(xh, yh, zh):
        out = f(xh, yh, zh)
        return np.broadcast_to(np.asarray(out, dtype=np.complex128), np.broadcast(xh, yh, zh).shape)

    ylm.expr = expr
    return ylm


def get_legendre(ell, r_xhat, r_yhat, r_zhat):
    """L_ell(r.khat) as a callable (``math_evaluator.py:134-197``)."""
    ell = int(ell)
    if ell < 0:
        raise ValueError("ell must be non-negative")
    if not np.isclose(r_xhat ** 2 + r_yhat ** 2 + r_zhat ** 2, 1.0):
        raise ValueError("(r_xhat, r_yhat, r_zhat) must form a unit vector")
    kx, ky, kz = sp.symbols("k_xhat k_yhat k_zhat", real=True)
    c = sp.Symbol("cos_theta", real=True)
    expr = sp.simplify(sp.legendre(ell, c).subs(c, r_xhat * kx + r_yhat * ky + r_zhat * kz))
    f = sp.lambdify((kx, ky, kz), expr, "numpy")

    def leg(a, b, d):
        return np.broadcast_to(np.asarray(f(a, b, d), dtype=np.float64), np.broadcast(a, b, d).shape)

    return leg


def get_kgrid(k):
    """khat grid and |k| (``math_evaluator.py:12-22``): khat(0) := 0."""
    kgrid = [np.real(kk) for kk in k]
    knorm = np.sqrt(sum(np.real(kk) ** 2 for kk in k))
    full = np.broadcast_shapes(*[kk.shape for kk in kgrid])
    knorm = np.broadcast_to(knorm, full)
    with np.errstate(invalid="ignore", divide="ignore"):
        out = [np.broadcast_to(kk, full) / knorm for kk in kgrid]
    if knorm[0, 0, 0] == 0.0:
        for kk in out:
            kk[0, 0, 0] = 0.0
    return out, knorm


def get_xgrid(x, boxcenter, boxsize, nmesh):
    """xhat of cell centres (``math_evaluator.py:25-32``): +1/2 cell + boxcenter."""
    off = 0.5 * np.asarray(boxsize, dtype=float) / np.asarray(nmesh) + np.asarray(boxcenter, dtype=float)
    xg = [np.real(x[i]) + off[i] for i in range(3)]
    xnorm = np.sqrt(sum(xx ** 2 for xx in xg))
    return [xx / xnorm for xx in xg]


# ---------------------------------------------------------------- compensation
def _sinc(w):
    return np.sinc(0.5 * w / np.pi)


_P = {"ngp": 1, "cic": 2, "tsc": 3, "pcs": 4}


def compensate_sinc(w, v, sampler):
    """CompensateNGP/CIC/TSC/PCS (``math_evaluator.py:206-248``)."""
    p = _P[sampler]
    for i in range(3):
        v = v / _sinc(w[i]) ** p
    return v


def _alias_c(s, sampler):
    if sampler == "cic":
        return 1 - 2.0 / 3 * s
    if sampler == "tsc":
        return 1 - s + 2.0 / 15 * s ** 2
    if sampler == "pcs":
        return 1 - 4.0 / 3.0 * s + 2.0 / 5.0 * s ** 2 - 4.0 / 315.0 * s ** 3
    return np.ones_like(s)


def compensate_shotnoise(w, v, sampler):
    """Compensate*Shotnoise (``math_evaluator.py:251-275``): / sqrt(C_p(w))."""
    if sampler == "ngp":
        return v
    for i in range(3):
        s = np.sin(0.5 * w[i]) ** 2
        v = v / _alias_c(s, sampler) ** 0.5
    return v


def compensate_bk_noise(w, v, sampler):
    """Compensate_bk_noise_* (``math_evaluator.py:278-309``): * C_p / sinc^{2p}."""
    if sampler == "ngp":
        return v
    p = _P[sampler]
    for i in range(3):
        s = np.sin(0.5 * w[i]) ** 2
        t = _sinc(w[i]) ** p
        v = v * _alias_c(s, sampler) / t ** 2
    return v


def get_compensation(interlaced, sampler):
    """``get_compensation`` (``math_evaluator.py:980-1004``)."""
    if sampler not in _P:
        raise ValueError("compensation for window %s is not defined" % sampler)
    if interlaced:
        return lambda w, v: compensate_sinc(w, v, sampler)
    return lambda w, v: compensate_shotnoise(w, v, sampler)


# ---------------------------------------------------------------- binning
def get_kbin_count(k_bins, k_edge, knorm):
    """Mode count and sum |k| per bin (``math_evaluator.py:452-481``)."""
    knorm = np.ascontiguousarray(knorm.ravel())
    valid = np.logical_and(knorm >= k_edge[0], knorm < k_edge[-1])
    knorm = knorm[valid]
    idx = np.searchsorted(k_edge, knorm, side="right") - 1
    return (np.bincount(idx, minlength=k_bins).astype("f8"),
            np.bincount(idx, weights=knorm, minlength=k_bins).astype("f8"))


def radial_binning(kfield, k_bins, k_edge, knorm):
    """Per-bin sum of a field (``math_evaluator.py:504-539``)."""
    kfield = np.ascontiguousarray(kfield.ravel())
    knorm = np.ascontiguousarray(knorm.ravel())
    valid = np.logical_and(knorm >= k_edge[0], knorm < k_edge[-1])
    knorm = knorm[valid]
    kfield = kfield[valid]
    idx = np.searchsorted(k_edge, knorm, side="right") - 1
    if np.iscomplexobj(kfield):
        out = (np.bincount(idx, weights=np.real(kfield), minlength=k_bins)
               + 1j * np.bincount(idx, weights=np.imag(kfield), minlength=k_bins))
        return out.astype(kfield.dtype, copy=False)
    return np.bincount(idx, weights=kfield, minlength=k_bins).astype(kfield.dtype, copy=False)


# ---------------------------------------------------------------- 3j / q_ell
def wigner_3j(j1, j2, j3, m1, m2, m3):
    from sympy.physics.wigner import wigner_3j as w3j
    return np.float64(w3j(j1, j2, j3, m1, m2, m3).evalf())


def get_magnetic_configs_box(ell_1, ell_2, L):
    """(m1, m2, 0) with m1 >= 0 (``math_evaluator.py:312-331``)."""
    cfg, tj = [], []
    for m1 in range(0, ell_1 + 1):
        for m2 in range(-ell_2, ell_2 + 1):
            if m1 + m2 == 0:
                cfg.append((m1, m2, 0))
                tj.append(wigner_3j(ell_1, ell_2, L, m1, m2, 0))
    return cfg, tj


def get_magnetic_configs_survey(ell_1, ell_2, L):
    """(m1, m2, M) up to and including (0,0,0) (``math_evaluator.py:334-355``)."""
    cfg, tj = [], []
    for M in range(-L, L + 1):
        for m1 in range(-ell_1, ell_1 + 1):
            m2 = -M - m1
            if -ell_2 <= m2 <= ell_2:
                cfg.append((m1, m2, M))
                tj.append(wigner_3j(ell_1, ell_2, L, m1, m2, M))
                if (m1, m2, M) == (0, 0, 0):
                    return cfg, tj
    return cfg, tj


@functools.lru_cache(maxsize=None)
def _assoc_legendre_fn(ell, m):
    mu = sp.Symbol("mu")
    return sp.lambdify(mu, sp.assoc_legendre(ell, m, mu), "numpy")


def assoc_legendre(ell, m, mu):
    """sympy ``assoc_legendre(ell, m, mu)`` (Condon-Shortley phase included)."""
    return float(np.asarray(_assoc_legendre_fn(int(ell), int(m))(mu), dtype=np.float64))


def get_associated_legendre_coefficients_eff(ell, m, k1, k2, k3_eff, mode="13"):
    """``math_evaluator.py:418-446``."""
    res = np.zeros(len(k3_eff))
    for i, k3 in enumerate(k3_eff):
        if abs(k1 - k2) <= k3 <= (k1 + k2):
            if mode == "13":
                mu = (k3 ** 2 + k1 ** 2 - k2 ** 2) / (2 * k1 * k3)
            else:
                mu = (k3 ** 2 + k2 ** 2 - k1 ** 2) / (2 * k2 * k3)
            res[i] = assoc_legendre(ell, m, mu)
    res *= (-1) ** (ell + m)
    return res


def get_associated_legendre_coefficients(ell, m, k1, k2, k_min, k_max, kbin, mode="13"):
    """``math_evaluator.py:389-415``: as the effective-k form, at the k3 bin
    centres of linspace(2 k_min, 2 k_max, 2 kbin + 1)[:-1]."""
    e = np.linspace(2 * k_min, 2 * k_max, kbin * 2 + 1)[:-1]
    c = 0.5 * (e[1:] + e[:-1])
    res = np.zeros(len(c))
    for i, k3 in enumerate(c):
        if abs(k1 - k2) <= k3 <= (k1 + k2):
            if mode == "13":
                mu = (k3 ** 2 + k1 ** 2 - k2 ** 2) / (2 * k1 * k3)
            else:
                mu = (k3 ** 2 + k2 ** 2 - k1 ** 2) / (2 * k2 * k3)
            res[i] = assoc_legendre(ell, m, mu)
    res *= (-1) ** (ell + m)
    return res


def get_q_ells(i, j, k_center, k_min, k_max, k_bins, ell_1, ell_2, L, k3_bins):
    """``get_q_ells`` (``math_evaluator.py:560-589``): q at the bin centres
    (``shotnoise_ana_all_eff=False``)."""
    q = np.zeros(k3_bins).astype("f8")
    lmin = min(ell_1, ell_2)
    for xx in range(-lmin, lmin + 1):
        tj = wigner_3j(ell_1, ell_2, L, xx, -xx, 0)
        a13 = get_associated_legendre_coefficients(ell_1, xx, k_center[i], k_center[j], k_min, k_max, k_bins, "13")
        a23 = get_associated_legendre_coefficients(ell_2, -xx, k_center[i], k_center[j], k_min, k_max, k_bins, "23")
        q += (-1) ** xx * tj * a13 * a23
    return q


def get_valid_k3_bins(k1, k2, k_min, k_max, kbin):
    """``math_evaluator.py:448-452``."""
    e = np.linspace(2 * k_min, 2 * k_max, kbin * 2 + 1)[:-1]
    c = 0.5 * (e[1:] + e[:-1])
    return np.logical_and(c >= abs(k1 - k2), c <= (k1 + k2))


def get_q_ells_eff(i, j, k_eff, ell_1, ell_2, L, k3_bins, k3_eff):
    """``get_q_ells_eff`` (``math_evaluator.py:592-616``)."""
    q = np.zeros(k3_bins).astype("f8")
    lmin = min(ell_1, ell_2)
    for xx in range(-lmin, lmin + 1):
        tj = wigner_3j(ell_1, ell_2, L, xx, -xx, 0)
        a13 = get_associated_legendre_coefficients_eff(ell_1, xx, k_eff[i], k_eff[j], k3_eff, "13")
        a23 = get_associated_legendre_coefficients_eff(ell_2, -xx, k_eff[i], k_eff[j], k3_eff, "23")
        q += (-1) ** xx * tj * a13 * a23
    return q
