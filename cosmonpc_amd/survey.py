"""Survey-geometry meshes and estimators on the GPU.

Mirrors, with the same arguments, result keys and normalisations:

* ``get_mesh_pk_survey``  (``mesh_generator.py:227-538``)
* ``get_mesh_bk_survey``  (``mesh_generator.py:541-1004``)
* ``get_N_field``         (``mesh_generator.py:1007-1128``)
* ``compute_particle_sums_pk`` / ``_bk`` (``mesh_generator.py:1131-1268``)
* ``calculate_power_spectrum_survey`` + ``get_G_ell``
  (``clustering_estimator.py:205-563``, ``:3104-3134``)
* ``calculate_bk_sugi_survey`` (``clustering_estimator.py:1558-3101``)

A survey field lives in HBM as two full N^3 complex meshes: its x-space values
(the reference's ``rfield``) and its spectrum ``r2c(rfield)``.  Full complex
meshes are needed because the interlaced field is not Hermitian on the Nyquist
planes: the reference keeps the resulting imaginary part of ``rfield``
(pmesh ``dtype="complex128"``), and the x-space Y_lm products below carry it
into every k bin.  Catalog sums run in FP64 on the GPU; the host keeps only
O(k_bins^2) bookkeeping.

GPU-only options (from ``stat_attrs`` / kwargs; the reference ignores them):
  precision   : "fp64" (default) | "fp32"
  contraction : "gemm_fp64" (fp64 default) | "pairwise" | "bf16x3" (fp32
                default) | "fp16x3" | "fp8x6"
"""
import logging
import math

import numpy as np
import torch

from . import angular as A
from .catalog import read_survey_catalog
from .device import PREC, get_context
from .estimators import (ShellBuilder, _config_matrix, _gemm_mode, contraction_grid)

LAST_TIMINGS = {}


def _conj_ylm_poly(ell, m, racah):
    """conj(Y_lm) = (-1)^m Y_l,-m."""
    return A.ylm_poly(ell, -m, racah, factor=(-1.0) ** m)


class SurveyField:
    """A survey field in HBM: ``x`` the full complex mesh of the reference's
    ``rfield`` and ``k`` its spectrum r2c(rfield) (forward FFT / N^3)."""

    def __init__(self, ctx, k, n, boxsize, prec, sampler, interlaced, boxcenter, catalogs=None):
        self.k = k
        self.x = ctx.full(n, prec)
        ctx.fft_c2c(k, self.x, n, prec, +1)
        self.n = n
        self.boxsize = float(boxsize)
        self.prec = prec
        self.sampler = sampler
        self.interlaced = bool(interlaced)
        self.boxcenter = np.asarray(boxcenter, dtype=np.float64)
        self.catalogs = catalogs or {}


def _paint_halves(ctx, cat, w, n, L, sampler, interlaced, prec, boxcenter):
    """rocFFT half spectra of the plain (and shifted) paint of one catalog."""
    m0 = ctx.mesh(n, prec)
    m1 = ctx.mesh(n, prec) if interlaced else None
    ctx.paint(cat.pos, w, n, L, sampler, interlaced, prec, m0, m1, boxcenter)
    s1 = ctx.spectrum(n, prec)
    ctx.r2c(m0, s1, n, prec)
    s2 = None
    if interlaced:
        s2 = ctx.spectrum(n, prec)
        ctx.r2c(m1, s2, n, prec)
    return s1, s2


def _spectrum_of(ctx, parts, n, L, prec, sampler, interlaced):
    """sum_j scale_j * r2c(pm_painter(catalog_j)) as a full spectrum; ``parts``
    holds (s1, s2, scale) from _paint_halves."""
    out = ctx.full(n, prec)
    tmp = None
    for j, (s1, s2, scale) in enumerate(parts):
        if j == 0:
            ctx.full_spectrum(s1, s2, out, n, L, prec, interlaced, 0, sampler, scale / float(n) ** 3)
        else:
            if tmp is None:
                tmp = ctx.full(n, prec)
            ctx.full_spectrum(s1, s2, tmp, n, L, prec, interlaced, 0, sampler, scale / float(n) ** 3)
            ctx.kcombine(tmp, out, n, L, prec, accumulate=True)
    return out


def _xspace(ctx, k, n, prec):
    x = ctx.full(n, prec)
    ctx.fft_c2c(k, x, n, prec, +1)
    return x


# ---------------------------------------------------------------- catalog sums
def compute_particle_sums_pk(ctx, data, randoms, correlation_mode):
    """``mesh_generator.py:1131-1184`` (FP64 on the GPU)."""
    d = ctx.particle_sums(data.pos, data.weight, data.weight_fkp, data.nz)
    r = ctx.particle_sums(randoms.pos, randoms.weight, randoms.weight_fkp, randoms.nz)
    nd, nr = (d[3], r[3]) if correlation_mode == "auto" else (0.0, 0.0)
    return (d[0], r[0], d[1], r[1], d[2], r[2], nd, nr)


def compute_particle_sums_bk(ctx, data, randoms, correlation_mode, angu_config):
    """``mesh_generator.py:1187-1268``: alpha sums, I_33 and
    S_LM[M] = sum (W Wfkp)^3 conj(y_LM(xhat)), M = 0..L (FP64 on the GPU)."""
    L = angu_config[2]
    polys = [_conj_ylm_poly(L, M, True) for M in range(L + 1)] if correlation_mode == "auto" else []
    d = ctx.particle_sums(data.pos, data.weight, data.weight_fkp, data.nz, polys)
    r = ctx.particle_sums(randoms.pos, randoms.weight, randoms.weight_fkp, randoms.nz, polys)
    if polys:
        sd = d[5::2] + 1j * d[6::2]
        sr = r[5::2] + 1j * r[6::2]
    else:
        sd, sr = 0.0, 0.0
    return [d[0], r[0], d[1], r[1], r[4], sd, sr]


def _alpha(sums, alpha_scheme):
    if alpha_scheme == "pypower":
        logging.info("Using same alpha scheme as pypower.")
        return sums[0] / sums[1]
    if alpha_scheme == "nbodykit":
        logging.info("Using same alpha scheme as nbodykit.")
        return sums[2] / sums[3]
    raise ValueError("alpha_scheme must be either 'pypower' or 'nbodykit'.")


class _Reader:
    """Reads (and keeps) the survey catalogs of one task; ``arrays`` may map a
    catalog key to the structured array a FITS read would return."""

    def __init__(self, ctx, catalogs, column_names, z_range, comp_weight_plan, para_cosmo, normalization_scheme,
                 arrays=None):
        self.ctx, self.catalogs, self.arrays = ctx, catalogs, arrays or {}
        self.args = (column_names, z_range, comp_weight_plan, para_cosmo)
        self.ns = normalization_scheme
        self.cache = {}
        self.boxcenter = None

    def box_center(self):
        if self.boxcenter is None:
            cat, bc = self._read("randoms_a", None)
            self.cache["randoms_a"] = cat
            self.boxcenter = bc
        return self.boxcenter

    def _read(self, key, boxcenter):
        column_names, z_range, plan, cosmo = self.args
        ctype = "randoms" if key.startswith("randoms") else "data"
        return read_survey_catalog(self.ctx, self.catalogs.get(key) if self.catalogs else None, column_names, z_range,
                                   plan, cosmo, ctype, self.ns, boxcenter=boxcenter, rows=self.arrays.get(key))

    def get(self, key):
        if key not in self.cache:
            self.cache[key] = self._read(key, self.box_center())
        return self.cache[key]


# ---------------------------------------------------------------- pk meshes
def get_mesh_pk_survey(catalogs, correlation_mode, nmesh, geometry, column_names, boxsize, sampler, interlaced,
                       z_range, comp_weight_plan, para_cosmo, comm=None, normalization_scheme="particle",
                       alpha_scheme="pypower", precision="fp64", arrays=None):
    """``mesh_generator.py:227-538``: returns (mesh_attrs, field_a, field_b)."""
    ctx = get_context()
    prec = PREC[precision]
    n = int(nmesh[0])
    L = float(boxsize[0])
    vpc = float(np.prod(np.asarray(boxsize, dtype=float)) / np.prod(np.asarray(nmesh, dtype=float)))
    rd = _Reader(ctx, catalogs, column_names, z_range, comp_weight_plan, para_cosmo, normalization_scheme, arrays)
    attrs = {"boxcenter": rd.box_center()}
    bc = attrs["boxcenter"]
    ra, da = rd.get("randoms_a"), rd.get("data_a")
    sa = compute_particle_sums_pk(ctx, da, ra, correlation_mode)
    alpha_a = _alpha(sa, alpha_scheme)
    logging.info(f"alpha_a = {alpha_a}")
    I_rand = alpha_a * sa[5]
    I_data = sa[4]
    attrs.update({"alpha_a": alpha_a, "I_rand": I_rand, "I_data": I_data})

    def halves(cat):
        return _paint_halves(ctx, cat, ctx.pair_weights(cat.weight, cat.weight_fkp), n, L, sampler, interlaced, prec,
                             bc)

    hd_a, hr_a = halves(da), halves(ra)
    k_a = _spectrum_of(ctx, [(*hd_a, 1.0 / vpc), (*hr_a, -alpha_a / vpc)], n, L, prec, sampler, interlaced)
    fa = SurveyField(ctx, k_a, n, L, prec, sampler, interlaced, bc)
    if correlation_mode == "auto":
        fb = fa
    else:
        rb, db = rd.get("randoms_b"), rd.get("data_b")
        sb = compute_particle_sums_pk(ctx, db, rb, correlation_mode)
        alpha_b = _alpha(sb, alpha_scheme)
        attrs["alpha_b"] = alpha_b
        hd_b, hr_b = halves(db), halves(rb)
        k_b = _spectrum_of(ctx, [(*hd_b, 1.0 / vpc), (*hr_b, -alpha_b / vpc)], n, L, prec, sampler, interlaced)
        fb = SurveyField(ctx, k_b, n, L, prec, sampler, interlaced, bc)
    if normalization_scheme == "particle":
        if correlation_mode == "auto":
            attrs["N0"] = (sa[6] + alpha_a ** 2 * sa[7]) / I_rand
        else:
            raise ValueError("Particle-normalization scheme will not work for cross-correlation!")
    else:
        # mixed-mesh normalisation on the x-space meshes D/dV and R/dV
        def xmesh(h):
            return _xspace(ctx, _spectrum_of(ctx, [(*h, 1.0 / vpc)], n, L, prec, sampler, interlaced), n, prec)

        if correlation_mode == "auto":
            I_mesh = alpha_a * ctx.triple_sum(prec, xmesh(hd_a), xmesh(hr_a)) * vpc
        else:
            I_mesh = 0.5 * vpc * (attrs["alpha_b"] * ctx.triple_sum(prec, xmesh(hd_a), xmesh(hr_b))
                                  + alpha_a * ctx.triple_sum(prec, xmesh(hd_b), xmesh(hr_a)))
        I_mesh = np.float64(np.real(I_mesh))
        attrs["I_mesh"] = I_mesh
        attrs["N0"] = (sa[6] + alpha_a ** 2 * sa[7]) / I_mesh if correlation_mode == "auto" else 0.0
    return attrs, fa, fb


# ---------------------------------------------------------------- pk estimator
def _G_ell(ctx, field, ell, comp):
    """``get_G_ell`` (``clustering_estimator.py:3104-3134``):
    G_l = 4 pi V/(2l+1) comp(k) sum_{m=0..l} (2 - delta_m0) Y_lm(khat)
    r2c(F Y*_lm(xhat)) with non-Racah Y_lm."""
    n, L, prec = field.n, field.boxsize, field.prec
    V = L ** 3
    G = ctx.full(n, prec)
    tmp = ctx.full(n, prec)
    X = ctx.full(n, prec)
    for m in range(ell + 1):
        wx = A.ylm_poly(ell, 0, False) if m == 0 else _conj_ylm_poly(ell, m, False)
        ctx.xweight(field.x, tmp, n, L, prec, field.boxcenter, wx)
        ctx.fft_c2c(tmp, X, n, prec, -1)
        coef = (1.0 if m == 0 else 2.0) * 4 * np.pi * V / (2 * ell + 1) / float(n) ** 3
        ctx.kcombine(X, G, n, L, prec, comp=comp, sampler=field.sampler, poly=A.ylm_poly(ell, m, False), coef=coef,
                     accumulate=m > 0)
    return G


def high_order_strategy(poles, high_order_mode, correlation_mode):
    """``get_high_order_strategy`` (``clustering_estimator.py:243-331``):
    which exact and fast (P4b, P6b, P8b) multipoles to form."""
    supported = ([0, 2, 4], [0, 2, 4, 6], [0, 2, 4, 6, 8])
    if high_order_mode == "default":
        logging.info("[high-order mode] using exact multipoles only.")
        return None
    if correlation_mode != "auto" or poles not in supported:
        logging.info("[high-order mode] disabled (auto-correlation and poles "
                     f"{list(supported)} only).")
        return None
    high = {p for p in poles if p in (4, 6, 8)}
    fast, exact, internal = set(), set(), set()
    if high_order_mode == "fast":
        fast = {(0, 2, 4): {4}, (0, 2, 4, 6): {6}, (0, 2, 4, 6, 8): {6, 8}}.get(tuple(poles), set())
        exact = {(0, 2, 4): set(), (0, 2, 4, 6): {4}, (0, 2, 4, 6, 8): {4}}.get(tuple(poles), set())
    elif high_order_mode == "compare":
        fast, exact = set(high), set(high)
    if any(p in fast for p in (6, 8)):
        internal.add(4)  # P6b and P8b use the exact P4
    return {"mode": high_order_mode, "compute_fast_poles": fast, "compute_exact_poles": exact | internal,
            "visible_exact_poles": exact, "need_f2_cache": any(p in fast for p in (4, 6, 8)),
            "need_f4_cache": any(p in fast for p in (6, 8))}


def calculate_power_spectrum_survey(stat_attrs, rfield_a, rfield_b, correlation_mode, comm=None, **kwargs):
    """``clustering_estimator.py:205-563``, including the ``high_order_mode``
    fast / compare estimators P4b, P6b, P8b built from
    F_l = (G_l(k) + conj G_l(-k)) / 2 (:429-444, :487-550)."""
    ctx = get_context()
    poles = stat_attrs["poles"]
    strat = high_order_strategy(poles, stat_attrs.get("high_order_mode", "default"), correlation_mode)
    k_min, k_max, k_bins = stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"]
    interlaced = stat_attrs["interlaced"]
    N0 = stat_attrs["N0"]
    if stat_attrs["normalization_scheme"] == "particle":
        I_norm = stat_attrs["I_rand"]
    else:
        I_norm = stat_attrs["I_mesh"]
    k_edge = np.linspace(k_min, k_max, k_bins + 1)
    n, L, prec = rfield_a.n, rfield_a.boxsize, rfield_a.prec
    V = L ** 3
    comp = 1 if interlaced else 2
    ca = ctx.full(n, prec)
    ctx.kcombine(rfield_a.k, ca, n, L, prec, comp=comp, sampler=rfield_a.sampler, coef=V)
    if correlation_mode == "auto":
        cb = ca
    else:
        cb = ctx.full(n, prec)
        ctx.kcombine(rfield_b.k, cb, n, L, prec, comp=comp, sampler=rfield_b.sampler, coef=V)
    count, ksum, _, _ = ctx.bin(None, None, n, L, prec, k_edge)
    k_num = count
    with np.errstate(invalid="ignore", divide="ignore"):
        results = {"k_eff": ksum / k_num, "k_num": k_num, "I_norm": I_norm}
    F2, F4 = None, None
    for pole in poles:
        if pole == 0:
            if correlation_mode == "auto":
                s = ctx.full_bin(ca, n, L, prec, k_edge, B=ca, opb=2).real
            else:
                s = ctx.full_bin(ca, n, L, prec, k_edge, B=cb, opb=2)
        else:
            store = True
            if pole in (4, 6, 8) and strat is not None:
                store = pole in strat["compute_exact_poles"]
            cache2 = pole == 2 and strat is not None and strat["need_f2_cache"]
            cache4 = pole == 4 and strat is not None and strat["need_f4_cache"]
            if not store and not cache2 and not cache4:
                continue
            G = _G_ell(ctx, rfield_b, pole, comp)
            if cache2:
                F2 = ctx.reflect_conj(G, n, prec)
            if cache4:
                F4 = ctx.reflect_conj(G, n, prec)
            if not store:
                del G
                continue
            s = ctx.full_bin(ca, n, L, prec, k_edge, B=G, opb=2)
            del G
        with np.errstate(invalid="ignore", divide="ignore"):
            res = s / k_num
        res = res * ((2 * pole + 1) / I_norm)
        if pole == 0:
            res = res - N0
        else:
            res = np.imag(res) if pole & 1 else np.real(res)
        results[f"P{pole}"] = res
    if strat is not None and strat["compute_fast_poles"]:
        fast = strat["compute_fast_poles"]
        with np.errstate(invalid="ignore", divide="ignore"):
            if 4 in fast:
                P22 = ctx.full_bin(F2, n, L, prec, k_edge, B=F2, opb=2) / k_num
                P22 = P22 * (9 / I_norm) - N0 * 9 / 5
                results["P4b"] = np.real(35 / 18 * P22 - results["P2"] - 7 / 2 * results["P0"])
            if 6 in fast or 8 in fast:
                s42 = ctx.full_bin(F4, n, L, prec, k_edge, B=F2, opb=2)
                P42 = s42 / k_num * (13 / I_norm)
                if 6 in fast:
                    results["P6b"] = np.real(11 / 5 * P42 - 52 / 63 * results["P4"] - 286 / 175 * results["P2"])
                if 8 in fast:
                    P44 = ctx.full_bin(F4, n, L, prec, k_edge, B=F4, opb=2) / k_num * (17 / I_norm)
                    P44 = P44 - N0 * 17 / 9
                    results["P8b"] = np.real(1287 / 490 * P44 - 374 / 245 * P42 - 3553 / 15435 * results["P4"]
                                             - 1326 / 8575 * results["P2"] - 2431 / 490 * results["P0"])
        if strat["mode"] == "fast":
            for pole in fast:
                results.pop(f"P{pole}", None)
    return results


# ---------------------------------------------------------------- bk meshes
def get_mesh_bk_survey(catalogs, correlation_mode, tracer_type, angu_config, nmesh, geometry, column_names, boxsize,
                       sampler, interlaced, z_range, comp_weight_plan, para_cosmo, comm=None,
                       normalization_scheme="particle", alpha_scheme="pypower", precision="fp64", arrays=None):
    """``mesh_generator.py:541-1004``: returns (mesh_attrs, field_a, field_b,
    field_c); each field keeps its (device) catalogs for the N-field."""
    assert geometry == "survey-like", "get_mesh_bk_survey only supports survey-like geometry."
    assert correlation_mode in ["auto", "cross"], "correlation_mode must be 'auto' or 'cross'."
    assert tracer_type in ["aaa", "aab", "abb", "abc"], "tracer_type must be one of 'aaa', 'aab', 'abb', 'abc'."
    assert normalization_scheme in ["particle", "mesh", "mixed-mesh"], \
        "normalization_scheme must be one of 'particle', 'mesh', 'mixed-mesh'."
    if correlation_mode == "auto":
        assert tracer_type == "aaa", "For auto-correlation, tracer_type must be 'aaa'."
    ctx = get_context()
    prec = PREC[precision]
    n = int(nmesh[0])
    L = float(boxsize[0])
    vpc = float(np.prod(np.asarray(boxsize, dtype=float)) / np.prod(np.asarray(nmesh, dtype=float)))
    attrs = {"normalization_scheme": normalization_scheme}
    rd = _Reader(ctx, catalogs, column_names, z_range, comp_weight_plan, para_cosmo, normalization_scheme, arrays)
    bc = rd.box_center()
    attrs["boxcenter"] = bc
    labels = ["a"] if correlation_mode == "auto" else (["a", "b", "c"] if tracer_type == "abc" else ["a", "b"])
    fields, halves = {}, {}
    for lab in labels:
        r, d = rd.get(f"randoms_{lab}"), rd.get(f"data_{lab}")
        sums = compute_particle_sums_bk(ctx, d, r, correlation_mode, angu_config)
        alpha = _alpha(sums, alpha_scheme)
        attrs[f"alpha_{lab}"] = alpha
        attrs[f"I_33_{lab}"] = sums[4] * alpha
        attrs[f"S_LM_{lab}"] = sums[5] - alpha ** 3 * sums[6]
        hd = _paint_halves(ctx, d, ctx.pair_weights(d.weight, d.weight_fkp), n, L, sampler, interlaced, prec, bc)
        hr = _paint_halves(ctx, r, ctx.pair_weights(r.weight, r.weight_fkp), n, L, sampler, interlaced, prec, bc)
        k = _spectrum_of(ctx, [(*hd, 1.0 / vpc), (*hr, -alpha / vpc)], n, L, prec, sampler, interlaced)
        fields[lab] = SurveyField(ctx, k, n, L, prec, sampler, interlaced, bc, catalogs={"data": d, "randoms": r})
        halves[lab] = (hd, hr)

    def xmesh(lab, which, scale):
        h = halves[lab][0 if which == "d" else 1]
        return _xspace(ctx, _spectrum_of(ctx, [(*h, scale)], n, L, prec, sampler, interlaced), n, prec)

    if normalization_scheme == "particle":
        assert correlation_mode == "auto", "Particle-normalization scheme is only applicable for auto-correlation."
        attrs["I_norm"] = attrs["I_33_a"]
    elif normalization_scheme == "mesh":
        assert correlation_mode == "auto", "Mesh-normalization scheme is only applicable for auto-correlation."
        ra_x = xmesh("a", "r", attrs["alpha_a"] / vpc)
        attrs["I_norm"] = np.complex128(vpc * ctx.triple_sum(prec, ra_x, ra_x, ra_x))
    else:
        if correlation_mode == "auto":
            d, r = rd.get("data_a"), rd.get("randoms_a")
            nr = len(r)
            idx = torch.arange(nr, device=ctx.device)
            r_odd, r_even = r.take(idx[idx % 2 == 1]), r.take(idx[idx % 2 == 0])
            if alpha_scheme == "pypower":
                sd = ctx.particle_sums(d.pos, d.weight, d.weight_fkp)[0]
                so = ctx.particle_sums(r_odd.pos, r_odd.weight, r_odd.weight_fkp)[0]
                se = ctx.particle_sums(r_even.pos, r_even.weight, r_even.weight_fkp)[0]
            elif alpha_scheme == "nbodykit":
                sd = ctx.particle_sums(d.pos, d.weight, d.weight_fkp)[1]
                so = ctx.particle_sums(r_odd.pos, r_odd.weight, r_odd.weight_fkp)[1]
                se = ctx.particle_sums(r_even.pos, r_even.weight, r_even.weight_fkp)[1]
            else:
                raise ValueError("alpha_scheme must be either 'pypower' or 'nbodykit'.")
            a_odd, a_even = sd / so, sd / se

            def xcat(cat, scale):
                h = _paint_halves(ctx, cat, ctx.pair_weights(cat.weight, cat.weight_fkp), n, L, sampler, interlaced,
                                  prec, bc)
                return _xspace(ctx, _spectrum_of(ctx, [(*h, scale)], n, L, prec, sampler, interlaced), n, prec)

            attrs["I_norm"] = np.complex128(vpc * ctx.triple_sum(prec, xmesh("a", "d", 1.0 / vpc),
                                                                 xcat(r_odd, a_odd / vpc), xcat(r_even, a_even / vpc)))
        else:
            al = {lab: attrs[f"alpha_{lab}"] for lab in labels}
            if tracer_type == "abc":
                t = ctx.triple_sum(prec, xmesh("a", "r", al["a"] / vpc), xmesh("b", "r", al["b"] / vpc),
                                   xmesh("c", "r", al["c"] / vpc))
            elif tracer_type == "aab":
                t = ctx.triple_sum(prec, xmesh("a", "d", 1.0 / vpc), xmesh("a", "r", al["a"] / vpc),
                                   xmesh("b", "r", al["b"] / vpc))
            else:
                t = ctx.triple_sum(prec, xmesh("a", "r", al["a"] / vpc), xmesh("b", "r", al["b"] / vpc),
                                   xmesh("b", "d", 1.0 / vpc))
            attrs["I_norm"] = np.complex128(vpc * t)
    del halves
    return attrs, fields["a"], fields.get("b"), fields.get("c")


def get_N_field(ctx, field, alpha):
    """``mesh_generator.py:1007-1128``: N = D_{w^2}/dV + alpha^2 R_{w^2}/dV,
    painted with the squared total weights of the field's own catalogs."""
    n, L, prec = field.n, field.boxsize, field.prec
    vpc = L ** 3 / float(n) ** 3
    d, r = field.catalogs["data"], field.catalogs["randoms"]
    hd = _paint_halves(ctx, d, ctx.pair_weights(d.weight, d.weight_fkp, 2), n, L, field.sampler, field.interlaced,
                       prec, field.boxcenter)
    hr = _paint_halves(ctx, r, ctx.pair_weights(r.weight, r.weight_fkp, 2), n, L, field.sampler, field.interlaced,
                       prec, field.boxcenter)
    k = _spectrum_of(ctx, [(*hd, 1.0 / vpc), (*hr, alpha ** 2 / vpc)], n, L, prec, field.sampler, field.interlaced)
    return SurveyField(ctx, k, n, L, prec, field.sampler, field.interlaced, field.boxcenter)


# ---------------------------------------------------------------- bk estimator
def _complex_on_grid(ctx, spec, n, prec, m, nmax):
    """x-space (re, im) real meshes, flattened, of a full spectrum, on the
    contraction grid m (spectrum cut at |n_d| <= nmax when m < n)."""
    if m != n:
        spec = ctx.resample_full(spec, n, m, prec, nmax)
    x = ctx.full(m, prec)
    ctx.fft_c2c(spec, x, m, prec, +1)
    re, im = ctx.split_complex(x, prec)
    return [(re.view(-1), 1.0 + 0j), (im.view(-1), 1j)]


def _half_of(full, n):
    return full[:, :, : n // 2 + 1].contiguous()


def calculate_bk_sugi_survey(rfield_a, rfield_b, rfield_c, correlation_mode, stat_attrs, comm=None, **kwargs):
    """Survey Sugiyama bispectrum (``clustering_estimator.py:1558-3101``)."""
    ctx = get_context()
    dvm = stat_attrs.get("data_vector_mode", "diagonal")
    sn_mode = stat_attrs.get("shotnoise_mode", "ana")
    block_size = stat_attrs.get("block_size", 1 if dvm == "diagonal" else "full")
    ell_1, ell_2, L = stat_attrs["angu_config"]
    k_min, k_max, k_bins = stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"]
    tracer_type = stat_attrs["tracer_type"]
    I_norm = stat_attrs["I_norm"]
    n, Lbox, prec = rfield_a.n, rfield_a.boxsize, rfield_a.prec
    sampler = rfield_a.sampler
    V = Lbox ** 3
    vpc = V / float(n) ** 3
    contraction = _gemm_mode(stat_attrs.get("contraction", "gemm_fp64" if prec == 0 else "bf16x3"), prec)
    shell_method = stat_attrs.get("shell_method", "synth")
    zsynth = stat_attrs.get("zsynth", "fused")
    N_ells = (2 * ell_1 + 1) * (2 * ell_2 + 1) * (2 * L + 1)
    H_ells = A.wigner_3j(ell_1, ell_2, L, 0, 0, 0)
    mcfg, tjv = A.get_magnetic_configs_survey(ell_1, ell_2, L)
    k_edge = np.linspace(k_min, k_max, k_bins + 1)
    if dvm == "full":
        if block_size == "full":
            pass
        elif not (isinstance(block_size, int) and 1 <= block_size <= k_bins):
            raise ValueError("For full mode, block_size must be 1, 'full', or an integer in [1, k_bins].")
    elif dvm != "diagonal":
        raise ValueError("data_vector_mode must be either 'diagonal' or 'full'.")
    p = {"ngp": 1, "cic": 2, "tsc": 3, "pcs": 4}[sampler]

    def spec_V(field):
        out = ctx.full(n, prec)
        ctx.kcombine(field.k, out, n, Lbox, prec, comp=1, sampler=sampler, coef=V)
        return out

    ca = spec_V(rfield_a)
    cb = ca if correlation_mode == "auto" else spec_V(rfield_b)
    count, ksum, _, _ = ctx.bin(None, None, n, Lbox, prec, k_edge)
    k_num = count
    with np.errstate(invalid="ignore", divide="ignore"):
        k_eff = ksum / k_num
    if correlation_mode == "auto":
        f3 = rfield_a
    else:
        f3 = rfield_c if tracer_type == "abc" else rfield_b

    m, nmax = contraction_grid(n, Lbox, k_max, contraction, stat_attrs.get("shell_grid", "reduced"))
    grid_factor = (n / m) ** 3
    kinfo = (m, Lbox, prec, k_edge, k_bins)
    builder = ShellBuilder(ctx, m, Lbox, prec, k_edge, method=shell_method)

    def half_on_grid(full):
        h = _half_of(full, n)
        return ctx.resample(h, n, m, prec, nmax) if m != n else h

    da_m = half_on_grid(ca)
    db_m = da_m if tracer_type in ("aaa", "aab") else half_on_grid(cb)

    tmp = ctx.full(n, prec)
    X = ctx.full(n, prec)

    def G_LM(M):
        """c2r(comp(r2c(F3 y_LM(xhat)))) (:1706-1717), as complex parts on grid m."""
        ctx.xweight(f3.x, tmp, n, Lbox, prec, f3.boxcenter, A.ylm_poly(L, M, True))
        ctx.fft_c2c(tmp, X, n, prec, -1)
        ctx.kcombine(X, X, n, Lbox, prec, comp=1, sampler=sampler, coef=1.0 / float(n) ** 3)
        return _complex_on_grid(ctx, X, n, prec, m, nmax)

    def run_configs(gfun, src1, src2, sym_tracer, conj_w, post):
        """Configuration loop with the reference's swap reuse and the
        (0,0,0)-once / 2 Re accumulation (:1773-1913, :1984-2052)."""
        total = np.zeros((k_bins, k_bins) if dvm == "full" else k_bins, dtype=np.complex128)
        cfg_idx = {c: i for i, c in enumerate(mcfg)}
        processed = set()
        cur_M, Gs = None, None
        same_ell = ell_1 == ell_2
        for ss, (m1, m2, M) in enumerate(mcfg):
            if (m1, m2, M) in processed:
                continue
            if M != cur_M:
                Gs, cur_M = gfun(M), M
            sym_eq = sym_tracer and same_ell and m1 == m2
            sym_cj = sym_tracer and same_ell and m1 == -m2
            sub = grid_factor * _config_matrix(ctx, builder, src1, src2, Gs, kinfo, ell_1, m1, ell_2, m2, sym_eq,
                                               sym_cj, contraction, block_size, dvm, conj_w, zsynth)
            sub = post(tjv[ss] * sub)
            swap = (m2, m1, M)
            has_swap = sym_tracer and same_ell and swap in cfg_idx and swap != (m1, m2, M)
            add_swap = has_swap and swap not in processed
            add = sub
            if add_swap:
                add = 2.0 * sub if dvm == "diagonal" else sub + sub.T
            total += add if (m1, m2, M) == (0, 0, 0) else 2 * np.real(add)
            if add_swap:
                processed.add(swap)
            processed.add((m1, m2, M))
        return total

    same_tracer_12 = correlation_mode == "auto" or tracer_type == "aab"
    total_res = run_configs(G_LM, da_m, db_m, same_tracer_12, False, lambda s: s)

    # ------------------------------------------------------------ shot noise
    k_center = 0.5 * (k_edge[1:] + k_edge[:-1])
    k3_edge = np.linspace(2 * k_min, 2 * k_max, k_bins * 2 + 1)[:-1]
    k3_center = 0.5 * (k3_edge[1:] + k3_edge[:-1])
    k3_bins = len(k3_center)
    k_num_k3, ksum3, _, _ = ctx.bin(None, None, n, Lbox, prec, k3_edge)
    with np.errstate(invalid="ignore", divide="ignore"):
        k3_eff = ksum3 / k_num_k3
    use_sbar = correlation_mode == "auto" and tracer_type == "aaa"
    rf_shot, N_shot, S_LM_shot = None, None, None
    if tracer_type != "abc":
        flag = "b" if tracer_type == "abb" else "a"
        src_field = {"a": rfield_a, "b": rfield_b}[flag]
        N_shot = get_N_field(ctx, src_field, stat_attrs[f"alpha_{flag}"])
        rf_shot = rfield_b if tracer_type == "aab" else rfield_a
        S_LM_shot = stat_attrs.get(f"S_LM_{flag}", None)

    def s_lm_array(S):
        if S is None:
            return np.zeros(L + 1, dtype="complex128")
        if np.isscalar(S):
            a = np.zeros(L + 1, dtype="complex128")
            a[0] = S
            return a
        return np.asarray(S, dtype="complex128")

    def weighted_V(field_x, poly):
        """V comp(r2c(field * poly(xhat))) as a full spectrum (in X)."""
        ctx.xweight(field_x, tmp, n, Lbox, prec, rfield_a.boxcenter, poly)
        ctx.fft_c2c(tmp, X, n, prec, -1)
        ctx.kcombine(X, X, n, Lbox, prec, comp=1, sampler=sampler, coef=V / float(n) ** 3)
        return X

    def S1_like():
        """``clustering_estimator.py:2231-2320``."""
        out = np.zeros(k_bins, dtype="complex128")
        if not (ell_1 == L and ell_2 == 0) or rf_shot is None or N_shot is None:
            return out
        cd = spec_V(rf_shot)
        S_arr = s_lm_array(S_LM_shot) if use_sbar else None
        acc = ctx.full(n, prec)
        for M in range(L + 1):
            mw = 0.5 if M == 0 else 1.0
            cn = weighted_V(N_shot.x, A.ylm_poly(L, M, True))
            ctx.kcombine(cd, acc, n, Lbox, prec, B=cn, opb=2, sampler=sampler,
                         shot=S_arr[M] if use_sbar else None, poly=A.ylm_poly(L, M, True), coef=mw,
                         accumulate=M > 0)
        s = 2.0 * np.real(ctx.full_bin(acc, n, Lbox, prec, k_edge))
        ok = k_num > 0
        out[ok] = (2 * L + 1) * s[ok] / k_num[ok]
        return out

    def Q_L_like():
        """``clustering_estimator.py:2322-2401``."""
        if rf_shot is None or N_shot is None:
            return np.zeros(k3_bins, dtype="complex128")
        S_arr = s_lm_array(S_LM_shot) if use_sbar else None
        cn0 = spec_V(N_shot)
        acc = ctx.full(n, prec)
        for M in range(L + 1):
            mw = 0.5 if M == 0 else 1.0
            cdl = weighted_V(rf_shot.x, _conj_ylm_poly(L, M, True))
            ctx.kcombine(cdl, acc, n, Lbox, prec, B=cn0, opb=2, sampler=sampler,
                         shot=S_arr[M] if use_sbar else None, poly=A.ylm_poly(L, M, True), coef=mw,
                         accumulate=M > 0)
        return (2.0 * np.real(ctx.full_bin(acc, n, Lbox, prec, k3_edge))).astype("complex128")

    def S3_ana(Q):
        """``clustering_estimator.py:2403-2491``."""
        wkn = np.zeros(k3_bins, dtype="float64")
        ok = k_num_k3 > 0
        wkn[ok] = k_num_k3[ok] / k3_center[ok]
        # every (i, j) of the data vector at once
        if dvm == "diagonal":
            q = A.q_ells_pairs(k_eff[:k_bins], k_eff[:k_bins], ell_1, ell_2, L, k3_eff)
            numer = np.sum(q * Q / k3_center, axis=-1)
            denom = np.array([np.sum(wkn[: 2 * i + 1]) for i in range(k_bins)])
        else:
            q = A.q_ells_pairs(k_eff[:k_bins, None], k_eff[None, :k_bins], ell_1, ell_2, L, k3_eff)
            valid = A.valid_k3_bins_pairs(k_center[:, None], k_center[None, :], k_min, k_max, k_bins)
            numer = np.sum(q * Q / k3_center, axis=-1)
            denom = np.sum(wkn * valid, axis=-1)
        s3 = np.where(denom > 0, numer / np.where(denom > 0, denom, 1.0), 0.0).astype("complex128")
        return s3 * (H_ells * N_ells)

    def S3_fft():
        """``clustering_estimator.py:2510-2781``: shells of conj(Y) * mask
        contracted with c2r(V comp r2c(F y*_LM) conj(V comp r2c(N)) - S C_bk)."""
        S_arr = s_lm_array(S_LM_shot) if use_sbar else None

        def s_lm_value(M):
            if S_arr is None:
                return None
            if M >= 0:
                return S_arr[M]
            return ((-1) ** M) * np.conj(S_arr[-M])

        cn0 = spec_V(N_shot)
        src = ctx.full(n, prec)

        def source(M):
            cdl = weighted_V(rf_shot.x, _conj_ylm_poly(L, M, True))
            ctx.kcombine(cdl, src, n, Lbox, prec, B=cn0, opb=2, sampler=sampler, shot=s_lm_value(M))
            return _complex_on_grid(ctx, src, n, prec, m, nmax)

        kk = k_num.astype(np.float64)
        with np.errstate(invalid="ignore", divide="ignore"):
            norm = np.outer(kk, kk) if dvm == "full" else kk ** 2

        def post(sub):
            with np.errstate(invalid="ignore", divide="ignore"):
                return np.where(norm > 0, sub / np.where(norm > 0, norm, 1.0), 0.0)

        tot = run_configs(source, None, None, True, True, post)
        return tot * (H_ells * N_ells) / float(n) ** 3

    shape = (k_bins,) if dvm == "diagonal" else (k_bins, k_bins)
    SN0 = 0.0 + 0.0j
    SN1 = np.zeros(shape, dtype="complex128")
    SN2 = np.zeros(shape, dtype="complex128")
    SN3 = np.zeros(shape, dtype="complex128")
    SN3_fft = None
    if tracer_type != "abc":
        if tracer_type == "aaa" and (ell_1, ell_2, L) == (0, 0, 0):
            S0 = stat_attrs.get("S_LM_a", 0.0)
            SN0 = S0[0] if isinstance(S0, np.ndarray) else S0
        if tracer_type in ["aaa", "abb"] and ell_1 == L and ell_2 == 0:
            s1 = S1_like()
        else:
            s1 = np.zeros(k_bins, dtype="complex128")
        s2 = s1.copy() if (tracer_type == "aaa" and (ell_1, ell_2, L) == (0, 0, 0)) \
            else np.zeros(k_bins, dtype="complex128")
        if dvm == "diagonal":
            SN1, SN2 = s1, s2
        else:
            SN1[:] = s1[:, None]
            SN2[:] = s2[None, :]
        if tracer_type != "abb":
            if sn_mode == "both":
                SN3 = S3_ana(Q_L_like())
                SN3_fft = S3_fft()
            elif sn_mode == "ana":
                SN3 = S3_ana(Q_L_like())
            elif sn_mode == "fft":
                SN3 = S3_fft()
            else:
                raise ValueError("shotnoise_mode must be either 'ana', 'fft', or 'both'.")
        elif sn_mode == "both":
            SN3_fft = np.zeros(shape, dtype="complex128")
    del tmp, X

    with np.errstate(invalid="ignore", divide="ignore"):
        total_res = total_res * (N_ells * H_ells / I_norm)
        if dvm == "full":
            total_res = total_res * (1.0 / np.outer(k_num, k_num))
        else:
            total_res = total_res * (1.0 / (k_num ** 2))
    total_res = total_res * vpc
    total_shot_noise = (SN0 + SN1 + SN2 + SN3) / I_norm
    sn_terms = {"SN0": SN0 / I_norm, "SN1": SN1 / I_norm, "SN2": SN2 / I_norm, "SN3": SN3 / I_norm}
    if sn_mode == "both":
        sn_terms["SN3_FFT"] = SN3_fft / I_norm
    return {"B_sugi": total_res - total_shot_noise, "Bk_raw": total_res, "SN_terms": sn_terms,
            "Shot_noise": total_shot_noise, "I_norm": I_norm, "k_eff": k_eff, "nmodes": k_num}
