"""Pins the oracle's survey restatement: comoving-distance closed forms,
catalog processing, particle sums on the line of sight, and the
distant-observer limits in which the survey estimators reduce to the box
estimators (themselves pinned by closed forms and brute-force triangle sums
in test_oracle_box.py)."""
import numpy as np
import pytest

from oracle.cnpc_ref import angular as OA
from oracle.cnpc_ref import box as B
from oracle.cnpc_ref import catalog as OC
from oracle.cnpc_ref import conventions as C
from oracle.cnpc_ref import survey as S

from helpers import COSMO, NO_COMP, oracle_survey_cats, survey_rows

CH = 299792.458 / 100.0


def test_comoving_distance_closed_forms():
    z = np.array([0.0, 0.1, 0.5, 1.0, 3.0])
    # Einstein-de Sitter: chi = 2 c/H0 (1 - 1/sqrt(1+z))
    np.testing.assert_allclose(OC.comoving_distance(z, 1.0), 2 * CH * (1 - 1 / np.sqrt(1 + z)), rtol=1e-13,
                               atol=1e-12)
    # empty limit (Omega_m -> 0): chi = c z / H0
    np.testing.assert_allclose(OC.comoving_distance(z, 1e-14), CH * z, rtol=1e-12, atol=1e-12)


def test_survey_catalog_processing():
    d, _ = survey_rows(500, 500, 11, extra_cols=True)
    cat = OC.survey_catalog(d, (0.46, 0.49), {"scheme": "boss", "name_alias": None}, COSMO, "data")
    keep = (d["Z"] > 0.46) & (d["Z"] < 0.49)
    assert len(cat) == keep.sum() and 0 < len(cat) < len(d)
    sub = d[keep]
    np.testing.assert_allclose(cat["WEIGHT"], sub["WEIGHT_SYSTOT"] * (sub["WEIGHT_NOZ"] + sub["WEIGHT_CP"] - 1.0))
    r = np.linalg.norm(cat["Position"], axis=1)
    np.testing.assert_allclose(r, OC.comoving_distance(sub["Z"], COSMO["Omega0"]), rtol=1e-14)
    np.testing.assert_allclose(np.rad2deg(np.arcsin(cat["Position"][:, 2] / r)), sub["DEC"], rtol=1e-12)
    np.testing.assert_allclose(np.rad2deg(np.arctan2(cat["Position"][:, 1], cat["Position"][:, 0])), sub["RA"],
                               rtol=1e-12)
    ecat = OC.survey_catalog(d, None, {"scheme": "eboss", "name_alias": None}, COSMO, "data")
    np.testing.assert_allclose(ecat["WEIGHT"], d["WEIGHT_SYSTOT"] * d["WEIGHT_NOZ"] * d["WEIGHT_CP"])
    rcat = OC.survey_catalog(d, None, {"scheme": "boss", "name_alias": None}, COSMO, "randoms")
    assert np.all(rcat["WEIGHT"] == 1.0)
    acat = OC.survey_catalog(d, None, {"scheme": None, "name_alias": "WEIGHT_SYSTOT"}, COSMO, "data")
    np.testing.assert_array_equal(acat["WEIGHT"], d["WEIGHT_SYSTOT"])
    with pytest.raises(ValueError, match="Cannot specify both"):
        OC.survey_catalog(d, None, {"scheme": "boss", "name_alias": "X"}, COSMO, "data")
    no_nz = np.zeros(len(d), dtype=[("RA", "f8"), ("DEC", "f8"), ("Z", "f8"), ("WEIGHT_FKP", "f8")])
    for nm in no_nz.dtype.names:
        no_nz[nm] = d[nm]
    with pytest.raises(ValueError, match="NZ column is missing"):
        OC.survey_catalog(no_nz, None, NO_COMP, COSMO, "data", "particle")
    assert np.all(OC.survey_catalog(no_nz, None, NO_COMP, COSMO, "data", "mixed-mesh")["NZ"] == 1.0)


def _axis_cat(w, wf, nz, zpos):
    out = np.zeros(len(w), dtype=[("Position", "f8", (3,)), ("WEIGHT", "f8"), ("WEIGHT_FKP", "f8"), ("NZ", "f8")])
    out["Position"][:, 2] = zpos
    out["WEIGHT"], out["WEIGHT_FKP"], out["NZ"] = w, wf, nz
    return out


def test_particle_sums_on_the_axis():
    """Objects on +z: y_LM(zhat) = delta_M0 (Racah), so S_LM[0] = sum w^3 and
    S_LM[M>0] = 0; the other sums are their definitions."""
    rng = np.random.default_rng(3)
    w, wf, nz = 0.5 + rng.random(50), 0.1 + rng.random(50), 1e-4 * rng.random(50)
    d = _axis_cat(w, wf, nz, 100 + rng.random(50))
    r = _axis_cat(w[:20], wf[:20], nz[:20], 200 + rng.random(20))
    for L in (0, 2, 4):
        out = S.compute_particle_sums_bk(d, r, "auto", [L, 0, L])
        np.testing.assert_allclose(out[5][0], np.sum((w * wf) ** 3), rtol=1e-13)
        np.testing.assert_allclose(np.abs(out[5][1:]), 0.0, atol=1e-12)
        np.testing.assert_allclose(out[4], np.sum(nz[:20] ** 2 * w[:20] ** 2 * wf[:20] ** 3), rtol=1e-13)
    pk = S.compute_particle_sums_pk(d, r, "auto")
    np.testing.assert_allclose(pk[:4], [np.sum(w * wf), np.sum(w[:20] * wf[:20]), np.sum(w), np.sum(w[:20])],
                               rtol=1e-13)
    np.testing.assert_allclose(pk[4], np.sum(w * nz * wf ** 2), rtol=1e-13)
    np.testing.assert_allclose(pk[6], np.sum((w * wf) ** 2), rtol=1e-13)
    # a general direction: |S_LM| against an explicit scipy Ylm sum
    from scipy.special import sph_harm_y

    p = rng.normal(size=(30, 3))
    dd = _axis_cat(w[:30], wf[:30], nz[:30], 0.0)
    dd["Position"] = p
    out = S.compute_particle_sums_bk(dd, r, "auto", [2, 0, 2])
    th = np.arccos(p[:, 2] / np.linalg.norm(p, axis=1))
    ph = np.arctan2(p[:, 1], p[:, 0])
    for M in range(3):
        y = sph_harm_y(2, M, th, ph) * np.sqrt(4 * np.pi / 5)
        np.testing.assert_allclose(out[5][M], np.sum((w[:30] * wf[:30]) ** 3 * np.conj(y)), rtol=1e-12)


def _far_catalogs(D, seed, n_d=1500, n_r=6000, L=200.0):
    """FKP catalogs in a cube around (0, 0, D): a distant observer."""
    rng = np.random.default_rng(seed)

    def cat(nobj):
        out = np.zeros(nobj, dtype=[("Position", "f8", (3,)), ("WEIGHT", "f8"), ("WEIGHT_FKP", "f8"), ("NZ", "f8")])
        out["Position"] = rng.random((nobj, 3)) * (0.9 * L) + np.array([0.0, 0.0, D]) - 0.45 * L
        out["WEIGHT"] = 0.5 + rng.random(nobj)
        out["WEIGHT_FKP"] = 0.2 + rng.random(nobj)
        out["NZ"] = 1e-3 * (1 + rng.random(nobj))
        return out

    return {"data_a": cat(n_d), "randoms_a": cat(n_r)}


@pytest.mark.parametrize("interlaced", [False, True])
def test_pk_survey_distant_observer(interlaced):
    """xhat -> zhat: sum_m (2 - d_m0) Y_lm(khat) Y*_lm(zhat) = (2l+1)/(4 pi)
    L_l(khat.zhat), so P_l = (2l+1)/I bin(|F_k|^2 L_l(mu)) / k_num with the
    plane-parallel Legendre weight."""
    n, L = 16, 200.0
    cats = _far_catalogs(1e12, 5, L=L)
    attrs, fa, fb = S.get_mesh_pk_survey(cats, "auto", n, L, "tsc", interlaced, "particle")
    sa = dict(attrs, poles=[0, 1, 2, 3, 4], boxsize=[L] * 3, nmesh=[n] * 3, k_min=0.02, k_max=0.2, k_bins=5,
              sampler="tsc", interlaced=interlaced, normalization_scheme="particle")
    res = S.calculate_power_spectrum_survey(sa, fa, fb, "auto")
    comp = OA.get_compensation(interlaced, "tsc")
    ca = comp(C.circular_axes(n), C.r2c(fa)) * L ** 3
    kgrid, knorm = OA.get_kgrid(C.k_axes(n, L))
    edges = np.linspace(0.02, 0.2, 6)
    k_num, _ = OA.get_kbin_count(5, edges, knorm)
    for ell in range(5):
        leg = OA.get_legendre(ell, 0.0, 0.0, 1.0)(*kgrid)
        s = OA.radial_binning(np.abs(ca) ** 2 * leg, 5, edges, knorm) / k_num * (2 * ell + 1) / attrs["I_rand"]
        if ell == 0:
            s = s - attrs["N0"]
            np.testing.assert_allclose(res["P0"], s, rtol=1e-8)
        elif ell % 2 == 0:
            np.testing.assert_allclose(res[f"P{ell}"], s, rtol=1e-7, atol=1e-8 * np.max(np.abs(s)))
        else:
            # odd poles vanish for a real |F|^2 L_l with L_l(-mu) = -L_l(mu)
            assert np.max(np.abs(res[f"P{ell}"])) < 1e-7 * np.max(np.abs(res["P0"]))


@pytest.mark.parametrize("angu,dvm", [([0, 0, 0], "full"), ([2, 0, 2], "diagonal"), ([1, 1, 0], "full"),
                                      ([2, 2, 2], "diagonal")])
def test_bk_survey_signal_distant_observer_equals_box(angu, dvm):
    """xhat -> zhat: y_LM(xhat) -> delta_M0, G_LM -> delta_M0 c2r(comp r2c F)
    and the survey configuration sum collapses onto the box one, so the survey
    signal is V^2 times the box signal of the field delta = F."""
    n, L = 16, 200.0
    cats = _far_catalogs(1e12, 9, L=L)
    attrs, fa, _, _ = S.get_mesh_bk_survey(cats, "auto", "aaa", angu, n, L, "cic", False, "particle")
    sa = dict(attrs, angu_config=angu, boxsize=[L] * 3, nmesh=[n] * 3, k_min=0.03, k_max=0.12, k_bins=3,
              sampler="cic", interlaced=False, tracer_type="aaa", data_vector_mode=dvm)
    sig = S.calculate_bk_sugi_survey(fa, None, None, "auto", sa, cats, return_signal_only=True)
    vpc = L ** 3 / n ** 3
    NZ = 1.0
    rbox = (fa + NZ) * vpc
    bsa = dict(sa, NZ_a=NZ, N_gal_a=1.0)
    box = B.calculate_bk_sugi_box(rbox, None, None, "auto", bsa, return_signal_only=True)
    np.testing.assert_allclose(sig, (L ** 3) ** 2 * box, rtol=1e-7, atol=1e-7 * np.max(np.abs(sig)))


@pytest.mark.parametrize("poles,mode", [([0, 2, 4], "fast"), ([0, 2, 4, 6], "compare"), ([0, 2, 4, 6, 8], "compare"),
                                        ([0, 2, 4, 6, 8], "fast")])
def test_pk_survey_fast_modes_distant_observer(poles, mode):
    """Distant observer: F_l = G_l and L_2^2, L_4 L_2, L_4^2 expand in
    Legendre polynomials with the coefficients behind P4b/P6b/P8b, so the fast
    estimators equal the exact P4/P6/P8 (and the output keys follow the
    reference's fast/compare tables)."""
    n, L = 16, 200.0
    cats = _far_catalogs(1e12, 7, L=L)
    attrs, fa, fb = S.get_mesh_pk_survey(cats, "auto", n, L, "cic", True, "particle")
    sa = dict(attrs, poles=poles, boxsize=[L] * 3, nmesh=[n] * 3, k_min=0.02, k_max=0.2, k_bins=5, sampler="cic",
              interlaced=True, normalization_scheme="particle")
    exact = S.calculate_power_spectrum_survey(dict(sa, poles=[0, 2, 4, 6, 8]), fa, fb, "auto")
    res = S.calculate_power_spectrum_survey(dict(sa, high_order_mode=mode), fa, fb, "auto")
    fast = {(0, 2, 4): [4], (0, 2, 4, 6): [6], (0, 2, 4, 6, 8): [6, 8]}[tuple(poles)] if mode == "fast" \
        else [p for p in poles if p >= 4]
    for p in fast:
        scale = np.max(np.abs(exact[f"P{p}"])) + np.max(np.abs(exact["P0"])) * 1e-3
        np.testing.assert_allclose(res[f"P{p}b"], exact[f"P{p}"], rtol=0, atol=1e-6 * scale)
    keys = {k for k in res if k.startswith("P")}
    want = {f"P{p}" for p in poles} | {f"P{p}b" for p in fast}
    if mode == "fast":
        want -= {f"P{p}" for p in fast}
    assert keys == want


def _lattice_randoms(n, L, center, sub=4):
    """A periodic lattice of spacing H/sub filling the box: every paint
    window sees the same weights around every mesh point, so the painted
    randoms are exactly uniform (no k != 0 modes)."""
    H = L / n
    g = (np.arange(sub * n) - sub * n // 2) * (H / sub)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    out = np.zeros(X.size, dtype=[("Position", "f8", (3,)), ("WEIGHT", "f8"), ("WEIGHT_FKP", "f8"), ("NZ", "f8")])
    out["Position"] = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()]) + np.asarray(center)
    out["WEIGHT"], out["WEIGHT_FKP"], out["NZ"] = 1.0, 1.0, 1e-3
    return out


def test_N_field_closed_forms():
    """N = D_{(W W_fkp)^2}/dV + alpha^2 R_{(W W_fkp)^2}/dV
    (``mesh_generator.py:1007-1128``): mass conservation with non-trivial
    weights pins the squares and alpha^2; one data object on a mesh point
    (CIC) over lattice randoms pins the cell and the 1/dV."""
    n, L = 16, 200.0
    vpc = L ** 3 / n ** 3
    cats = _far_catalogs(1e6, 21, n_d=400, n_r=1500, L=L)
    d, r = cats["data_a"], cats["randoms_a"]
    bc = S.survey_boxcenter(r)
    alpha = 0.37
    for interlaced in (False, True):
        N = S.get_N_field(cats, "a", alpha, n, L, "tsc", interlaced, bc)
        want = np.sum((d["WEIGHT"] * d["WEIGHT_FKP"]) ** 2) + alpha ** 2 * np.sum((r["WEIGHT"] * r["WEIGHT_FKP"]) ** 2)
        np.testing.assert_allclose(np.sum(np.real(N)) * vpc, want, rtol=1e-12)
    rl = _lattice_randoms(n, L, (0.0, 0.0, 1e6))
    bc = S.survey_boxcenter(rl)
    one = np.zeros(1, dtype=rl.dtype)
    one["Position"] = bc + np.array([3, -5, 2]) * (L / n)
    one["WEIGHT"], one["WEIGHT_FKP"] = 1.5, 0.8
    N = np.real(S.get_N_field({"data_a": one, "randoms_a": rl}, "a", alpha, n, L, "cic", False, bc))
    uniform = alpha ** 2 * len(rl) / L ** 3  # lattice randoms of unit weight
    exp = np.full((n, n, n), uniform)
    exp[3, -5, 2] += (1.5 * 0.8) ** 2 / vpc
    np.testing.assert_allclose(N, exp, rtol=1e-12, atol=1e-12 * np.max(exp))


@pytest.mark.parametrize("angu", [[0, 0, 0], [2, 0, 2]])
@pytest.mark.parametrize("sn", ["ana", "fft"])
def test_bk_survey_shot_noise_distant_observer_equals_box(angu, sn):
    """Distant observer, unit weights and lattice randoms (exactly uniform
    paint): F = D/dV - alpha R/dV and the N-field D/dV + alpha^2 R/dV share
    every k != 0 mode, y_LM(xhat) -> delta_M0 and S_LM -> delta_M0 S_0.  Then
    the survey S1-like term (``clustering_estimator.py:2231-2320``), Q_L
    (``:2322-2401``) with the analytic S3 (``:2403-2491``), and the FFT S3
    (``:2510-2781``) reduce to the box SN1, SN2 and S3 of the same field
    with SN0 = S_0 (``:1386-1468``, ``:1081-1384``; pinned in
    test_oracle_box.py).  k_min = 0.01 k_f, so the box's k3 bins (from 2 k_min)
    hold the same modes as the k bins the survey S1 uses."""
    n, L, D = 16, 200.0, 1e8
    kf = 2 * np.pi / L
    rng = np.random.default_rng(4)
    rl = _lattice_randoms(n, L, (0.0, 0.0, D))
    dat = np.zeros(1500, dtype=rl.dtype)
    dat["Position"] = rng.random((1500, 3)) * (0.9 * L) + np.array([0.0, 0.0, D]) - 0.45 * L
    dat["WEIGHT"], dat["WEIGHT_FKP"], dat["NZ"] = 1.0, 1.0, 1e-3
    cats = {"data_a": dat, "randoms_a": rl}
    kmin, kmax, K = 0.01 * kf, 3.61 * kf, 3
    attrs, fa, _, _ = S.get_mesh_bk_survey(cats, "auto", "aaa", angu, n, L, "tsc", False, "particle")
    sa = dict(attrs, angu_config=angu, boxsize=[L] * 3, nmesh=[n] * 3, k_min=kmin, k_max=kmax, k_bins=K,
              sampler="tsc", interlaced=False, tracer_type="aaa", data_vector_mode="full", shotnoise_mode=sn)
    sur = S.calculate_bk_sugi_survey(fa, None, None, "auto", sa, cats)
    S0 = complex(np.atleast_1d(attrs["S_LM_a"])[0])
    assert abs(S0.imag) < 1e-6 * abs(S0.real)
    NZ = 1.0
    rbox = (np.real(fa) + NZ) * L ** 3 / n ** 3
    bsa = dict(sa, NZ_a=NZ, N_gal_a=S0.real)
    box = B.calculate_bk_sugi_box(rbox, None, None, "auto", bsa)
    terms = ["SN0", "SN1", "SN2", "SN3"]
    scale = max(np.max(np.abs(sur["SN_terms"][t] * sur["I_norm"])) for t in terms)
    for t in terms:
        got = np.asarray(sur["SN_terms"][t] * sur["I_norm"])
        exp = np.asarray(box["SN_terms"][t] * box["I_norm"])
        np.testing.assert_allclose(got, exp, rtol=0, atol=2e-6 * scale, err_msg=t)
    # the terms are not all trivially zero
    assert np.max(np.abs(sur["SN_terms"]["SN3"] * sur["I_norm"])) > 1e-3 * scale
    if angu == [2, 0, 2]:
        assert np.max(np.abs(sur["SN_terms"]["SN1"] * sur["I_norm"])) > 1e-3 * scale


def test_planck18_flrw_pins():
    """The Planck18 branch's E(z) and distances: E(0) = 1 (flat), the photon
    density 2.4728e-5 h^-2 at 2.7255 K, the massless-neutrino ratio
    7/8 (4/11)^(4/3) N_eff, the non-relativistic limit of the 0.06 eV
    species (Omega_nu h^2 = m / 93.14 eV within 1.5%: the per-species
    N_eff / 3 enters as a factor), the fitting constant 0.3173 against
    180 zeta(3) / (7 pi^4), and the no-radiation, massless limits against the
    flat-LCDM and closed-form distances."""
    from scipy.special import zeta

    p = OC.P18
    h = p["H0"] / 100
    og = OC.omega_gamma(p["H0"], p["Tcmb0"])
    assert abs(og * h * h / 2.4728e-5 - 1) < 2e-4
    np.testing.assert_allclose(OC.efunc_flrw(0.0, **p), 1.0, rtol=0, atol=1e-15)
    r0 = OC.nu_density_ratio(0.0, p["Tcmb0"], p["Neff"], (0.0, 0.0, 0.0))
    np.testing.assert_allclose(r0, 7 / 8 * (4 / 11) ** (4 / 3) * p["Neff"], rtol=1e-12)
    assert abs(0.3173 / (180 * zeta(3) / (7 * np.pi ** 4)) - 1) < 1e-3
    massive = OC.nu_density_ratio(0.0, p["Tcmb0"], p["Neff"], p["m_nu"]) - 2 / 3 * r0
    assert abs(og * massive * h * h / (0.06 / 93.14) - 1) < 0.015
    # no photons: the flat-LCDM distance (times h / (H0/100) = 1 in Mpc/h)
    z = np.array([0.1, 0.55, 1.0, 2.5])
    np.testing.assert_allclose(OC.comoving_distance_flrw(z, 1.0, 100.0, 0.3, 1e-30, 3.046, (0.0, 0.0, 0.0)),
                               OC.comoving_distance(z, 0.3), rtol=1e-12)
    # matter + radiation only (a 19 K photon bath, massless neutrinos, so
    # Om + Or = 1 and no Lambda): with a = 1/(1+z),
    # chi = c/H0 * 2 (sqrt(Om + Or) - sqrt(Om a + Or)) / Om
    H0 = 100.0
    orad = OC.omega_gamma(H0, 19.0) * (1 + 7 / 8 * (4 / 11) ** (4 / 3) * 3.0)
    om = 1.0 - orad
    assert 0.05 < orad < 0.5
    chi = OC.comoving_distance_flrw(z, 1.0, H0, om, 19.0, 3.0, (0.0, 0.0, 0.0))
    exact = CH * 2.0 * (np.sqrt(om + orad) - np.sqrt(om / (1 + z) + orad)) / om
    np.testing.assert_allclose(chi, exact, rtol=1e-10)
    # the Planck18 distances: smaller than the no-radiation ones of the
    # same Om0 (radiation and massive neutrinos raise E) by a small amount
    d18 = OC.comoving_distance_flrw(z, h)
    dl = OC.comoving_distance(z, p["Om0"])
    assert np.all(d18 < dl) and np.all(dl / d18 - 1 < 2e-3)
