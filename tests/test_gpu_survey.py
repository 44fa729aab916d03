"""Survey-geometry parity: the HIP path against the oracle on identical seeded
catalogs (FP64: 1e-10 relative per bin, or relative to max|value| for bins
that cross zero; FP32 fields with split-bf16 contraction: 5e-6 of max|B|)."""
import numpy as np
import pytest

from helpers import COSMO, NO_COMP, oracle_survey_cats, rel_to_max, survey_rows

pytestmark = pytest.mark.gpu

N, L = 24, 300.0
TOL = 1e-10


def _rows(tracers, seed=21):
    out = {}
    for i, t in enumerate(tracers):
        d, r = survey_rows(2500, 10000, seed + 17 * i)
        out[f"data_{t}"], out[f"randoms_{t}"] = d, r
    return out


def _close(a, b, tol=TOL, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    if np.max(np.abs(b)) == 0:
        assert np.max(np.abs(a)) == 0, what
        return
    err = rel_to_max(a, b)
    assert err < tol, f"{what}: {err:.3e}"


def test_radecz_planck18_default(gpu_ctx):
    """No 'Omega0' in the cosmology: the default Planck18 branch (photons,
    two massless and one 0.06 eV neutrino; distances in Mpc times the
    config's h) against the oracle's FLRW quadrature, to 1e-12."""
    from cosmonpc_amd.catalog import read_survey_catalog
    from oracle.cnpc_ref import catalog as OC

    d, _ = survey_rows(3000, 10, 6)
    for h in (0.6766, 0.7):
        cat, bc = read_survey_catalog(gpu_ctx, None, ["RA", "DEC", "Z"], (0.45, 0.5), NO_COMP, {"h": h}, "data",
                                      rows=d)
        ref = OC.survey_catalog(d, (0.45, 0.5), NO_COMP, {"h": h}, "data")
        np.testing.assert_allclose(cat.pos.cpu().numpy(), ref["Position"], rtol=1e-12, atol=1e-9)


def test_radecz_and_particle_sums(gpu_ctx):
    import torch

    from cosmonpc_amd import angular as A
    from cosmonpc_amd.catalog import read_survey_catalog
    from cosmonpc_amd.survey import compute_particle_sums_bk, compute_particle_sums_pk
    from oracle.cnpc_ref import catalog as OC
    from oracle.cnpc_ref import survey as S

    d, r = survey_rows(2500, 10000, 5, extra_cols=True)
    plan = {"scheme": "boss", "name_alias": None}
    cat, bc = read_survey_catalog(gpu_ctx, None, ["RA", "DEC", "Z"], (0.46, 0.49), plan, COSMO, "data", rows=d)
    ref = OC.survey_catalog(d, (0.46, 0.49), plan, COSMO, "data")
    np.testing.assert_allclose(cat.pos.cpu().numpy(), ref["Position"], rtol=1e-13, atol=1e-10)
    np.testing.assert_array_equal(cat.weight.cpu().numpy(), ref["WEIGHT"])
    np.testing.assert_allclose(bc, 0.5 * (ref["Position"].min(0) + ref["Position"].max(0)), rtol=1e-13)
    rcat = read_survey_catalog(gpu_ctx, None, ["RA", "DEC", "Z"], None, plan, COSMO, "randoms", boxcenter=bc,
                               rows=r)
    rref = OC.survey_catalog(r, None, plan, COSMO, "randoms")
    got = compute_particle_sums_pk(gpu_ctx, cat, rcat, "auto")
    exp = S.compute_particle_sums_pk(ref, rref, "auto")
    np.testing.assert_allclose(got, exp, rtol=1e-12)
    for L_ in (0, 2, 4):
        g = compute_particle_sums_bk(gpu_ctx, cat, rcat, "auto", [L_, 0, L_])
        e = S.compute_particle_sums_bk(ref, rref, "auto", [L_, 0, L_])
        np.testing.assert_allclose(g[:5], e[:5], rtol=1e-12)
        np.testing.assert_allclose(g[5], e[5], rtol=1e-11, atol=1e-12 * np.abs(e[5]).max())
        np.testing.assert_allclose(g[6], e[6], rtol=1e-11, atol=1e-12 * np.abs(e[6]).max())
    del torch, A


PK_CASES = [
    ("auto", "particle", "tsc", True, [0, 2, 4]),
    ("auto", "mixed-mesh", "cic", False, [0, 1, 2, 3, 4]),
    ("auto", "particle", "pcs", True, [1, 2]),
    ("cross", "mixed-mesh", "pcs", True, [0, 2]),
    ("cross", "mixed-mesh", "ngp", False, [0, 1, 3]),
]


@pytest.mark.parametrize("corr,ns,sampler,interlaced,poles", PK_CASES)
def test_pk_survey_vs_oracle(gpu_ctx, corr, ns, sampler, interlaced, poles):
    from cosmonpc_amd import survey as GS
    from oracle.cnpc_ref import survey as S

    print(f"[pk survey] {corr} {ns} {sampler} interlaced={interlaced} poles={poles}", flush=True)
    rows = _rows(["a"] if corr == "auto" else ["a", "b"])
    kw = dict(k_min=0.01, k_max=0.24, k_bins=8, poles=poles, boxsize=[L] * 3, nmesh=[N] * 3, sampler=sampler,
              interlaced=interlaced, normalization_scheme=ns)
    attrs, fa, fb = GS.get_mesh_pk_survey(None, corr, [N] * 3, "survey-like", ["RA", "DEC", "Z"], [L] * 3, sampler,
                                          interlaced, None, NO_COMP, COSMO, normalization_scheme=ns, arrays=rows)
    res = GS.calculate_power_spectrum_survey(dict(attrs, **kw), fa, fb, corr)
    cats = oracle_survey_cats(rows, ns=ns)
    oattrs, oa, ob = S.get_mesh_pk_survey(cats, corr, N, L, sampler, interlaced, ns)
    ores = S.calculate_power_spectrum_survey(dict(oattrs, **kw), oa, ob, corr)
    for key in oattrs:
        np.testing.assert_allclose(attrs[key], oattrs[key], rtol=1e-12, err_msg=key)
    np.testing.assert_array_equal(res["k_num"], ores["k_num"])
    np.testing.assert_allclose(res["k_eff"], ores["k_eff"], rtol=1e-14)
    for p in poles:
        assert np.iscomplexobj(res[f"P{p}"]) == np.iscomplexobj(ores[f"P{p}"]), p
        _close(res[f"P{p}"], ores[f"P{p}"], what=f"P{p}")


@pytest.mark.parametrize("poles,mode", [([0, 2, 4], "fast"), ([0, 2, 4, 6], "fast"), ([0, 2, 4, 6, 8], "compare"),
                                        ([0, 2, 4, 6, 8], "fast")])
def test_pk_survey_high_order_modes(gpu_ctx, poles, mode):
    from cosmonpc_amd import survey as GS
    from oracle.cnpc_ref import survey as S

    print(f"[pk survey high-order] {poles} {mode}", flush=True)
    rows = _rows(["a"], seed=61)
    kw = dict(k_min=0.01, k_max=0.24, k_bins=8, poles=poles, boxsize=[L] * 3, nmesh=[N] * 3, sampler="tsc",
              interlaced=True, normalization_scheme="particle", high_order_mode=mode)
    attrs, fa, fb = GS.get_mesh_pk_survey(None, "auto", [N] * 3, "survey-like", ["RA", "DEC", "Z"], [L] * 3, "tsc",
                                          True, None, NO_COMP, COSMO, normalization_scheme="particle", arrays=rows)
    res = GS.calculate_power_spectrum_survey(dict(attrs, **kw), fa, fb, "auto")
    cats = oracle_survey_cats(rows)
    oattrs, oa, ob = S.get_mesh_pk_survey(cats, "auto", N, L, "tsc", True, "particle")
    ores = S.calculate_power_spectrum_survey(dict(oattrs, **kw), oa, ob, "auto")
    assert set(res) == set(ores)
    for key in ores:
        if key.startswith("P"):
            _close(res[key], ores[key], what=key)


BK_CASES = [
    ("auto", "aaa", [0, 0, 0], "full", "particle", "both", "tsc", True, "gemm_fp64"),
    ("auto", "aaa", [2, 0, 2], "diagonal", "mesh", "ana", "cic", False, "gemm_fp64"),
    ("auto", "aaa", [1, 1, 0], "full", "mixed-mesh", "fft", "pcs", True, "pairwise"),
    ("auto", "aaa", [2, 0, 2], "full", "particle", "both", "tsc", True, "gemm_fp64"),
    ("cross", "aab", [0, 0, 0], "full", "mixed-mesh", "both", "cic", True, "gemm_fp64"),
    ("cross", "abb", [2, 0, 2], "diagonal", "mixed-mesh", "both", "tsc", False, "gemm_fp64"),
    ("cross", "abc", [1, 1, 0], "full", "mixed-mesh", "ana", "cic", True, "gemm_fp64"),
]


@pytest.mark.parametrize("corr,tracer,angu,dvm,ns,sn,sampler,interlaced,contraction", BK_CASES)
def test_bk_survey_vs_oracle(gpu_ctx, corr, tracer, angu, dvm, ns, sn, sampler, interlaced, contraction):
    from cosmonpc_amd import survey as GS
    from oracle.cnpc_ref import survey as S

    print(f"[bk survey] {corr} {tracer} {angu} {dvm} {ns} {sn} {sampler} interlaced={interlaced} {contraction}",
          flush=True)
    labels = {"aaa": ["a"], "aab": ["a", "b"], "abb": ["a", "b"], "abc": ["a", "b", "c"]}[tracer]
    rows = _rows(labels, seed=31)
    kw = dict(k_min=0.02, k_max=0.1, k_bins=4, angu_config=angu, boxsize=[L] * 3, nmesh=[N] * 3, sampler=sampler,
              interlaced=interlaced, tracer_type=tracer, data_vector_mode=dvm, shotnoise_mode=sn,
              contraction=contraction)
    attrs, fa, fb, fc = GS.get_mesh_bk_survey(None, corr, tracer, angu, [N] * 3, "survey-like", ["RA", "DEC", "Z"],
                                              [L] * 3, sampler, interlaced, None, NO_COMP, COSMO,
                                              normalization_scheme=ns, arrays=rows)
    res = GS.calculate_bk_sugi_survey(fa, fb, fc, corr, dict(attrs, **kw))
    cats = oracle_survey_cats(rows, ns=ns)
    oattrs, oa, ob, oc = S.get_mesh_bk_survey(cats, corr, tracer, angu, N, L, sampler, interlaced, ns)
    ores = S.calculate_bk_sugi_survey(oa, ob, oc, corr, dict(oattrs, **kw), cats)
    for key in oattrs:
        if key == "normalization_scheme":
            assert attrs[key] == oattrs[key]
            continue
        np.testing.assert_allclose(attrs[key], oattrs[key], rtol=1e-11, err_msg=key)
    np.testing.assert_array_equal(res["nmodes"], ores["nmodes"])
    for key in ("Bk_raw", "B_sugi", "Shot_noise"):
        _close(res[key], ores[key], what=key)
    for key, v in ores["SN_terms"].items():
        if np.max(np.abs(v)) > 0:
            _close(res["SN_terms"][key], v, what=key)
        else:
            assert np.max(np.abs(res["SN_terms"][key])) == 0, key


def test_bk_survey_fp32_split_bf16(gpu_ctx):
    """FP32 fields + split-bf16 MFMA contraction against the FP64 oracle:
    5e-6 of max|B|."""
    from cosmonpc_amd import survey as GS
    from oracle.cnpc_ref import survey as S

    rows = _rows(["a"], seed=41)
    angu = [2, 0, 2]
    kw = dict(k_min=0.02, k_max=0.1, k_bins=4, angu_config=angu, boxsize=[L] * 3, nmesh=[N] * 3, sampler="tsc",
              interlaced=True, tracer_type="aaa", data_vector_mode="full", shotnoise_mode="ana",
              contraction="bf16x3")
    attrs, fa, fb, fc = GS.get_mesh_bk_survey(None, "auto", "aaa", angu, [N] * 3, "survey-like",
                                              ["RA", "DEC", "Z"], [L] * 3, "tsc", True, None, NO_COMP, COSMO,
                                              precision="fp32", arrays=rows)
    res = GS.calculate_bk_sugi_survey(fa, fb, fc, "auto", dict(attrs, **kw))
    cats = oracle_survey_cats(rows)
    oattrs, oa, ob, oc = S.get_mesh_bk_survey(cats, "auto", "aaa", angu, N, L, "tsc", True, "particle")
    ores = S.calculate_bk_sugi_survey(oa, ob, oc, "auto", dict(oattrs, **kw), cats)
    err = rel_to_max(res["Bk_raw"], ores["Bk_raw"])
    print(f"survey bk fp32 split-bf16 rel max|B| {err:.2e}", flush=True)
    assert err < 5e-6


def test_run_stats_survey_boundary(gpu_ctx, tmp_path):
    """run_stats on survey configs: same files and result dicts as the oracle."""
    from cosmonpc_amd import run_stats
    from oracle.cnpc_ref import survey as S

    rows = _rows(["a"], seed=51)
    base = dict(correlation_mode="auto", geometry="survey-like",
                catalogs={"data_a": str(tmp_path / "gal.fits"), "randoms_a": str(tmp_path / "ran.fits")},
                column_names=["RA", "DEC", "Z", "WEIGHT_FKP", "NZ"], nmesh=[N] * 3, boxsize=[L] * 3,
                sampler="tsc", interlaced=True, k_min=0.02, k_max=0.1, k_bins=4, compensation=True,
                normalization_scheme="particle", cosmology=dict(COSMO), z_range=None, comp_weight_plan=dict(NO_COMP),
                output_dir=str(tmp_path / "out"), arrays=rows)
    cfg = dict(base, statistic="pk", poles=[0, 2, 4])
    run_stats(cfg)
    out = np.load(tmp_path / "out" / "pk_res_auto_gal.npy", allow_pickle=True).item()
    cats = oracle_survey_cats(rows)
    oattrs, oa, ob = S.get_mesh_pk_survey(cats, "auto", N, L, "tsc", True, "particle")
    ores = S.calculate_power_spectrum_survey(dict(oattrs, poles=[0, 2, 4], boxsize=[L] * 3, nmesh=[N] * 3,
                                                  k_min=0.02, k_max=0.1, k_bins=4, sampler="tsc", interlaced=True,
                                                  normalization_scheme="particle"), oa, ob, "auto")
    for p in (0, 2, 4):
        _close(out[f"P{p}"], ores[f"P{p}"], what=f"P{p}")
    assert out["high_order_mode"] == "default" and "z_range" in out
    cfg = dict(base, statistic="bk_sugi", tracer_type="aaa", angu_config=[0, 0, 0], data_vector_mode="diagonal",
               shotnoise_mode="ana", arrays=rows)
    run_stats(cfg)
    out = np.load(tmp_path / "out" / f"bk_sugi_res_diagonal_000_auto_aaa_{tmp_path.name}_gal.npy",
                  allow_pickle=True).item()
    kw = dict(k_min=0.02, k_max=0.1, k_bins=4, angu_config=[0, 0, 0], boxsize=[L] * 3, nmesh=[N] * 3,
              sampler="tsc", interlaced=True, tracer_type="aaa", data_vector_mode="diagonal", shotnoise_mode="ana")
    oattrs, oa, ob, oc = S.get_mesh_bk_survey(cats, "auto", "aaa", [0, 0, 0], N, L, "tsc", True, "particle")
    ores = S.calculate_bk_sugi_survey(oa, ob, oc, "auto", dict(oattrs, **kw), cats)
    _close(out["B_sugi"], ores["B_sugi"], what="B_sugi")
    assert "apply_rsd" not in out and out["block_size"] == 1
