"""GPU parity of the workload legs against the FP64 oracle on identical
seeded inputs, each at the largest size the oracle finishes in seconds:

* pk box, 128^3, CIC interlaced, 30 bins over 0.005-0.305, 2e5 lognormal +
  RSD particles, poles 0/2/4 (the full small-pk workload): 1e-10 per bin;
* bk_sugi box diagonal, TSC interlaced, [0,0,0] and [2,0,2], 19 bins over
  0.01-0.20 at 128^3 (the 256^3 workload's binning on half its grid): FP64
  1e-10; FP32 paint / FFT / Ylm with split-bf16, its error against FP64;
* survey pk (FP32 bulk) against the FP64 oracle;
* batched mocks (pk 0/2/4 + bk_sugi [0,0,0] diagonal, CIC interlaced, uneven
  catalog sizes) against the oracle's box estimators per mock;
* box bk cross-correlations aab / abb / abc.
"""
import numpy as np
import pytest

from tests.helpers import bk_attrs, catalog, pk_attrs, rel_err, rel_to_max, weights

pytestmark = pytest.mark.gpu


def _box():
    from oracle.cnpc_ref import box

    return box


def _log(msg):
    print(msg, flush=True)


def _per_bin(got, ref, tol):
    """Relative per bin; bins whose reference crosses zero (|ref| below 1e-3
    of max) relative to max |ref|."""
    got, ref = np.asarray(got, dtype=np.complex128), np.asarray(ref, dtype=np.complex128)
    big = np.abs(ref) > 1e-3 * np.max(np.abs(ref))
    assert rel_err(got[big], ref[big]) < tol
    assert rel_to_max(got, ref) < tol


def test_pk_box_128_full_workload(gpu_ctx):
    from cosmonpc_amd.estimators import calculate_power_spectrum_box
    from cosmonpc_amd.mesh import paint_field

    B = _box()
    n, L = 128, 1000.0
    _log("pk box 128^3 cic-i, 30 bins, 2e5 particles")
    pos = catalog(200000, L, 101, ngrid=256)
    w = np.ones(len(pos))
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, "cic", True)
    sa = dict(attrs)
    sa.update(pk_attrs(n, L, 0.005, 0.305, 30, [0, 2, 4], "cic", True))
    ref = B.calculate_power_spectrum_box(fa, None, "auto", sa)
    got = calculate_power_spectrum_box(paint_field(gpu_ctx, pos, w, n, L, "cic", True, 0), None, "auto", sa)
    np.testing.assert_array_equal(got["k_num"], ref["k_num"])
    assert rel_err(got["k_eff"], ref["k_eff"]) < 1e-12
    for p in (0, 2, 4):
        _per_bin(got[f"P{p}"], ref[f"P{p}"], 1e-10)


def _bk_diag(gpu_ctx, angu, prec, contraction, n=128, L=1000.0):
    from cosmonpc_amd.estimators import calculate_bk_sugi_box
    from cosmonpc_amd.mesh import paint_field

    B = _box()
    pos = catalog(250000, L, 102, ngrid=128)
    w = np.ones(len(pos))
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, "tsc", True)
    sa = dict(attrs)
    sa.update(bk_attrs(n, L, 0.01, 0.20, 19, angu, "diagonal", "tsc", True, contraction=contraction))
    ref = B.calculate_bk_sugi_box(fa, None, None, "auto", sa)
    got = calculate_bk_sugi_box(paint_field(gpu_ctx, pos, w, n, L, "tsc", True, prec), None, None, "auto", sa)
    return got, ref


@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2)])
def test_bk_box_diagonal_tsc_fp64(gpu_ctx, angu):
    _log(f"bk diagonal {angu} tsc-i 128^3 fp64")
    got, ref = _bk_diag(gpu_ctx, angu, 0, "gemm_fp64")
    np.testing.assert_array_equal(got["nmodes"], ref["nmodes"])
    for key in ("Bk_raw", "B_sugi"):
        assert rel_to_max(got[key], ref[key]) < 1e-10, key
    for key in ("SN0", "SN1", "SN2", "SN3"):
        a, b = np.asarray(got["SN_terms"][key]), np.asarray(ref["SN_terms"][key])
        if np.max(np.abs(b)) > 0:
            assert rel_to_max(a, b) < 1e-10, key
        else:
            assert np.max(np.abs(a)) == 0, key


@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2)])
def test_bk_box_diagonal_tsc_fp32(gpu_ctx, angu):
    """FP32 paint, FFTs, Ylm weights and shells; the diagonal sums in FP32
    with FP64 partials.  Error against the FP64 oracle relative to max|B|:
    measured ~1e-6; bar 1e-5."""
    _log(f"bk diagonal {angu} tsc-i 128^3 fp32")
    got, ref = _bk_diag(gpu_ctx, angu, 1, "bf16x3")
    np.testing.assert_array_equal(got["nmodes"], ref["nmodes"])
    for key in ("Bk_raw", "B_sugi"):
        err = rel_to_max(got[key], ref[key])
        _log(f"  {key} rel max|B| {err:.2e}")
        assert err < 1e-5, key


def test_pk_survey_fp32(gpu_ctx):
    """Survey pk with FP32 paint / FFT / Ylm (alpha, I22 and N0 in FP64)
    against the FP64 oracle: attrs to 1e-12, poles to 1e-5 of max|P_l|."""
    from cosmonpc_amd import survey as GS
    from oracle.cnpc_ref import survey as S
    from tests.helpers import COSMO, NO_COMP, oracle_survey_cats, survey_rows

    N, L = 32, 300.0
    d, r = survey_rows(4000, 20000, 91)
    rows = {"data_a": d, "randoms_a": r}
    kw = dict(k_min=0.01, k_max=0.3, k_bins=10, poles=[0, 2, 4], boxsize=[L] * 3, nmesh=[N] * 3, sampler="tsc",
              interlaced=True, normalization_scheme="particle")
    attrs, fa, fb = GS.get_mesh_pk_survey(None, "auto", [N] * 3, "survey-like", ["RA", "DEC", "Z"], [L] * 3, "tsc",
                                          True, None, NO_COMP, COSMO, normalization_scheme="particle", arrays=rows,
                                          precision="fp32")
    res = GS.calculate_power_spectrum_survey(dict(attrs, **kw), fa, fb, "auto")
    cats = oracle_survey_cats(rows)
    oattrs, oa, ob = S.get_mesh_pk_survey(cats, "auto", N, L, "tsc", True, "particle")
    ores = S.calculate_power_spectrum_survey(dict(oattrs, **kw), oa, ob, "auto")
    for key in oattrs:
        np.testing.assert_allclose(attrs[key], oattrs[key], rtol=1e-12, err_msg=key)
    np.testing.assert_array_equal(res["k_num"], ores["k_num"])
    for p in (0, 2, 4):
        err = rel_to_max(res[f"P{p}"], ores[f"P{p}"])
        _log(f"survey pk fp32 P{p} rel max {err:.2e}")
        assert err < 1e-5, p


def test_run_mocks_against_oracle(gpu_ctx, tmp_path):
    """run_mocks over uneven catalogs (batched paint, one launch per chunk)
    against the oracle's box pk and bk_sugi for each mock: FP64, 1e-10."""
    from cosmonpc_amd.mocks import run_mocks

    B = _box()
    n, L = 32, 500.0
    base = dict(correlation_mode="auto", geometry="box-like", column_names=["x", "y", "z"], rsd=[0, 0, 1],
                apply_rsd=False, nmesh=[n] * 3, boxsize=[L] * 3, sampler="cic", interlaced=True, compensation=True,
                cosmology={"h": 0.6766, "Omega0": 0.30966}, redshift_box=0.5, output_dir=str(tmp_path / "out"))
    pk = dict(base, statistic="pk", k_min=0.005, k_max=0.185, k_bins=9, poles=[0, 2, 4])
    bk = dict(base, statistic="bk_sugi", k_min=0.02, k_max=0.1, k_bins=4, angu_config=[0, 0, 0], tracer_type="aaa",
              data_vector_mode="diagonal", shotnoise_mode="ana")
    arrays = {}
    for i, npart in enumerate((5000, 1200, 8000, 2600)):
        path = str(tmp_path / "mocks" / f"mock_{i}.npy")
        arrays[path] = (catalog(npart, L, 200 + i, ngrid=32), weights(npart, seed=300 + i))
    got = run_mocks([pk, bk], list(arrays), batch=3, arrays=arrays)
    for path, (pos, w) in arrays.items():
        attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, "cic", True)
        sa = dict(attrs)
        sa.update(pk_attrs(n, L, 0.005, 0.185, 9, [0, 2, 4], "cic", True))
        ref = B.calculate_power_spectrum_box(fa, None, "auto", sa)
        res_pk, res_bk = got[path]
        np.testing.assert_array_equal(res_pk["k_num"], ref["k_num"])
        for p in (0, 2, 4):
            _per_bin(res_pk[f"P{p}"], ref[f"P{p}"], 1e-10)
        sb = dict(attrs)
        sb.update(bk_attrs(n, L, 0.02, 0.1, 4, (0, 0, 0), "diagonal", "cic", True))
        refb = B.calculate_bk_sugi_box(fa, None, None, "auto", sb)
        for key in ("Bk_raw", "B_sugi", "Shot_noise"):
            assert rel_to_max(res_bk[key], refb[key]) < 1e-10, (path, key)


@pytest.mark.parametrize("tracer", ["aab", "abb", "abc"])
@pytest.mark.parametrize("dvm", ["full", "diagonal"])
def test_bk_box_cross_matches_oracle(gpu_ctx, tracer, dvm):
    """Box cross-correlation bispectra (``clustering_estimator.py:749-1061``
    cross branches, shot-noise tables ``:1386-1468``, I_norm ``:1497-1521``).
    abc has no shot noise and no FFT S3: with shotnoise_mode 'both' the
    reference divides None by I_norm (``:1534``) and raises TypeError, and so
    does the drop-in."""
    from cosmonpc_amd.estimators import calculate_bk_sugi_box
    from cosmonpc_amd.mesh import paint_field

    B = _box()
    n, L = 32, 300.0
    cats = {"a": (catalog(20000, L, 111, ngrid=n), weights(20000, 1)),
            "b": (catalog(15000, L, 112, ngrid=n), weights(15000, 2)),
            "c": (catalog(12000, L, 113, ngrid=n), weights(12000, 3))}
    _log(f"bk box cross {tracer} {dvm}")
    attrs, fa, fb, fc = B.get_mesh_box(cats, [n] * 3, [L] * 3, "tsc", True, correlation_mode="cross",
                                       tracer_type=tracer)
    g = {t: paint_field(gpu_ctx, *cats[t], n, L, "tsc", True, 0) for t in cats if t in "ab" or tracer == "abc"}
    sn = "fft" if tracer == "abc" else "both"
    for angu in ((0, 0, 0), (2, 0, 2)):
        sa = dict(attrs)
        sa.update(bk_attrs(n, L, 0.03, 0.15, 6, angu, dvm, "tsc", True, tracer_type=tracer, sn=sn,
                           contraction="gemm_fp64"))
        ref = B.calculate_bk_sugi_box(fa, fb, fc, "cross", sa)
        got = calculate_bk_sugi_box(g["a"], g["b"], g.get("c"), "cross", sa)
        assert rel_to_max(got["Bk_raw"], ref["Bk_raw"]) < 1e-10, angu
        assert rel_to_max(got["B_sugi"], ref["B_sugi"]) < 1e-10, angu
        assert got["I_norm"] == pytest.approx(ref["I_norm"], rel=1e-14)
        assert sorted(got["SN_terms"]) == sorted(ref["SN_terms"])
        for key in ref["SN_terms"]:
            a, b = np.asarray(got["SN_terms"][key]), np.asarray(ref["SN_terms"][key])
            if np.max(np.abs(b)) > 0:
                assert rel_to_max(a, b) < 1e-10, (angu, key)
            else:
                assert np.max(np.abs(a)) == 0, (angu, key)
    if tracer == "abc":
        sa["shotnoise_mode"] = "both"
        with pytest.raises(TypeError):
            B.calculate_bk_sugi_box(fa, fb, fc, "cross", sa)
        with pytest.raises(TypeError):
            calculate_bk_sugi_box(g["a"], g["b"], g["c"], "cross", sa)


def test_run_stats_box_hdf5_catalog(gpu_ctx, tmp_path):
    """run_stats on a box catalog in HDF5 (written by h5py, read without
    it): pk of the file's x/y/z and w columns against the oracle, FP64,
    1e-10 per bin; the saved result names the catalog like the reference."""
    import os

    from cosmonpc_amd import run_stats

    B = _box()
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hdf5")
    exp = np.load(os.path.join(gold, "expected.npz"))
    pos = np.column_stack([exp[f"box_default/{a}"] for a in ("x", "y", "z")])
    w = exp["box_default/w"]
    n, L = 32, 1000.0
    cfg = dict(statistic="pk", correlation_mode="auto", geometry="box-like",
               catalogs={"data_a": os.path.join(gold, "box_default.h5"), "randoms_a": None, "data_b": None,
                         "randoms_b": None},
               column_names=["x", "y", "z", "w"], rsd=[0, 0, 1], apply_rsd=False, nmesh=[n] * 3, boxsize=[L] * 3,
               sampler="tsc", interlaced=True, k_min=0.005, k_max=0.09, k_bins=8, poles=[0, 2], compensation=True,
               cosmology={"h": 0.6766, "Omega0": 0.30966}, redshift_box=0.5, output_dir=str(tmp_path / "out"))
    _log("run_stats box_default.h5")
    got = run_stats(cfg)
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, "tsc", True)
    sa = dict(attrs)
    sa.update(pk_attrs(n, L, 0.005, 0.09, 8, [0, 2], "tsc", True))
    ref = B.calculate_power_spectrum_box(fa, None, "auto", sa)
    np.testing.assert_array_equal(got["k_num"], ref["k_num"])
    for p in (0, 2):
        _per_bin(got[f"P{p}"], ref[f"P{p}"], 1e-10)
    assert os.path.exists(os.path.join(str(tmp_path / "out"), "pk_res_auto_box_default.npy"))
