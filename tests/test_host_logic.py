"""Host logic of the run_stats boundary: validation errors (types and
messages of the reference's param_helper.py / task_executor.py), config
normalisation, and failures raised before any GPU work."""
import copy
import os

import numpy as np
import pytest

from cosmonpc_amd import param_helper as P


def _base(**kw):
    c = {
        "statistic": "pk", "correlation_mode": "auto", "geometry": "box-like",
        "catalogs": {"data_a": "a.npy", "randoms_a": None, "data_b": None, "randoms_b": None},
        "column_names": ["x", "y", "z"], "rsd": [0, 0, 1], "apply_rsd": False,
        "nmesh": [128, 128, 128], "boxsize": [1000.0, 1000.0, 1000.0], "sampler": "cic", "interlaced": True,
        "k_min": 0.005, "k_max": 0.305, "k_bins": 30, "poles": [0, 2, 4], "compensation": True,
        "cosmology": {"h": 0.6766, "Omega0": 0.30966}, "redshift_box": 0.5, "output_dir": "/tmp/cnpc_out",
    }
    c.update(kw)
    return c


def test_validate_config_aliases_and_defaults():
    c = _base(**{"shotnoise-mode": "fft", "use_fast_mode": False})
    P.validate_config(c)
    assert c["shotnoise_mode"] == "fft" and "shotnoise-mode" not in c
    assert c["high_order_mode"] == "default" and "use_fast_mode" not in c
    c = _base()
    P.validate_config(c)
    assert c["shotnoise_mode"] == "ana" and c["high_order_mode"] == "default"
    c = _base(high_order_mode="fast", use_fast_mode=True, geometry="survey-like", poles=[0, 2, 4])
    P.validate_config(c)
    assert "use_fast_mode" not in c


@pytest.mark.parametrize("kw,exc,msg", [
    ({"sampler": "xyz"}, AssertionError, "sampler should be 'ngp', 'cic', 'pcs', or 'tsc'"),
    ({"nmesh": [64, 64]}, AssertionError, "nmesh must be a list or tuple of three elements"),
    ({"nmesh": [64, 64, 32]}, AssertionError, "All elements in nmesh must be equal"),
    ({"nmesh": [64.0, 64, 64]}, AssertionError, "All elements in nmesh must be positive ints"),
    ({"boxsize": [1000.0, -1.0, 1000.0]}, AttributeError, "__name__"),
    ({"shotnoise_mode": "x"}, ValueError, "shotnoise_mode must be either 'ana', 'fft', or 'both'"),
    ({"high_order_mode": "x"}, ValueError, "high_order_mode must be one of 'default', 'fast', or 'compare'"),
    ({"high_order_mode": "fast"}, ValueError, "only supported for survey-like pk measurements"),
    ({"interlaced": 1}, TypeError, "Config field 'interlaced' must be a boolean (True/False), got int: 1"),
])
def test_validate_config_errors(kw, exc, msg):
    c = _base(**kw)
    with pytest.raises(exc) as e:
        P.validate_config(c)
    assert msg in str(e.value)


def test_validate_poles_and_sugi_poles():
    P.validate_poles([0, 2, 4])
    with pytest.raises(ValueError, match="non-negative integers"):
        P.validate_poles([0, -2])
    with pytest.raises(ValueError, match="duplicate"):
        P.validate_poles([0, 0])
    with pytest.raises(ValueError, match="ascending"):
        P.validate_poles([2, 0])
    P.validate_sugi_poles([2, 0, 2], "box-like")
    for cfg, msg in (([0, 2, 2], "ell1 must be greater"), ([1, 0, 0], "must be even"), ([1, 1, 1], "must be even"),
                     ([2, 0, 4], "triangle condition")):
        with pytest.raises(ValueError, match=msg):
            P.validate_sugi_poles(cfg, "box-like")
    with pytest.raises(ValueError, match="geometry must be"):
        P.validate_sugi_poles([0, 0, 0], "sphere")


def test_validate_tracer_and_catalog_check():
    P.validate_tracer("aaa", "auto")
    with pytest.raises(AssertionError, match="For auto-correlation"):
        P.validate_tracer("aab", "auto")
    with pytest.raises(AssertionError, match="tracer_type must be one of"):
        P.validate_tracer("xyz", "cross")
    with pytest.raises(AssertionError, match="data_b catalog must be provided"):
        P.catalog_check({"data_a": "a", "data_b": None}, "box-like", "cross", "pk")
    with pytest.raises(AssertionError, match="randoms_a catalog must be provided for survey-like auto"):
        P.catalog_check({"data_a": "a", "randoms_a": None}, "survey-like", "auto", "pk")
    with pytest.raises(AssertionError, match="data_c catalog must be provided"):
        P.catalog_check({"data_a": "a", "data_b": "b"}, "box-like", "cross", "bk_sugi", tracer_type="abc")


def test_run_task_errors_before_gpu_work():
    from cosmonpc_amd.task_executor import run_task

    c = _base(k_max=0.3)
    c["nmesh"] = [32, 32, 32]  # k_Nyq = 0.1005
    with pytest.raises(ValueError, match="is larger than the Nyquist frequency"):
        run_task(**copy.deepcopy(c))
    c = _base(poles=[1, 3])
    with pytest.raises(ValueError, match="All poles in"):
        run_task(**c)
    c = _base(statistic="bk_sugi", tracer_type="aaa", angu_config=[0, 0, 0], k_max=0.2)
    c["nmesh"] = [64, 64, 64]  # 2 k_max = 0.4 >= k_Nyq = 0.201
    with pytest.raises(ValueError, match="2 \\* k_max"):
        run_task(**c)
    c = _base(statistic="bk_sugi", tracer_type="aaa", angu_config=[0, 0, 0], k_max=0.05,
              data_vector_mode="both")
    with pytest.raises(ValueError, match="data_vector_mode must be either"):
        run_task(**c)
    c = _base(statistic="bk_sugi", tracer_type="aab", angu_config=[0, 0, 0], k_max=0.05)
    with pytest.raises(AssertionError, match="For auto-correlation"):
        run_task(**c)


def test_catalog_reader_box_npy(tmp_path):
    from cosmonpc_amd import catalog, synth

    pos = np.random.default_rng(0).uniform(0, 10, (50, 3))
    w = np.linspace(1, 2, 50)
    f = tmp_path / "cat.npy"
    cols = synth.write_npy_catalog(str(f), pos, w)
    p, ww = catalog.read_box_catalog(str(f), cols)
    np.testing.assert_array_equal(p, pos)
    np.testing.assert_array_equal(ww, w)
    p, ww = catalog.read_box_catalog(str(f), ["x", "y", "z", "other"])
    assert np.all(ww == 1.0)
    with pytest.raises(ValueError, match="Unsupported data file type: txt"):
        catalog.read_box_catalog("cat.txt", cols)


def test_catalog_reader_box_fits_and_rsd(tmp_path):
    """Box FITS catalogs (catalog_processor.py:388-449): X/Y/Z or x/y/z
    positions, weight 'w', RSD shift (1+z)/(100 E(z)) v along the LOS with a
    periodic wrap; E(z) = (1+z)^1.5 for Omega_m = 1."""
    from cosmonpc_amd import catalog, fits

    rng = np.random.default_rng(1)
    pos = rng.uniform(0, 100, (200, 3))
    vel = rng.normal(0, 300, (200, 3))
    w = rng.uniform(0.5, 1.5, 200)
    f = tmp_path / "box.fits"
    fits.write_table(str(f), {"X": pos[:, 0], "Y": pos[:, 1], "Z": pos[:, 2], "VX": vel[:, 0], "VY": vel[:, 1],
                              "VZ": vel[:, 2], "w": w})
    cols = ["X", "Y", "Z", "VX", "VY", "VZ", "w"]
    p, ww = catalog.read_box_catalog(str(f), cols)
    np.testing.assert_array_equal(p, pos)
    np.testing.assert_array_equal(ww, w)
    zb = 0.5
    p, _ = catalog.read_box_catalog(str(f), cols, apply_rsd=True, para_cosmo={"h": 0.7, "Omega0": 1.0},
                                    boxsize=[100.0] * 3, redshift_box=zb, los=[0, 0, 1])
    fac = (1 + zb) / ((1 + zb) ** 1.5 * 100)
    np.testing.assert_allclose(p[:, 2], np.mod(pos[:, 2] + fac * vel[:, 2], 100.0), rtol=1e-13)
    np.testing.assert_array_equal(p[:, :2], pos[:, :2])
    with pytest.raises(ValueError, match="LOS must be a 3d unit vector"):
        catalog.read_box_catalog(str(f), cols, apply_rsd=True, para_cosmo={"h": 0.7, "Omega0": 0.3},
                                 boxsize=[100.0] * 3, redshift_box=zb, los=[0, 1, 1])
    # columns not requested are not read: no weight column -> 1
    p, ww = catalog.read_box_catalog(str(f), ["X", "Y", "Z"])
    assert np.all(ww == 1.0)


def test_fits_format_cards_and_foreign_file(tmp_path):
    """Header cards follow the FITS fixed format (keyword in columns 1-8,
    '= ' in 9-10, values right-justified to column 30), and a table written
    by an independent byte-level writer reads back."""
    import struct

    from cosmonpc_amd import fits

    f = tmp_path / "a.fits"
    fits.write_table(str(f), {"RA": np.array([1.5, 2.5]), "N": np.array([3, 4], dtype=np.int32)})
    raw = f.read_bytes()
    assert len(raw) % 2880 == 0
    assert raw[:80].decode() == "SIMPLE  =" + " " * 20 + "T" + " " * 50
    ext = raw[2880:2880 + 2880].decode()
    assert ext[:80].startswith("XTENSION= 'BINTABLE'")
    assert "TFORM1  = 'D       '" in ext and "TFORM2  = 'J       '" in ext

    def card(s):
        return s.ljust(80).encode()

    hdr0 = card("SIMPLE  =                    T") + card("BITPIX  =                    8") + \
        card("NAXIS   =                    0") + card("END")
    hdr1 = b"".join(card(c) for c in [
        "XTENSION= 'BINTABLE'", "BITPIX  =                    8", "NAXIS   =                    2",
        "NAXIS1  =                   12", "NAXIS2  =                    3", "PCOUNT  =                    0",
        "GCOUNT  =                    1", "TFIELDS =                    2", "TTYPE1  = 'Z       '",
        "TFORM1  = 'D       '", "TTYPE2  = 'WEIGHT_FKP'", "TFORM2  = 'E       '", "END"])
    data = b"".join(struct.pack(">df", z, w) for z, w in [(0.5, 1.0), (0.6, 0.25), (0.7, 2.0)])

    def pad(b, c):
        return b + c * ((-len(b)) % 2880)

    g = tmp_path / "b.fits"
    g.write_bytes(pad(hdr0, b" ") + pad(hdr1, b" ") + pad(data, b"\0"))
    t = fits.read_table(str(g), ["WEIGHT_FKP", "Z", "missing"])
    assert t.dtype.names == ("WEIGHT_FKP", "Z")
    np.testing.assert_array_equal(t["Z"], [0.5, 0.6, 0.7])
    np.testing.assert_array_equal(t["WEIGHT_FKP"], np.array([1.0, 0.25, 2.0], dtype=np.float32))


def test_run_mocks_validation_and_resume(tmp_path):
    """The batched mock driver accepts only box-like auto configs that share
    one mesh, and skips catalogs whose outputs all exist (no GPU needed)."""
    from cosmonpc_amd.mocks import LAST_TIMINGS, run_mocks

    out = tmp_path / "out"
    pk = _base(output_dir=str(out))
    bk = _base(output_dir=str(out), statistic="bk_sugi", tracer_type="aaa", angu_config=[0, 0, 0],
               data_vector_mode="diagonal", k_max=0.1)
    with pytest.raises(ValueError, match="box-like auto-correlation configs only"):
        run_mocks([_base(correlation_mode="cross", catalogs={"data_a": "a", "data_b": "b"})], ["m.npy"])
    with pytest.raises(ValueError, match="must share 'nmesh'"):
        run_mocks([pk, dict(bk, nmesh=[64] * 3)], ["m.npy"])
    with pytest.raises(ValueError, match="at least one config"):
        run_mocks([], ["m.npy"])
    with pytest.raises(ValueError, match="batch must be"):
        run_mocks([pk], ["m.npy"], batch=0)
    with pytest.raises(ValueError, match="data_vector_mode must be"):
        run_mocks([dict(bk, data_vector_mode="band")], ["m.npy"])
    out.mkdir()
    mocks = [str(tmp_path / "mocks" / f"mock_{i}.npy") for i in range(3)]
    for m in mocks:
        np.save(out / f"pk_res_auto_{os.path.splitext(os.path.basename(m))[0]}.npy", {})
        np.save(out / f"bk_sugi_res_diagonal_000_auto_aaa_mocks_{os.path.splitext(os.path.basename(m))[0]}.npy", {})
    assert run_mocks([pk, bk], mocks) == {}
    assert LAST_TIMINGS["mocks"] == 0


def test_autotune_cache_and_disable(tmp_path, monkeypatch):
    """Disabled tuning returns the first candidate without running anything;
    cached choices are looked up per shape and persist through the JSON
    cache file."""
    import json

    from cosmonpc_amd import autotune

    monkeypatch.setattr(autotune, "_CHOICES", {})
    monkeypatch.setattr(autotune, "_LOADED", False)
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"gemm_blocks:(1, 2)": 4096}))
    monkeypatch.setenv("CNPC_TUNE_CACHE", str(path))
    monkeypatch.setenv("CNPC_TUNE", "0")
    ran = []
    assert autotune.lookup("gemm_blocks", (1, 2)) == 4096
    assert autotune.choose("gemm_blocks", (1, 2), [2048, 1024], ran.append) == 4096
    assert autotune.choose("synth_half_batch", (400, 99, 30), [4, 8], ran.append) == 4
    assert ran == [] and autotune.lookup("synth_half_batch", (400, 99, 30)) is None


def test_zbatch_pairs_same_and_pair_requests(monkeypatch):
    """A B == A request on shells X and a B != A request (Y, X) with the same
    G, X buffer, rows and grid become one two-output launch; a request with
    another G, or rows that differ, is launched on its own.  The frozen
    autotuner times nothing."""
    from cosmonpc_amd import autotune, estimators

    calls = []

    class FakeCtx:
        def contract_zgemm(self, *z, nblk=0, with_bb=False):
            calls.append((z[4], z[7], with_bb))
            return ("C", "Cbb") if with_bb else "C"

    monkeypatch.setattr(estimators, "_zgemm_blocks", lambda ctx, z, bb=False: 0)
    G, G2, X, Y, Z = object(), object(), object(), object(), object()
    kz = np.arange(4, dtype=np.int32)
    b = estimators.ZBatch(FakeCtx())
    same = b.request((40, 500.0, 0.1, G, X, 4, kz, None, 4, None))
    pair = b.request((40, 500.0, 0.1, G, Y, 4, kz, X, 4, kz))
    other_g = b.request((40, 500.0, 0.1, G2, Z, 4, kz, X, 4, kz))
    lone = b.request((40, 500.0, 0.1, G, Z, 3, kz[:3], None, 3, None))
    with autotune.frozen():
        b.run()
    assert (Y, X, True) in calls and len([c for c in calls if c[2]]) == 1
    assert pair.value == "C" and same.value == "Cbb"
    assert other_g.value == "C" and lone.value == "C"
    assert len(calls) == 3 and b.requests == []


def test_q_ells_memo_rejects_nearby_keys():
    """The memo of q_l(k1, k2, k3) serves a hit only for k values equal to
    atomic-order noise (1e-13 relative); k values that share float32 images
    but differ beyond that get their own q."""
    from cosmonpc_amd import angular as A
    from oracle.cnpc_ref import angular as R

    A._Q_CACHE.clear()
    k_eff = np.array([0.05, 0.07, 0.09, 0.11])
    e = np.linspace(2 * 0.04, 2 * 0.12, 2 * 4 + 1)[:-1]
    k3 = 0.5 * (e[1:] + e[:-1])
    q0 = A.get_q_ells_eff(1, 2, k_eff, 2, 0, 2, len(k3), k3)
    np.testing.assert_allclose(q0, R.get_q_ells_eff(1, 2, k_eff, 2, 0, 2, len(k3), k3), rtol=1e-13, atol=1e-15)
    # last-bit noise: served from the memo
    noisy = k_eff * (1 + 2e-16)
    assert np.array_equal(A.get_q_ells_eff(1, 2, noisy, 2, 0, 2, len(k3), k3), q0)
    assert len(A._Q_CACHE) == 1
    # the same float32 images, 1e-9 apart: recomputed at the new values
    moved = k_eff * (1 + 1e-9)
    assert np.array_equal(moved.astype(np.float32), k_eff.astype(np.float32))
    q1 = A.get_q_ells_eff(1, 2, moved, 2, 0, 2, len(k3), k3)
    np.testing.assert_allclose(q1, R.get_q_ells_eff(1, 2, moved, 2, 0, 2, len(k3), k3), rtol=1e-13, atol=1e-15)
    assert not np.array_equal(q1, q0)


@pytest.mark.parametrize("kind", ["library", "torch"])
def test_block_size_retry_on_device_oom(monkeypatch, kind):
    """A device allocation failure while a configuration's shell blocks are
    built (a library call returning hipErrorOutOfMemory, or torch's
    out-of-memory error) rebuilds the blocks with half the block size, down
    to one bin; the signal matrix is the one of the unconstrained run.  Other
    errors, and an out-of-memory error at one bin, propagate."""
    import torch

    from cosmonpc_amd import estimators
    from cosmonpc_amd._lib import HIP_ERROR_OUT_OF_MEMORY, CnpcError

    monkeypatch.setattr(estimators, "_fit_block", lambda k_bins, b, ctx, reserve=0.25: k_bins)
    n, k_bins = 4, 6

    class FakeCtx:
        def __init__(self, limit, err=None):
            self.limit, self.err, self.rows = limit, err, []

        def mesh(self, n, prec, rows=None):
            self.rows.append(rows)
            if self.err is not None:
                raise self.err
            if self.limit is not None and rows > self.limit:
                if kind == "library":
                    raise CnpcError("cnpc_shell_synth failed (status 2): injected", HIP_ERROR_OUT_OF_MEMORY)
                raise torch.OutOfMemoryError("injected")
            return torch.zeros((rows, n, n, n), dtype=torch.float64)

        def contract_pairs(self, prec, G, R1, R2, ia, ib):
            return np.array([float((G * R1[a] * R2[b]).sum()) for a, b in zip(ia, ib)])

    class FakeBuilder:
        method = "filter"

        def build(self, src, poly, bins, out):
            for r, b in enumerate(bins):
                out[r] = torch.from_numpy(np.random.default_rng(b).standard_normal(out.shape[1]))

    G = torch.from_numpy(np.random.default_rng(99).standard_normal(n ** 3))
    kinfo = (n, 100.0, 0, np.linspace(0.01, 0.07, k_bins + 1), k_bins)

    def run(ctx):
        return estimators._config_plan(ctx, FakeBuilder(), None, None, [(G, 1.0)], kinfo, 0, 0, 0, 0, True, False,
                                       "pairwise", "full", "full")()

    ref_ctx = FakeCtx(None)
    ref = run(ref_ctx)
    assert ref_ctx.rows == [k_bins]
    ctx = FakeCtx(2)
    got = run(ctx)
    assert ctx.rows[:2] == [6, 3] and set(ctx.rows[2:]) == {1}
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        run(FakeCtx(None, err=ValueError("not a memory error")))
    with pytest.raises((CnpcError, torch.OutOfMemoryError)):
        run(FakeCtx(0))


def test_planck18_parameters_and_missing_h():
    """The product's Planck18 numbers (Omega_gamma, Omega_de, the neutrino
    prefactor and y of the 0.06 eV species) agree with the oracle's FLRW
    restatement; without a config, or without 'h', the survey reader fails
    on para_cosmo["h"] as the reference does (TypeError / KeyError)."""
    from cosmonpc_amd.catalog import PLANCK18, flrw_params, read_survey_catalog
    from oracle.cnpc_ref import catalog as OC

    from helpers import NO_COMP, survey_rows

    p = PLANCK18
    cosmo, dh = flrw_params(p["H0"], p["Om0"], p["Tcmb0"], p["Neff"], p["m_nu"])
    og = OC.omega_gamma(p["H0"], p["Tcmb0"])
    np.testing.assert_allclose(cosmo[1], og, rtol=1e-12)
    onu = og * OC.nu_density_ratio(0.0, p["Tcmb0"], p["Neff"], p["m_nu"])
    np.testing.assert_allclose(cosmo[2], 1 - p["Om0"] - og - onu, rtol=1e-12)
    assert cosmo[4:6] == [2.0, 1.0] and cosmo[7:] == [0.0, 0.0]
    np.testing.assert_allclose(dh, 299792.458 / 67.66, rtol=1e-15)
    for z in (0.0, 0.7, 3.0):  # E(z) of the product's numbers == the oracle's
        q = 1 + z
        nu = cosmo[4] + (1 + (0.3173 * cosmo[6] / q) ** 1.83) ** 0.54644808743
        e = np.sqrt(cosmo[0] * q ** 3 + cosmo[1] * (1 + cosmo[3] * nu) * q ** 4 + cosmo[2])
        np.testing.assert_allclose(e, OC.efunc_flrw(z, **OC.P18), rtol=1e-13)
    d, _ = survey_rows(50, 10, 3)
    with pytest.raises(TypeError):
        read_survey_catalog(None, None, ["RA", "DEC", "Z"], None, NO_COMP, None, "data", rows=d)
    with pytest.raises(KeyError):
        read_survey_catalog(None, None, ["RA", "DEC", "Z"], None, NO_COMP, {"Omega0": 0.3}, "data", rows=d)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says
    so, before touching a device."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
