"""Batched mock driver on the GPU: every output of run_mocks equals what
run_stats writes for the same catalog alone (FP64, 1e-10 relative to max|.|
per entry: the two paths differ only in the order of the atomic additions),
and a second run resumes by skipping the finished mocks."""
import os

import numpy as np
import pytest

from tests.helpers import catalog, rel_to_max, weights

pytestmark = pytest.mark.gpu


def _cfgs(out):
    base = dict(correlation_mode="auto", geometry="box-like", column_names=["x", "y", "z"], rsd=[0, 0, 1],
                apply_rsd=False, nmesh=[32] * 3, boxsize=[500.0] * 3, sampler="cic", interlaced=True,
                compensation=True, cosmology={"h": 0.6766, "Omega0": 0.30966}, redshift_box=0.5,
                output_dir=str(out))
    pk = dict(base, statistic="pk", k_min=0.01, k_max=0.19, k_bins=9, poles=[0, 2, 4])
    bk = dict(base, statistic="bk_sugi", k_min=0.02, k_max=0.1, k_bins=4, angu_config=[2, 0, 2],
              tracer_type="aaa", data_vector_mode="full", shotnoise_mode="both")
    return pk, bk


def _arrays(tmp_path):
    arrays = {}
    for i, npart in enumerate((3000, 700, 0, 5000, 1200)):
        path = str(tmp_path / "mocks" / f"mock_{i}.npy")
        pos = catalog(max(npart, 1), 500.0, 20 + i, ngrid=32)[:npart]
        arrays[path] = (pos, weights(npart, seed=30 + i))
    return arrays


def _assert_close(got, want, what):
    """Result entries agree: floating arrays to 1e-10 of their max |.|,
    floats to 1e-12, containers recursively, everything else exactly."""
    if isinstance(want, dict):
        assert sorted(got) == sorted(want), what
        for k in want:
            _assert_close(got[k], want[k], f"{what}.{k}")
    elif isinstance(want, (list, tuple)):
        assert len(got) == len(want), what
        for i, (g, w) in enumerate(zip(got, want)):
            _assert_close(g, w, f"{what}[{i}]")
    elif isinstance(want, np.ndarray) and want.dtype.kind in "fc":
        assert got.shape == want.shape, what
        if np.any(want != 0):
            assert rel_to_max(got, want) < 1e-10, what
        else:
            np.testing.assert_array_equal(got, want, err_msg=what)
    elif isinstance(want, (float, np.floating)):
        assert got == pytest.approx(want, rel=1e-12, nan_ok=True), what
    else:
        np.testing.assert_equal(got, want, err_msg=what)


def test_run_mocks_matches_run_stats(gpu_ctx, tmp_path):
    from cosmonpc_amd import run_stats
    from cosmonpc_amd.mocks import LAST_TIMINGS, run_mocks

    arrays = _arrays(tmp_path)
    paths = [p for p in arrays if len(arrays[p][1]) > 0]
    pk, bk = _cfgs(tmp_path / "batched")
    print(f"run_mocks over {len(paths)} mocks, batch 2", flush=True)
    got = run_mocks([pk, bk], paths, batch=2, arrays=arrays)
    assert LAST_TIMINGS["mocks"] == len(paths) and set(got) == set(paths)
    for path in paths:
        print(f"run_stats {os.path.basename(path)}", flush=True)
        for cfg, res in zip(_cfgs(tmp_path / "single"), got[path]):
            cfg = dict(cfg, catalogs={"data_a": path}, arrays={"data_a": arrays[path]})
            ref = run_stats(cfg)
            assert sorted(ref) == sorted(res)
            for key, val in ref.items():
                if key != "output_dir":
                    _assert_close(res[key], val, f"{os.path.basename(path)} {cfg['statistic']} {key}")
    names = sorted(os.listdir(tmp_path / "batched"))
    assert names == sorted(os.listdir(tmp_path / "single"))
    again = run_mocks([pk, bk], paths, batch=2, arrays=arrays)
    assert again == {} and LAST_TIMINGS["mocks"] == 0


def test_run_mocks_empty_catalog(gpu_ctx, tmp_path):
    """A mock with no particles paints an empty mesh inside a batch; its
    neighbours are unaffected (checked against a batch without it)."""
    from cosmonpc_amd.mocks import run_mocks

    arrays = _arrays(tmp_path)
    paths = list(arrays)
    pk, _ = _cfgs(tmp_path / "with_empty")
    with np.errstate(divide="ignore", invalid="ignore"):
        got = run_mocks([pk], paths, batch=len(paths), arrays=arrays)
    pk2, _ = _cfgs(tmp_path / "without")
    alone = run_mocks([pk2], [paths[3]], arrays=arrays)
    assert rel_to_max(got[paths[3]][0]["P0"], alone[paths[3]][0]["P0"]) < 1e-10
    assert got[paths[2]][0]["N_gal_a"] == 0


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
def test_run_mocks_shared_bk_signal(gpu_ctx, tmp_path, precision, tol):
    """bk_sugi [0,0,0] and [2,0,2] of one catalog in one run_mocks call share
    one multi-configuration signal pass (estimators.bk_box_signals); each
    result equals run_stats of that configuration alone: FP64 to 1e-10 (the
    same contractions), split-bf16 to 1e-5 of max|B| (FP32 summation order;
    its error against FP64 is ~1e-6)."""
    from cosmonpc_amd import run_stats
    from cosmonpc_amd.estimators import signal_group_key
    from cosmonpc_amd.mocks import run_mocks

    arrays = _arrays(tmp_path)
    path = [p for p in arrays if len(arrays[p][1]) > 0][0]
    _, bk = _cfgs(tmp_path / "batched")
    cfgs = [dict(bk, angu_config=a, precision=precision) for a in ([0, 0, 0], [2, 0, 2])]
    assert signal_group_key(cfgs[0]) == signal_group_key(cfgs[1]) is not None
    print(f"shared bk signal {precision}", flush=True)
    got = run_mocks(cfgs, [path], arrays=arrays)[path]
    for cfg, res in zip(cfgs, got):
        ref = run_stats(dict(cfg, output_dir=str(tmp_path / "single"), catalogs={"data_a": path},
                             arrays={"data_a": arrays[path]}))
        for key in ("Bk_raw", "B_sugi"):
            err = rel_to_max(res[key], ref[key])
            print(f"{cfg['angu_config']} {key} rel {err:.2e}", flush=True)
            assert err < tol, (cfg["angu_config"], key, err)
        np.testing.assert_array_equal(res["nmodes"], ref["nmodes"])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
def test_run_mocks_graph_replay(gpu_ctx, tmp_path, precision, tol):
    """The graph-recorded chunk pipeline (pk [0,2,4] + bk_sugi [0,0,0] and
    [2,0,2] diagonal with the analytic shot noise): chunks of 2 mocks, so the
    first chunk is measured by the warm-up before the recording, the second
    (a 5000-particle mock past the recorded capacity, beside an empty one)
    is recorded again, and the third (partial) is a replay.  Every output
    equals the eager driver's (graph=False): FP64 to 1e-10, FP32 to 1e-5 of
    max|.| (split-bf16 summation order); the empty mock's pk is NaN-free."""
    from cosmonpc_amd.mocks import LAST_TIMINGS, run_mocks

    arrays = _arrays(tmp_path)
    paths = list(arrays)
    pk, bk = _cfgs(tmp_path / "g")
    bk = dict(bk, data_vector_mode="diagonal", shotnoise_mode="ana", precision=precision)
    pk = dict(pk, precision=precision)
    cfgs = [pk] + [dict(bk, angu_config=a) for a in ([0, 0, 0], [2, 0, 2])]
    print(f"graph mocks {precision}", flush=True)
    with np.errstate(divide="ignore", invalid="ignore"):
        got = run_mocks(cfgs, paths, batch=2, arrays=arrays)
        assert LAST_TIMINGS["graph"] and LAST_TIMINGS["replayed_chunks_s"] > 0
        ref = run_mocks([dict(c, output_dir=str(tmp_path / "e")) for c in cfgs], paths, batch=2, arrays=arrays,
                        graph=False)
    assert not LAST_TIMINGS["graph"]
    for path in paths:
        for g, r in zip(got[path], ref[path]):
            for key in ("P0", "P2", "P4", "B_sugi", "Bk_raw", "Shot_noise", "k_eff", "nmodes", "I_norm"):
                if key not in r:
                    continue
                a, b = np.asarray(g[key]), np.asarray(r[key])
                if not np.all(np.isfinite(b)):
                    np.testing.assert_array_equal(np.isfinite(a), np.isfinite(b))
                    a, b = np.where(np.isfinite(b), a, 0), np.where(np.isfinite(b), b, 0)
                if np.any(b != 0):
                    assert rel_to_max(a, b) < tol, (os.path.basename(path), key)
                else:
                    assert np.all(a == 0), (os.path.basename(path), key)
