"""Pins the oracle's box estimators: closed-form power spectra, a
brute-force triangle sum for the bispectrum signal, and Poisson-catalog
shot-noise cancellation."""
import numpy as np
import pytest

from oracle.cnpc_ref import angular as OA
from oracle.cnpc_ref import box as B
from oracle.cnpc_ref import conventions as C


def _lattice_catalog(n, L, amp, kvec):
    """One particle at every mesh point, weight 1 + amp cos(k.x): with NGP
    the painted mesh is exactly that weight field."""
    H = L / n
    g = np.arange(n) * H
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])
    w = 1.0 + amp * np.cos(pos @ np.asarray(kvec))
    return pos, w


@pytest.mark.parametrize("axis,mu", [(2, 1.0), (0, 0.0)])
def test_pk_plane_wave_closed_form(axis, mu):
    n, L, amp = 16, 160.0, 0.3
    kf = 2 * np.pi / L
    kvec = np.zeros(3)
    kvec[axis] = 3 * kf
    pos, w = _lattice_catalog(n, L, amp, kvec)
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, "ngp", False)
    k_edge_lo, k_edge_hi, nb = 0.5 * kf, 4.5 * kf, 4
    sa = dict(attrs)
    sa.update(dict(poles=[0, 2, 4], boxsize=[L] * 3, nmesh=[n] * 3, k_min=k_edge_lo, k_max=k_edge_hi, k_bins=nb,
                   rsd=[0, 0, 1], sampler="ngp", interlaced=False))
    res = B.calculate_power_spectrum_box(fa, None, "auto", sa)
    V = L ** 3
    NZ = w.sum() / V
    assert NZ == pytest.approx(n ** 3 / V)
    # mode counts by brute force over the integer lattice
    nn = C.int_axis(n)
    kk = np.sqrt(nn[:, None, None] ** 2 + nn[None, :, None] ** 2 + nn[None, None, :] ** 2) * kf
    edges = np.linspace(k_edge_lo, k_edge_hi, nb + 1)
    cnt = [np.sum((kk >= edges[i]) & (kk < edges[i + 1])) for i in range(nb)]
    np.testing.assert_array_equal(res["k_num"], cnt)
    b = 2  # |k| = 3 kf sits in bin [2.5, 3.5) kf
    knum = res["k_num"][b]
    raw = 2 * (amp / 2) ** 2 / knum * n ** 6 / (V * NZ ** 2)  # = amp^2/2 V / k_num
    expect0 = np.full(nb, -1 / NZ)
    expect0[b] += raw
    np.testing.assert_allclose(res["P0"], expect0, rtol=1e-12, atol=1e-9 / NZ)
    P2 = 0.5 * (3 * mu ** 2 - 1)
    P4 = (35 * mu ** 4 - 30 * mu ** 2 + 3) / 8
    assert res["P2"][b] == pytest.approx(5 * raw * P2, rel=1e-12, abs=1e-9 * raw)
    assert res["P4"][b] == pytest.approx(9 * raw * P4, rel=1e-12, abs=1e-9 * raw)
    assert np.all(np.abs(np.delete(res["P2"], b)) < 1e-9 * raw)


def _brute_signal(rfield, attrs, angu, k_edge, dvm):
    """sum_x G F_i F_j = N^3 sum_{k1 in i, k2 in j} Y1(k1) d(k1) Y2(k2) d(k2) d(-k1-k2)."""
    n = attrs["nmesh"][0]
    L = attrs["boxsize"][0]
    vpc = L ** 3 / n ** 3
    d = OA.compensate_sinc(C.circular_axes(n), C.r2c(rfield / vpc - attrs["NZ_a"]), attrs["sampler"])
    kgrid, knorm = OA.get_kgrid(C.k_axes(n, L))
    l1, l2, Lm = angu
    cfg, tj = OA.get_magnetic_configs_box(l1, l2, Lm)
    K = len(k_edge) - 1
    idx = np.array(np.unravel_index(np.arange(n ** 3), (n, n, n))).T
    kn = knorm.ravel()
    total = np.zeros((K, K), dtype=complex)
    for s, (m1, m2, _) in enumerate(cfg):
        y1 = OA.get_Ylm(l1, m1, True)(*kgrid).ravel()
        y2 = OA.get_Ylm(l2, m2, True)(*kgrid).ravel()
        sub = np.zeros((K, K), dtype=complex)
        for i in range(K):
            mi = np.where((kn >= k_edge[i]) & (kn < k_edge[i + 1]))[0]
            for j in range(K):
                if dvm == "diagonal" and i != j:
                    continue
                mj = np.where((kn >= k_edge[j]) & (kn < k_edge[j + 1]))[0]
                k3 = (-(idx[mi][:, None, :] + idx[mj][None, :, :])) % n
                g = d[k3[..., 0], k3[..., 1], k3[..., 2]]
                val = (y1[mi] * d.ravel()[mi])[:, None] * (y2[mj] * d.ravel()[mj])[None, :] * g
                sub[i, j] = tj[s] * n ** 3 * val.sum()
        total += sub if (m1, m2) == (0, 0) else 2 * np.real(sub)
    return total if dvm == "full" else np.diag(total)


@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2), (1, 1, 0), (2, 2, 0)])
@pytest.mark.parametrize("dvm", ["full", "diagonal"])
def test_bk_signal_brute_force(angu, dvm):
    from tests.helpers import bk_attrs, catalog

    n, L = 12, 300.0
    pos = catalog(3000, L, 9, ngrid=12)
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, np.ones(len(pos)))}, [n] * 3, [L] * 3, "cic", True)
    sa = dict(attrs)
    sa.update(bk_attrs(n, L, 0.02, 0.06, 3, angu, dvm, "cic", True))
    got = B.calculate_bk_sugi_box(fa, None, None, "auto", sa, return_signal_only=True)
    exp = _brute_signal(fa, sa, angu, np.linspace(0.02, 0.06, 4), dvm)
    assert np.max(np.abs(got - exp)) < 1e-10 * np.max(np.abs(exp))


@pytest.mark.slow
def test_poisson_catalog_shot_noise_cancels():
    """A uniform Poisson catalog has zero power and zero bispectrum: the
    shot-noise terms must cancel the raw estimates (pins the relative
    normalisation of P_shot, SN0..SN3 and the signal)."""
    from tests.helpers import bk_attrs, pk_attrs

    rng = np.random.default_rng(1)
    n, L, N = 48, 1000.0, 200000
    pos = rng.uniform(0, L, (N, 3))
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, np.ones(N))}, [n] * 3, [L] * 3, "tsc", True)
    sa = dict(attrs)
    sa.update(pk_attrs(n, L, 0.02, 0.14, 6, [0], "tsc", True))
    pk = B.calculate_power_spectrum_box(fa, None, "auto", sa)
    # a Gaussian estimate with k_num/2 independent modes scatters by P/sqrt(k_num/2)
    sigma = pk["P_shot"] / np.sqrt(pk["k_num"] / 2)
    assert np.all(np.abs(pk["P0"]) < 4 * sigma)
    # without the shot-noise subtraction P0 would sit at P_shot, ~10-100 sigma away
    assert np.all(pk["P_shot"] > 10 * sigma)
    sa.update(bk_attrs(n, L, 0.02, 0.14, 6, (0, 0, 0), "full", "tsc", True))
    bk = B.calculate_bk_sugi_box(fa, None, None, "auto", sa)
    raw = np.real(bk["Bk_raw"])
    est = np.real(bk["B_sugi"])
    # The reference's shot-noise model is approximate (analytic S3 at
    # effective k, SN1/SN2 read on k3 bins that start at 2 k_min), so only
    # most of the raw signal cancels: ~0.1-0.2 residual here, 1.0 without it.
    assert np.mean(np.abs(est)) < 0.25 * np.mean(np.abs(raw))
    assert np.mean(np.abs(raw - np.real(bk["Shot_noise"]))) < 0.25 * np.mean(np.abs(raw))
