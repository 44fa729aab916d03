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


def _brute_s3_fft(ca, sa, angu, k_edge, dvm, SN0):
    """S3 of shotnoise_mode='fft' as a triangle sum: by the convolution
    theorem sum_x c2r(S) f_i f_j = N^3 sum_{k1 in i, k2 in j} conj Y1(k1)
    conj Y2(k2) S(-k1-k2), with S = V^2 |delta|^2 - SN0 C(k); divided by
    k_num_i k_num_j and scaled by H N_ells / N^3."""
    n = sa["nmesh"][0]
    L = sa["boxsize"][0]
    V = float(L) ** 3
    S = V ** 2 * ca * np.conj(ca) - OA.compensate_bk_noise(C.circular_axes(n), np.full(ca.shape, SN0 + 0j),
                                                           sa["sampler"])
    kgrid, knorm = OA.get_kgrid(C.k_axes(n, L))
    l1, l2, Lm = angu
    cfg, tj = OA.get_magnetic_configs_box(l1, l2, Lm)
    K = len(k_edge) - 1
    idx = np.array(np.unravel_index(np.arange(n ** 3), (n, n, n))).T
    kn = knorm.ravel()
    total = np.zeros((K, K), dtype=complex)
    for s, (m1, m2, _) in enumerate(cfg):
        y1 = np.conj(OA.get_Ylm(l1, m1, True)(*kgrid)).ravel()
        y2 = np.conj(OA.get_Ylm(l2, m2, True)(*kgrid)).ravel()
        sub = np.zeros((K, K), dtype=complex)
        for i in range(K):
            mi = np.where((kn >= k_edge[i]) & (kn < k_edge[i + 1]))[0]
            for j in range(K):
                if dvm == "diagonal" and i != j:
                    continue
                mj = np.where((kn >= k_edge[j]) & (kn < k_edge[j + 1]))[0]
                k3 = (-(idx[mi][:, None, :] + idx[mj][None, :, :])) % n
                sub[i, j] = tj[s] * n ** 3 * (y1[mi][:, None] * y2[mj][None, :] * S[k3[..., 0], k3[..., 1], k3[..., 2]]
                                              ).sum() / (len(mi) * len(mj))
        total += sub if (m1, m2) == (0, 0) else 2 * np.real(sub)
    N_ells = (2 * l1 + 1) * (2 * l2 + 1) * (2 * Lm + 1)
    total = total * OA.wigner_3j(l1, l2, Lm, 0, 0, 0) * N_ells / n ** 3
    return total if dvm == "full" else np.diag(total)


@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2), (1, 1, 0), (2, 2, 0)])
@pytest.mark.parametrize("dvm", ["full", "diagonal"])
def test_bk_s3_fft_brute_force(angu, dvm):
    """The FFT form of S3 (``clustering_estimator.py:1182-1384``) equals its
    triangle sum: pins the shell fields conj Y(k) mask, the source field, the
    conjugate-symmetry shortcuts for m2 = +-m1, the 2 Re rule over magnetic
    configurations and the H N / N^3 / (k_num_i k_num_j) scale."""
    from tests.helpers import bk_attrs, catalog

    n, L = 12, 300.0
    pos = catalog(3000, L, 9, ngrid=12)
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, np.ones(len(pos)))}, [n] * 3, [L] * 3, "cic", True)
    sa = dict(attrs)
    sa.update(bk_attrs(n, L, 0.02, 0.06, 3, angu, dvm, "cic", True, sn="fft"))
    res = B.calculate_bk_sugi_box(fa, None, None, "auto", sa)
    got = res["SN_terms"]["SN3"] * res["I_norm"]
    ca = B._bk_prepare(fa, sa["NZ_a"], L ** 3 / n ** 3, "cic", n)
    SN0 = attrs["N_gal_a"] if tuple(angu) == (0, 0, 0) else 0.0
    exp = _brute_s3_fft(ca, sa, angu, np.linspace(0.02, 0.06, 4), dvm, SN0)
    assert np.max(np.abs(got - exp)) < 1e-10 * np.max(np.abs(exp))


def _wave_catalog(rng, nmean, L, k0, amp):
    """Poisson sample (by thinning) of the intensity nmean/V (1 + amp cos(k0
    z + phase)) with a random phase."""
    phase = rng.uniform(0, 2 * np.pi)
    ntot = rng.poisson(nmean * (1 + amp))
    p = rng.uniform(0, L, (ntot, 3))
    keep = rng.uniform(0, 1 + amp, ntot) < 1 + amp * np.cos(k0 * p[:, 2] + phase)
    return p[keep]


def _wave_ensemble(angu, nreal):
    """B_raw, B_sugi and the shot-noise terms (x I_norm) of nreal Poisson
    samples of a single plane wave along z (32^3, TSC interlaced, 5 bins of
    1.5 k_f from 0.01 k_f, the wave at 4 k_f in bin 2, FFT S3)."""
    from tests.helpers import bk_attrs

    n, L = 32, 1000.0
    kf = 2 * np.pi / L
    kmin, dk, K = 0.01 * kf, 1.5 * kf, 5
    rng = np.random.default_rng(3)
    out = {"raw": [], "est": [], "SN1": [], "SN3": []}
    for _ in range(nreal):
        pos = _wave_catalog(rng, 200000, L, 4 * kf, 0.9)
        attrs, fa, _, _ = B.get_mesh_box({"a": (pos, np.ones(len(pos)))}, [n] * 3, [L] * 3, "tsc", True)
        sa = dict(attrs)
        sa.update(bk_attrs(n, L, kmin, kmin + K * dk, K, angu, "full", "tsc", True, sn="fft"))
        res = B.calculate_bk_sugi_box(fa, None, None, "auto", sa)
        out["raw"].append(np.real(res["Bk_raw"]))
        out["est"].append(np.real(res["B_sugi"]))
        out["SN1"].append(np.real(res["SN_terms"]["SN1"]))
        out["SN3"].append(np.real(res["SN_terms"]["SN3"]))
    return {k: np.array(v) for k, v in out.items()}


@pytest.mark.slow
@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2)])
def test_bk_shot_noise_single_wave_ensemble(angu):
    """Poisson samples of a fixed intensity 1 + A cos(k0 z): the intensity
    has no bispectrum (no three of +-k0 sum to zero), so by the Poisson
    moments E[rho1 rho2 rho3] = sum of the pair terms |Lambda(k_a)|^2 plus the
    N_gal term, and the estimator's mean must be the shot-noise model: mean
    B_sugi = 0 within the ensemble's standard error.  The P(k1) term lives in
    the k0 bin's row only and is (2L+1) times the Legendre-weighted power, so
    a wrong (2L+1) factor in SN1, or SN1 read one k3 bin off, leaves a
    residual of tens of standard errors (checked below by planting both
    mistakes in the oracle's outputs).  The k3 bins start at 2 k_min; with
    k_min = 0.01 k_f they hold the same modes as the k bins."""
    nreal = 40
    e = _wave_ensemble(angu, nreal)
    se = e["est"].std(0, ddof=1) / np.sqrt(nreal)
    scale = np.max(np.abs(e["raw"].mean(0)))
    z = np.abs(e["est"].mean(0)) / (se + 1e-3 * scale)
    assert np.all(z < 4.5), z
    row = 2  # the bin of |k0| = 4 k_f
    assert np.all(np.abs(e["raw"].mean(0)[row]) / se[row] > 5)
    ell = angu[2]

    def zmax(est):  # in standard errors of the estimator itself
        return np.max(np.abs(est.mean(0)) / (se + 1e-3 * scale))

    # planted: SN1 without its (2L+1) (for L = 0: SN1 and SN2 both dropped)
    wrong = e["est"] + e["SN1"] * (2 * ell / (2 * ell + 1) if ell else 2.0)
    assert zmax(wrong) > 10
    # planted: SN1 read one k3 bin higher
    shifted = e["est"] + e["SN1"] - np.roll(e["SN1"], -1, axis=1)  # axis 1: the k1 rows
    assert zmax(shifted) > 10


@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2)])
@pytest.mark.parametrize("dvm", ["full", "diagonal"])
def test_bk_s3_ana_model_smooth_spectrum(angu, dvm, monkeypatch):
    """The analytic S3 (``clustering_estimator.py:1081-1180``: q_l(k1, k2,
    k3) at effective k times the binned Legendre-weighted power, over the
    triangle-allowed k3 bins) is a continuum model of the exact k3 sum that
    the FFT form computes (pinned above).  For a field whose power is a
    smooth P(k, mu) = (1 + 0.8 L_2(mu)) k / (10 k_f) with exact amplitudes
    (random phases), the two agree to 10% of max |S3| (measured: 3-8%) on
    the bins away from the few-mode lowest shells.  A q_l at the wrong k3
    bin (planted in the oracle) misses by 35-48%."""
    from tests.helpers import bk_attrs

    n, L = 64, 1000.0
    kf = 2 * np.pi / L
    rng = np.random.default_rng(1)
    w = np.fft.fftn(rng.standard_normal((n, n, n)))
    w = w / np.abs(w)
    nn = np.fft.fftfreq(n, 1.0 / n)
    KX, KY, KZ = np.meshgrid(nn, nn, nn, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        mu = np.where(kk > 0, KZ / kk, 0.0)
    P = (1.0 + 0.8 * 0.5 * (3 * mu ** 2 - 1)) * kk / 10.0
    V = L ** 3
    dk = w * np.sqrt(P) / V
    dk[0, 0, 0] = 0.0
    NZ = 1e-3
    rfield = (np.real(np.fft.ifftn(dk) * n ** 3) + NZ) * V / n ** 3
    sa = {"N_gal_a": 1.0, "NZ_a": NZ}
    sa.update(bk_attrs(n, L, 0.01 * kf, 18.01 * kf, 6, angu, dvm, "ngp", False, sn="both"))

    def s3(res):
        a, f = res["SN_terms"]["SN3"], res["SN_terms"]["SN3_FFT"]
        sl = (slice(2, None), slice(2, None)) if dvm == "full" else slice(2, None)
        return np.real(a)[sl], np.real(f)[sl]

    a, f = s3(B.calculate_bk_sugi_box(rfield, None, None, "auto", sa))
    assert np.max(np.abs(a - f)) < 0.10 * np.max(np.abs(f))
    orig = OA.get_q_ells_eff

    def shifted(i, j, k_eff, l1, l2, Lm, nb, k3_eff):
        return np.roll(orig(i, j, k_eff, l1, l2, Lm, nb, k3_eff), 1)

    monkeypatch.setattr(OA, "get_q_ells_eff", shifted)
    a2, f2 = s3(B.calculate_bk_sugi_box(rfield, None, None, "auto", sa))
    assert np.max(np.abs(a2 - f2)) > 0.25 * np.max(np.abs(f2))
