"""GPU parity: the HIP path (through libcnpc_hip.so) against the FP64 oracle
on identical seeded inputs.

Tolerances: FP64 path 1e-10 relative per bin (relative to max|B| for
bispectrum entries, which cross zero); FP32 + split-bf16 / fp16 path 5e-6
relative to max|B| per multipole (split-fp8 3e-5).
"""
import numpy as np
import pytest

from tests.helpers import bk_attrs, catalog, pk_attrs, rel_err, rel_to_max, weights

pytestmark = pytest.mark.gpu

WINDOWS = ["ngp", "cic", "tsc", "pcs"]


def _oracle():
    from oracle.cnpc_ref import angular, box, conventions

    return conventions, angular, box


def _log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("sampler", WINDOWS)
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n", [32, 36])
def test_paint_matches_oracle(gpu_ctx, sampler, prec, n):
    """n = 32: sorted LDS-tiled paint; n = 36: scatter (atomic) paint."""
    import torch

    C, _, _ = _oracle()
    _log(f"paint {sampler} prec={prec} n={n}")
    L = 500.0
    pos = catalog(5000, L, 3, ngrid=32)
    w = weights(len(pos), seed=4)
    ctx = gpu_ctx
    m0, m1 = ctx.mesh(n, prec), ctx.mesh(n, prec)
    ctx.paint(torch.from_numpy(pos.copy()).cuda(), torch.from_numpy(w).cuda(), n, L, sampler, True, prec, m0, m1)
    ref0 = C.paint(pos, w, n, L, sampler).real
    ref1 = C.paint(pos, w, n, L, sampler, shift=0.5).real
    tol = 1e-12 if prec == 0 else 2e-6
    assert np.max(np.abs(m0.cpu().numpy() - ref0)) <= tol * np.max(ref0)
    assert np.max(np.abs(m1.cpu().numpy() - ref1)) <= tol * np.max(ref1)
    assert abs(m0.sum().item() - w.sum()) <= 1e-9 * w.sum() * (1 if prec == 0 else 1e4)


def test_paint_wraps_and_empty(gpu_ctx):
    import torch

    C, _, _ = _oracle()
    n, L = 16, 100.0
    pos = np.array([[0.0, 0.0, 0.0], [99.99, 99.99, 99.99], [50.0, 0.01, 99.0]])
    w = np.array([1.0, 2.0, 3.0])
    ctx = gpu_ctx
    m0, m1 = ctx.mesh(n, 0), ctx.mesh(n, 0)
    ctx.paint(torch.from_numpy(pos).cuda(), torch.from_numpy(w).cuda(), n, L, "pcs", True, 0, m0, m1)
    np.testing.assert_allclose(m0.cpu().numpy(), C.paint(pos, w, n, L, "pcs").real, atol=1e-14)
    np.testing.assert_allclose(m1.cpu().numpy(), C.paint(pos, w, n, L, "pcs", shift=0.5).real, atol=1e-14)
    e = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    ctx.paint(e, torch.zeros(0, dtype=torch.float64, device="cuda"), n, L, "cic", False, 0, m0)
    assert m0.abs().max().item() == 0.0


@pytest.mark.parametrize("sampler", WINDOWS)
@pytest.mark.parametrize("n", [64, 40, 36])
def test_paint_tiled_dense_clump_and_batch(gpu_ctx, sampler, n):
    """Tiled paint (16^3 tiles at n = 64, 8^3 at n = 40; n = 36 takes the
    atomic scatter paths) on a dense clump
    inside one tile, particles on tile and box edges and a uniform
    background, against the oracle's paint: FP64 to 1e-12 of max, FP32 to
    5e-5 (the LDS additions of ~3000 weights per cell in another order).
    The batch entry (the same catalog between an empty and a small one, one
    set of launches) must give the same meshes, and the empty catalog zero
    meshes."""
    import torch

    C, _, _ = _oracle()
    L = 10.0 * n
    rng = np.random.default_rng(5)
    clump = 3.0 * n + rng.random((3000, 3)) * 4.0
    edges = np.array([[0.0, 0.0, 0.0], [L - 1e-3] * 3, [80.0, 160.0, L - 0.01], [79.999, 80.0, 0.0],
                      [159.999, 159.999, 159.999]])
    pos = np.concatenate([clump, edges, rng.random((20000, 3)) * L])
    w = 0.5 + rng.random(len(pos))
    small = rng.random((37, 3)) * L
    ws = 0.5 + rng.random(37)
    ref0 = C.paint(pos, w, n, L, sampler).real
    ref1 = C.paint(pos, w, n, L, sampler, shift=0.5).real
    refs0 = C.paint(small, ws, n, L, sampler).real
    refs1 = C.paint(small, ws, n, L, sampler, shift=0.5).real
    ctx = gpu_ctx
    pd, wd = torch.from_numpy(pos).cuda(), torch.from_numpy(w).cuda()
    bpos = torch.from_numpy(np.concatenate([pos, small])).cuda()
    bw = torch.from_numpy(np.concatenate([w, ws])).cuda()
    for prec in (0, 1):
        tol = 1e-12 if prec == 0 else 5e-5
        a0, a1 = ctx.mesh(n, prec), ctx.mesh(n, prec)
        ctx.paint(pd, wd, n, L, sampler, True, prec, a0, a1)
        b0, b1 = ctx.mesh(n, prec, 3), ctx.mesh(n, prec, 3)
        ctx.paint_batch(bpos, bw, [0, 0, len(pos), len(pos) + len(small)], n, L, sampler, True, prec, b0, b1)
        for got, exp in ((a0, ref0), (a1, ref1), (b0[1], ref0), (b1[1], ref1), (b0[2], refs0), (b1[2], refs1)):
            got = got.double().cpu().numpy()
            assert np.max(np.abs(got - exp)) <= tol * np.max(np.abs(exp)), (sampler, prec)
        assert b0[0].abs().max().item() == 0 and b1[0].abs().max().item() == 0


@pytest.mark.parametrize("sampler", ["cic", "pcs"])
def test_paint_tiled_bitwise_reproducible(gpu_ctx, sampler):
    """The tiled paint sums in 64-bit fixed point, so the mesh does not
    depend on the order in which particles are added: the catalog in a
    shuffled order (another order inside every sorted bucket, other
    workgroup interleavings) and the same catalog inside a batch give
    bit-identical FP64 and FP32 meshes, which match the oracle; negative
    weights and weights spanning 1e-6..1e3 included."""
    import torch

    C, _, _ = _oracle()
    n, L = 64, 640.0
    rng = np.random.default_rng(11)
    pos = np.concatenate([190.0 + rng.random((4000, 3)) * 30.0, rng.random((30000, 3)) * L])
    w = rng.choice([-1.0, 1.0], len(pos)) * 10.0 ** rng.uniform(-6, 3, len(pos))
    perm = rng.permutation(len(pos))
    ctx = gpu_ctx
    for prec in (0, 1):
        outs = []
        for order in (np.arange(len(pos)), perm):
            a0, a1 = ctx.mesh(n, prec), ctx.mesh(n, prec)
            ctx.paint(torch.from_numpy(pos[order].copy()).cuda(), torch.from_numpy(w[order].copy()).cuda(), n, L,
                      sampler, True, prec, a0, a1)
            outs.append((a0.cpu().numpy(), a1.cpu().numpy()))
        b0, b1 = ctx.mesh(n, prec, 2), ctx.mesh(n, prec, 2)
        bp = torch.from_numpy(np.concatenate([pos[perm], pos])).cuda()
        bw = torch.from_numpy(np.concatenate([w[perm], w])).cuda()
        ctx.paint_batch(bp, bw, [0, len(pos), 2 * len(pos)], n, L, sampler, True, prec, b0, b1)
        for got in (outs[1], (b0[0].cpu().numpy(), b1[0].cpu().numpy()), (b0[1].cpu().numpy(), b1[1].cpu().numpy())):
            assert np.array_equal(got[0], outs[0][0]) and np.array_equal(got[1], outs[0][1]), (sampler, prec)
        ref0 = C.paint(pos, w, n, L, sampler).real
        tol = 1e-12 if prec == 0 else 2e-6
        assert np.max(np.abs(outs[0][0] - ref0)) <= tol * np.max(np.abs(ref0))


def test_fft_roundtrip_and_normalisation(gpu_ctx):
    import torch

    n = 32
    rng = np.random.default_rng(0)
    x = rng.standard_normal((n, n, n))
    ctx = gpu_ctx
    X = torch.from_numpy(x).cuda()
    S = ctx.spectrum(n, 0)
    ctx.r2c(X, S, n, 0)
    ref = np.fft.rfftn(x)
    assert np.max(np.abs(S.cpu().numpy() - ref)) <= 1e-12 * np.max(np.abs(ref))
    Y = ctx.mesh(n, 0)
    ctx.c2r(S, Y, n, 0)
    np.testing.assert_allclose(Y.cpu().numpy() / n ** 3, x, atol=1e-12)


def test_c2r_400_fixed_plan(gpu_ctx):
    """The FP32 c2r of a 400^3 grid (the contraction grid of the 512^3
    workload) runs on the fixed 20 x 20 line plan in three passes: against
    numpy's FP64 irfftn (times n^3) of an arbitrary half spectrum, whose
    kz = 0 and n/2 imaginary parts both drop; and a round trip of a real
    field through the FP64 r2c.  A batch of two equals two single calls bit
    for bit.  An index or sign mistake gives O(1) errors; FP32 rounding of
    the two 20-point stages per axis stays near 1e-7 of max|out|."""
    import torch

    n = 400
    ctx = gpu_ctx
    rng = np.random.default_rng(3)
    s = (rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))).astype(np.complex64)
    ref = np.fft.irfftn(s.astype(np.complex128), s=(n, n, n)) * float(n) ** 3
    S = ctx.spectrum(n, 1)
    S.copy_(torch.from_numpy(s))
    Y = ctx.mesh(n, 1)
    ctx.c2r(S, Y, n, 1)
    got = Y.cpu().numpy()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    _log(f"c2r 400^3 FP32 fixed plan vs numpy irfftn: rel err {err:.2e}")
    assert err < 2e-6, err
    del ref
    x = rng.standard_normal((n, n, n))
    X = torch.from_numpy(x).cuda()
    D = ctx.spectrum(n, 0)
    ctx.r2c(X, D, n, 0)
    S2 = ctx.spectrum(n, 1, 2)
    S2[0].copy_(torch.from_numpy(s))
    S2[1].copy_(D)
    single = ctx.spectrum(n, 1)
    single.copy_(D)
    Y1 = ctx.mesh(n, 1)
    ctx.c2r(single, Y1, n, 1)
    rt = Y1.double().cpu().numpy() / n ** 3
    err_rt = np.max(np.abs(rt - x)) / np.max(np.abs(x))
    _log(f"c2r 400^3 FP32 fixed plan round trip: rel err {err_rt:.2e}")
    assert err_rt < 2e-6, err_rt
    Y2 = ctx.mesh(n, 1, 2)
    ctx.c2r(S2, Y2, n, 1, 2)
    assert torch.equal(Y2[0], Y) and torch.equal(Y2[1], Y1)


@pytest.mark.parametrize("sampler,interlaced", [("cic", True), ("tsc", False), ("pcs", True), ("ngp", False)])
def test_pk_box_matches_oracle(gpu_ctx, sampler, interlaced):
    from cosmonpc_amd.estimators import calculate_power_spectrum_box
    from cosmonpc_amd.mesh import paint_field

    _, _, B = _oracle()
    _log(f"pk {sampler} interlaced={interlaced}")
    n, L = 64, 1000.0
    pos = catalog(40000, L, 5, ngrid=64)
    w = np.ones(len(pos))
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, sampler, interlaced)
    sa = dict(attrs)
    sa.update(pk_attrs(n, L, 0.005, 0.185, 18, [0, 1, 2, 4], sampler, interlaced, rsd=(0, 0.6, 0.8)))
    ref = B.calculate_power_spectrum_box(fa, None, "auto", sa)
    f = paint_field(gpu_ctx, pos, w, n, L, sampler, interlaced, 0)
    got = calculate_power_spectrum_box(f, None, "auto", sa)
    np.testing.assert_array_equal(got["k_num"], ref["k_num"])
    assert rel_err(got["k_eff"], ref["k_eff"]) < 1e-12
    for p in (0, 2, 4):
        assert rel_err(got[f"P{p}"], ref[f"P{p}"]) < 1e-10, p
    assert got["P1"].dtype == np.float128 and not np.any(got["P1"])


def test_pk_box_cross_matches_oracle(gpu_ctx):
    from cosmonpc_amd.estimators import calculate_power_spectrum_box
    from cosmonpc_amd.mesh import paint_field

    _, _, B = _oracle()
    n, L = 48, 800.0
    pa = catalog(20000, L, 6, ngrid=48)
    pb = catalog(15000, L, 7, ngrid=48)
    wa, wb = weights(len(pa), 1), weights(len(pb), 2)
    attrs, fa, fb, _ = B.get_mesh_box({"a": (pa, wa), "b": (pb, wb)}, [n] * 3, [L] * 3, "tsc", True,
                                      correlation_mode="cross")
    sa = dict(attrs)
    sa.update(pk_attrs(n, L, 0.0, 0.18, 12, [0, 2], "tsc", True))
    ref = B.calculate_power_spectrum_box(fa, fb, "cross", sa)
    ga = paint_field(gpu_ctx, pa, wa, n, L, "tsc", True, 0)
    gb = paint_field(gpu_ctx, pb, wb, n, L, "tsc", True, 0)
    got = calculate_power_spectrum_box(ga, gb, "cross", sa)
    np.testing.assert_array_equal(got["k_num"], ref["k_num"])
    for p in (0, 2):
        assert rel_to_max(got[f"P{p}"], ref[f"P{p}"]) < 1e-10


def _bk_case(gpu_ctx, angu, dvm, contraction, prec=0, sampler="tsc", interlaced=True, n=32, L=300.0,
             k=(0.03, 0.15, 6), npart=30000, seed=11, block_size=None, shell_method="synth", zsynth="fused"):
    # run_task requires 2 k_max < k_Nyq (task_executor.py:192-196); the
    # half-spectrum path relies on it, so every case respects it
    assert 2 * k[1] < np.pi * n / L
    from cosmonpc_amd.estimators import calculate_bk_sugi_box
    from cosmonpc_amd.mesh import paint_field

    _, _, B = _oracle()
    _log(f"bk {angu} {dvm} {contraction} prec={prec} {sampler} n={n} {shell_method} {zsynth}")
    pos = catalog(npart, L, seed, ngrid=n)
    w = np.ones(len(pos))
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, sampler, interlaced)
    extra = {"shell_method": shell_method, "zsynth": zsynth}
    if block_size is not None:
        extra["block_size"] = block_size
    sa = dict(attrs)
    sa.update(bk_attrs(n, L, k[0], k[1], k[2], angu, dvm, sampler, interlaced, contraction=contraction, **extra))
    ref = B.calculate_bk_sugi_box(fa, None, None, "auto", sa)
    f = paint_field(gpu_ctx, pos, w, n, L, sampler, interlaced, prec)
    got = calculate_bk_sugi_box(f, None, None, "auto", sa)
    return got, ref


@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2), (1, 1, 0), (2, 2, 2)])
@pytest.mark.parametrize("contraction", ["gemm_fp64", "pairwise"])
def test_bk_box_full_fp64_matches_oracle(gpu_ctx, angu, contraction):
    got, ref = _bk_case(gpu_ctx, angu, "full", contraction)
    np.testing.assert_array_equal(got["nmodes"], ref["nmodes"])
    assert rel_to_max(got["Bk_raw"], ref["Bk_raw"]) < 1e-10
    assert rel_to_max(got["B_sugi"], ref["B_sugi"]) < 1e-10
    for key in ("SN0", "SN1", "SN2", "SN3"):
        a, b = np.asarray(got["SN_terms"][key]), np.asarray(ref["SN_terms"][key])
        if np.max(np.abs(b)) > 0:
            assert rel_to_max(a, b) < 1e-10, key
        else:
            assert np.max(np.abs(a)) == 0, key


@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2), (2, 2, 0)])
def test_bk_box_diagonal_fp64_matches_oracle(gpu_ctx, angu):
    got, ref = _bk_case(gpu_ctx, angu, "diagonal", "gemm_fp64", sampler="pcs")
    assert rel_to_max(got["B_sugi"], ref["B_sugi"]) < 1e-10
    assert rel_to_max(got["Bk_raw"], ref["Bk_raw"]) < 1e-10


@pytest.mark.parametrize("angu,dvm", [((0, 0, 0), "full"), ((2, 0, 2), "full"), ((1, 1, 0), "diagonal"),
                                      ((2, 2, 2), "full")])
def test_bk_box_s3_fft_matches_oracle(gpu_ctx, angu, dvm):
    from cosmonpc_amd.estimators import calculate_bk_sugi_box
    from cosmonpc_amd.mesh import paint_field

    _, _, B = _oracle()
    n, L = 32, 300.0
    pos = catalog(30000, L, 11, ngrid=n)
    w = np.ones(len(pos))
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, "cic", True)
    sa = dict(attrs)
    sa.update(bk_attrs(n, L, 0.03, 0.15, 6, angu, dvm, "cic", True, sn="both", contraction="gemm_fp64"))
    ref = B.calculate_bk_sugi_box(fa, None, None, "auto", sa)
    got = calculate_bk_sugi_box(paint_field(gpu_ctx, pos, w, n, L, "cic", True, 0), None, None, "auto", sa)
    for key in ("SN3", "SN3_FFT"):
        assert rel_to_max(got["SN_terms"][key], ref["SN_terms"][key]) < 1e-10, key
    assert rel_to_max(got["B_sugi"], ref["B_sugi"]) < 1e-10


@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2), (2, 2, 0)])
def test_bk_box_reduced_grid_matches_oracle(gpu_ctx, angu):
    """Contraction on the reduced grid (64^3 -> 36^3 here) is the same sum."""
    from cosmonpc_amd.estimators import LAST_TIMINGS, contraction_grid

    assert contraction_grid(64, 1000.0, 0.05, "gemm_fp64")[0] == 36
    got, ref = _bk_case(gpu_ctx, angu, "full", "gemm_fp64", n=64, L=1000.0, k=(0.01, 0.05, 4), npart=40000,
                        sampler="cic")
    assert LAST_TIMINGS["grid"] == 36
    assert rel_to_max(got["Bk_raw"], ref["Bk_raw"]) < 1e-10
    assert rel_to_max(got["B_sugi"], ref["B_sugi"]) < 1e-10


@pytest.mark.parametrize("dvm", ["full", "diagonal"])
@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2)])
def test_bk_box_s3_ana_bin_centres(gpu_ctx, angu, dvm):
    """``shotnoise_ana_all_eff=False``: the analytic S3 with q at the bin
    centres (clustering_estimator.py:1102-1126, :1137-1160) against the
    oracle.  For l > 0 it differs from the effective-k default (q_000 is the
    triangle indicator, the same at centres and effective k here)."""
    from cosmonpc_amd.estimators import calculate_bk_sugi_box
    from cosmonpc_amd.mesh import paint_field

    _, _, B = _oracle()
    n, L = 32, 300.0
    pos = catalog(30000, L, 13, ngrid=n)
    w = np.ones(len(pos))
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, "tsc", True)
    sa = dict(attrs)
    sa.update(bk_attrs(n, L, 0.03, 0.15, 6, angu, dvm, "tsc", True, contraction="gemm_fp64",
                       shotnoise_ana_all_eff=False))
    ref = B.calculate_bk_sugi_box(fa, None, None, "auto", sa)
    f = paint_field(gpu_ctx, pos, w, n, L, "tsc", True, 0)
    got = calculate_bk_sugi_box(f, None, None, "auto", sa)
    assert rel_to_max(got["SN_terms"]["SN3"], ref["SN_terms"]["SN3"]) < 1e-10
    assert rel_to_max(got["B_sugi"], ref["B_sugi"]) < 1e-10
    if angu != (0, 0, 0):
        eff = calculate_bk_sugi_box(f, None, None, "auto", dict(sa, shotnoise_ana_all_eff=True))
        assert rel_to_max(eff["SN_terms"]["SN3"], got["SN_terms"]["SN3"]) > 1e-6


def test_bk_box_reduced_grid_s3_fft(gpu_ctx):
    from cosmonpc_amd.estimators import calculate_bk_sugi_box
    from cosmonpc_amd.mesh import paint_field

    _, _, B = _oracle()
    n, L = 64, 1000.0
    pos = catalog(40000, L, 12, ngrid=n)
    w = np.ones(len(pos))
    attrs, fa, _, _ = B.get_mesh_box({"a": (pos, w)}, [n] * 3, [L] * 3, "tsc", True)
    sa = dict(attrs)
    sa.update(bk_attrs(n, L, 0.01, 0.05, 4, (0, 0, 0), "full", "tsc", True, sn="fft", contraction="gemm_fp64"))
    ref = B.calculate_bk_sugi_box(fa, None, None, "auto", sa)
    got = calculate_bk_sugi_box(paint_field(gpu_ctx, pos, w, n, L, "tsc", True, 0), None, None, "auto", sa)
    assert rel_to_max(got["SN_terms"]["SN3"], ref["SN_terms"]["SN3"]) < 1e-10
    assert rel_to_max(got["B_sugi"], ref["B_sugi"]) < 1e-10


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,L,kmax", [(32, 300.0, 0.15), (40, 500.0, 0.12), (64, 1000.0, 0.09), (36, 1000.0, 0.05)])
def test_shell_synth_matches_filter_c2r(gpu_ctx, prec, n, L, kmax):
    """Pruned synthesis (sparse x-DFT + 2-D c2r) == shell filter + 3-D c2r,
    for real, imaginary and conjugated Y_lm weights and a null spectrum."""
    import torch

    from cosmonpc_amd import angular as A

    ctx = gpu_ctx
    rng = np.random.default_rng(n)
    x = torch.from_numpy(rng.standard_normal((n, n, n)).astype(np.float64 if prec == 0 else np.float32)).cuda()
    D = ctx.spectrum(n, prec)
    ctx.r2c(x, D, n, prec)
    edges = np.linspace(0.01, kmax, 7)
    nb = 6
    work = ctx.spectrum(n, prec, nb)
    for ell, m, part in [(0, 0, A.PART_RE), (2, 1, A.PART_IM), (3, 0, A.PART_RE), (4, -3, A.PART_RE)]:
        poly = A.ylm_poly(ell, m, True, part, factor=(1j if ell % 2 else 1.0))
        for src in (D, None):
            a = ctx.mesh(n, prec, nb)
            b = ctx.mesh(n, prec, nb)
            ctx.shell_synth(src, n, L, prec, edges, 0, nb, poly, work, a)
            ctx.shell_filter(src, n, L, prec, edges, 0, nb, poly, work)
            ctx.c2r(work, b, n, prec, nb)
            ah, bh = a.cpu().numpy(), b.cpu().numpy()
            tol = 1e-12 if prec == 0 else 2e-5
            assert np.max(np.abs(ah - bh)) <= tol * np.max(np.abs(bh)), (ell, m, src is None)


@pytest.mark.parametrize("n,L,kmax", [(40, 500.0, 0.12), (48, 600.0, 0.13), (80, 1000.0, 0.15), (200, 1000.0, 0.3),
                                      (400, 2000.0, 0.31)])
def test_zgemm_fused_matches_fp64_contraction(gpu_ctx, n, L, kmax):
    """Fused z synthesis + split-bf16 GEMM (half-synthesised shells in slots)
    against the FP64 real shells contracted by the FP64 GEMM: B == A with an
    odd row count (a lone last line), B != A with different weights and row
    counts, slots beyond the rows, a null spectrum.  The bar is the accuracy
    of the unfused split-bf16 path on the same inputs (FP32 real shells, the
    GEMM on them): the fused error must stay within twice it (or 1e-6 of
    max |C|); an indexing mistake gives O(1) errors."""
    import torch

    from cosmonpc_amd import angular as A

    ctx = gpu_ctx
    rng = np.random.default_rng(n + 7)
    x64 = torch.from_numpy(rng.standard_normal((n, n, n))).cuda()
    D64, D32 = ctx.spectrum(n, 0), ctx.spectrum(n, 1)
    ctx.r2c(x64, D64, n, 0)
    ctx.r2c(x64.float().contiguous(), D32, n, 1)
    G64 = torch.from_numpy(1.0 + 0.5 * rng.standard_normal(n ** 3)).cuda()
    G32 = G64.float().contiguous()
    edges = np.linspace(0.01, kmax, 8)
    nbins = 7
    R = ctx.synth_radius(L, kmax)
    polys = {0: A.shell_components(0, 0)[0][0], 2: A.shell_components(2, 0)[0][0]}

    def real64(src, poly, rows):
        out = ctx.mesh(n, 0, rows)
        ctx.shell_synth(src, n, L, 0, edges, 0, rows, poly, ctx.spectrum(n, 0, rows), out)
        return out.view(rows, n ** 3)

    def real32(src, poly, rows):
        out = ctx.mesh(n, 1, rows)
        ctx.shell_synth(src, n, L, 1, edges, 0, rows, poly, ctx.spectrum(n, 1, rows), out)
        return out.view(rows, n ** 3)

    def half32(src, poly, rows, nslots):
        H = ctx.half_shells(n, R, nslots)
        # rows and slots the synthesis leaves unwritten hold NaN: a read of
        # any of them (e.g. before the kz row counts are visible to every
        # wave) poisons the result
        H.fill_(complex(float("nan"), float("nan")))
        work = ctx.spectrum(n, 1, 4)
        kz = np.zeros(rows, dtype=np.int32)
        for b0 in range(0, rows, 4):
            c = min(4, rows - b0)
            kz[b0:b0 + c] = ctx.shell_synth_half(src, n, L, edges, b0, c, poly, work, H, b0)
        assert np.all(np.diff(kz) >= 0) and kz[-1] <= R + 1
        return H, kz

    cases = [("same odd rows", 0, 7, None, None, 9), ("A != B", 2, 6, 0, 5, 6), ("null spectrum", 0, 5, None, None, 5)]
    for what, la, na, lb, nb, nslots in cases:
        src64, src32 = (None, None) if what == "null spectrum" else (D64, D32)
        A64, A32 = real64(src64, polys[la], na), real32(src32, polys[la], na)
        HA, kza = half32(src32, polys[la], na, nslots)
        if lb is None:
            ref = ctx.contract_gemm("gemm_fp64", G64, A64).cpu().numpy()
            sep = ctx.contract_gemm("bf16x3", G32, A32).cpu().numpy()
            got = ctx.contract_zgemm(n, L, kmax, G32, HA, na, kza).cpu().numpy()
        else:
            B64, B32 = real64(src64, polys[lb], nb), real32(src32, polys[lb], nb)
            HB, kzb = half32(src32, polys[lb], nb, nb + 1)
            ref = ctx.contract_gemm("gemm_fp64", G64, A64, B64).cpu().numpy()
            sep = ctx.contract_gemm("bf16x3", G32, A32, B32).cpu().numpy()
            got = ctx.contract_zgemm(n, L, kmax, G32, HA, na, kza, HB, nb, kzb).cpu().numpy()
            # the same launch with B^T G B from the same z transforms
            got2, gbb = ctx.contract_zgemm(n, L, kmax, G32, HA, na, kza, HB, nb, kzb, with_bb=True)
            ref_bb = ctx.contract_gemm("gemm_fp64", G64, B64).cpu().numpy()
            sep_bb = ctx.contract_gemm("bf16x3", G32, B32).cpu().numpy()
            for tag, g2, r2, s2 in (("with_bb C", got2.cpu().numpy(), ref, sep),
                                    ("with_bb C_bb", gbb.cpu().numpy(), ref_bb, sep_bb)):
                sc = np.max(np.abs(r2))
                e2, e2s = np.max(np.abs(g2 - r2)) / sc, np.max(np.abs(s2 - r2)) / sc
                _log(f"zgemm n={n} {what} {tag}: rel err fused {e2:.2e}, separate {e2s:.2e}")
                assert e2 <= max(2 * e2s, 1e-6), (what, tag, e2, e2s)
        scale = np.max(np.abs(ref))
        err, err_sep = np.max(np.abs(got - ref)) / scale, np.max(np.abs(sep - ref)) / scale
        _log(f"zgemm n={n} {what}: rel err fused {err:.2e}, separate {err_sep:.2e}")
        assert err <= max(2 * err_sep, 1e-6), (what, err, err_sep)


def test_bk_box_fused_and_separate_z_agree(gpu_ctx):
    """The estimator with the fused z pass equals the one with real shells in
    HBM (both split-bf16) to 1e-6 of max|B| and the FP64 oracle to 5e-6."""
    got_f, ref = _bk_case(gpu_ctx, (2, 0, 2), "full", "bf16x3", prec=1)
    got_s, _ = _bk_case(gpu_ctx, (2, 0, 2), "full", "bf16x3", prec=1, zsynth="separate")
    assert rel_to_max(got_f["Bk_raw"], got_s["Bk_raw"]) < 1e-6
    assert rel_to_max(got_f["Bk_raw"], ref["Bk_raw"]) < 5e-6


def test_bk_box_filter_method_matches_oracle(gpu_ctx):
    got, ref = _bk_case(gpu_ctx, (2, 0, 2), "full", "gemm_fp64", shell_method="filter")
    assert rel_to_max(got["Bk_raw"], ref["Bk_raw"]) < 1e-10


def test_bk_box_blocked_matches_full(gpu_ctx):
    got, ref = _bk_case(gpu_ctx, (0, 0, 0), "full", "pairwise", block_size=3)
    assert rel_to_max(got["Bk_raw"], ref["Bk_raw"]) < 1e-10


@pytest.mark.parametrize("angu", [(0, 0, 0), (2, 0, 2)])
@pytest.mark.parametrize("contraction", ["bf16x3", "fp16x3", "fp8x6"])
def test_bk_box_fp32_split_error(gpu_ctx, angu, contraction):
    """Split-precision contractions against the FP64 oracle, relative to
    max|B| (FP32 paint / FFT / shells included): 5e-6 for the bf16 / fp16
    splits (measured 2.7e-7 - 8.6e-7), 3e-5 for the fp8 split (three 4-bit
    pieces; measured 2e-6 - 6.6e-6)."""
    got, ref = _bk_case(gpu_ctx, angu, "full", contraction, prec=1, n=48, L=800.0, k=(0.01, 0.09, 8),
                        npart=60000)
    err = rel_to_max(got["Bk_raw"], ref["Bk_raw"])
    _log(f"{contraction} {angu} error rel max|B| = {err:.3e}")
    assert err < (3e-5 if contraction == "fp8x6" else 5e-6)


def test_gemm_modes_agree_with_pairwise(gpu_ctx):
    """GEMM tiles (incl. rows beyond 32 and ragged K) against the pairwise sum."""
    import torch

    rng = np.random.default_rng(5)
    ncell = 64 * 1000 + 64
    na, nb = 37, 5
    A = rng.standard_normal((na, ncell))
    B = rng.standard_normal((nb, ncell))
    G = rng.standard_normal(ncell)
    ref = (A * G) @ B.T
    ctx = gpu_ctx
    Ad, Bd, Gd = (torch.from_numpy(x).cuda() for x in (A, B, G))
    C64 = ctx.contract_gemm("gemm_fp64", Gd, Ad, Bd, nblk=7).cpu().numpy()
    assert np.max(np.abs(C64 - ref)) < 1e-10 * np.max(np.abs(ref))
    ia, ib = np.meshgrid(np.arange(na), np.arange(nb), indexing="ij")
    P = ctx.contract_pairs(0, Gd, Ad, Bd, ia.ravel(), ib.ravel()).reshape(na, nb)
    assert np.max(np.abs(P - ref)) < 1e-10 * np.max(np.abs(ref))
    Cs = ctx.contract_gemm("gemm_fp64", Gd, Ad, None, nblk=13).cpu().numpy()
    refs = (A * G) @ A.T
    assert np.max(np.abs(Cs - refs)) < 1e-10 * np.max(np.abs(refs))
    A32, B32, G32 = (torch.from_numpy(x.astype(np.float32)).cuda() for x in (A, B, G))
    ref32 = (A.astype(np.float32).astype(np.float64) * G.astype(np.float32)) @ B.astype(np.float32).T
    for mode in ("bf16x3", "fp16x3"):
        s = 1.0 if mode == "bf16x3" else 2.0 ** 10
        C = ctx.contract_gemm(mode, G32, A32, B32, sa=s, sb=s, nblk=9).cpu().numpy()
        err = np.max(np.abs(C - ref32)) / np.sqrt(ncell)
        assert err < 1e-3, (mode, err)
