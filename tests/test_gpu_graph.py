"""hipGraph recording of the signal step (``cosmonpc_amd.graphs.StepGraph``).

Every replay is checked against the eager launches on the same input, with
eager work on the same context and stream between the replays, a replay
after the context's scratch area grew under the recording (the hold keeps
the recorded address alive), and a replay on a new input (so a recording
that did not recompute would fail).  The recorded step is the box signal
stage of the fused split-bf16 path: G_00 through rocFFT, the shell x/y
synthesis kernels and the fused contractions of two angular configurations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, L, EDGES = 48, 800.0, np.linspace(0.01, 0.09, 9)
ANGUS = [(0, 0, 0), (2, 0, 2)]


def _log(msg):
    print(msg, flush=True)


def _spectrum(ctx, seed):
    import torch

    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.standard_normal((N, N, N)).astype(np.float32)).cuda()
    D = ctx.spectrum(N, 1)
    ctx.r2c(x, D, N, 1)
    return D


def _enqueue(ctx, src, shared=True):
    """G_00 and the signal matrices of both configurations; ``shared``: one
    multi-configuration call (the unweighted shells synthesised once)."""
    from cosmonpc_amd.estimators import bk_box_g00, bk_box_signal_auto, bk_box_signal_multi

    G = bk_box_g00(ctx, src, N, 1)
    if shared:
        return bk_box_signal_multi(ctx, src, G, N, L, 1, EDGES, ANGUS, "bf16x3", defer=True)
    fins = [bk_box_signal_auto(ctx, src, G, N, L, 1, EDGES, a, "bf16x3", defer=True) for a in ANGUS]
    return lambda: [f() for f in fins]


def _eager(ctx, X, shared=True):
    src = X.clone()
    return _enqueue(ctx, src, shared)()


def test_shared_legs_equal_configurations_apart(gpu_ctx):
    """The multi-configuration call synthesises [0,0,0]'s shells once, reuses
    them as [2,0,2]'s second leg and takes [0,0,0]'s matrix from [2,0,2]'s
    contraction (B^T G B of the same z transforms): the results are the
    per-configuration ones up to the FP32 summation order of the split-bf16
    contraction (1e-5 of max |S|; the split error itself is ~1e-6)."""
    ctx = gpu_ctx
    X = _spectrum(ctx, 7)
    apart = _eager(ctx, X, shared=False)
    shared = _eager(ctx, X, shared=True)
    for a, b in zip(apart, shared):
        assert np.max(np.abs(a)) > 0
        err = np.max(np.abs(a - b)) / np.max(np.abs(a))
        _log(f"shared vs apart rel max diff {err:.2e}")
        assert err < 1e-5


def test_multi_configuration_pass_without_shared_legs(gpu_ctx):
    """Configurations with complex shell components and several magnetic
    configurations ([1,1,0], [2,2,2], [2,0,2] beside [0,0,0]) in one
    multi-configuration call equal the calls one at a time (1e-5 of max |S|:
    only the FP32 summation order of the contraction may differ)."""
    from cosmonpc_amd.estimators import bk_box_g00, bk_box_signal_auto, bk_box_signal_multi

    ctx = gpu_ctx
    X = _spectrum(ctx, 9)
    angus = [(1, 1, 0), (0, 0, 0), (2, 2, 2), (2, 0, 2)]
    src = X.clone()
    G = bk_box_g00(ctx, src, N, 1)
    multi = bk_box_signal_multi(ctx, src, G, N, L, 1, EDGES, angus, "bf16x3")
    for a, got in zip(angus, multi):
        want = bk_box_signal_auto(ctx, src, G, N, L, 1, EDGES, a, "bf16x3")
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        _log(f"multi {a}: rel max diff {err:.2e}")
        assert np.max(np.abs(want)) > 0 and err < 1e-5


def test_g00_from_consumed_spectrum(gpu_ctx):
    """G made inside the multi-configuration call from a spectrum the caller
    gives up (the inverse FFT runs on it after the shells have read it): the
    same matrices, bit for bit, as with G_00 made beforehand from a copy;
    configurations that do not all take the fused contraction ([1,1,0]'s
    complex components beside [0,0,0]) fall back to the copy and still
    match."""
    from cosmonpc_amd.estimators import bk_box_g00, bk_box_signal_multi

    ctx = gpu_ctx
    X = _spectrum(ctx, 11)
    for angus in (ANGUS, [(1, 1, 0), (0, 0, 0)]):
        src = X.clone()
        G = bk_box_g00(ctx, src, N, 1)
        want = bk_box_signal_multi(ctx, src, G, N, L, 1, EDGES, angus, "bf16x3")
        src2 = X.clone()
        got = bk_box_signal_multi(ctx, src2, None, N, L, 1, EDGES, angus, "bf16x3", consume_delta=True)
        for a, g, w in zip(angus, got, want):
            _log(f"G late {a}: max |diff| {np.max(np.abs(g - w)):.3e}")
            assert np.max(np.abs(w)) > 0 and np.array_equal(g, w), a


def test_graph_replays_match_eager(gpu_ctx):
    import torch

    from cosmonpc_amd.graphs import StepGraph

    ctx = gpu_ctx
    assert ctx.zgemm_supported(N, 8, 8, False), "the case must take the fused contraction"
    inputs = [_spectrum(ctx, s) for s in (3, 4, 5)]
    src = ctx.spectrum(N, 1)
    src.copy_(inputs[0])
    sg = StepGraph(ctx)
    _log(f"graph record n={N} angus={ANGUS}")
    fins = sg.record(lambda: _enqueue(ctx, src))
    try:
        for r, k in enumerate([0, 1, 2, 0]):
            src.copy_(inputs[k])
            _log(f"replay {r} input {k}")
            sg.replay()
            got = fins()
            want = _eager(ctx, inputs[k])  # eager work between replays
            for g, w in zip(got, want):
                assert np.max(np.abs(w)) > 0
                assert np.max(np.abs(g - w)) <= 1e-12 * np.max(np.abs(w)), f"replay {r} differs from eager"
            if r == 1:
                # the scratch area grows while the recording holds the old one
                ctx.reserve(256 << 20)
                torch.cuda.synchronize()
    finally:
        sg.close()
    # after close the context works as before
    got = _eager(ctx, inputs[1])
    assert np.all(np.isfinite(got[0]))


def test_graph_refuses_host_readback(gpu_ctx):
    """A call that reads back to the host fails cleanly while recording (the
    library checks the capture status before launching anything)."""
    import torch

    from cosmonpc_amd._lib import CnpcError
    from cosmonpc_amd.graphs import StepGraph

    ctx = gpu_ctx
    x = torch.full((4096,), -2.5, dtype=torch.float32, device=ctx.device)
    assert ctx.absmax(x) == 2.5
    sg = StepGraph(ctx)
    with pytest.raises(CnpcError, match="captured into a graph"):
        sg.record(lambda: ctx.absmax(x), warmup=False)
    sg.close()
    assert ctx.absmax(x) == 2.5


def test_multi_single_configuration_is_the_plain_call(gpu_ctx):
    """One configuration through the multi-configuration call is the plain
    call, bit for bit (nothing to share, the same launches)."""
    from cosmonpc_amd.estimators import bk_box_g00, bk_box_signal_auto, bk_box_signal_multi

    ctx = gpu_ctx
    src = _spectrum(ctx, 11)
    G = bk_box_g00(ctx, src, N, 1)
    for a in ANGUS:
        got = bk_box_signal_multi(ctx, src, G, N, L, 1, EDGES, [a], "bf16x3")[0]
        want = bk_box_signal_auto(ctx, src, G, N, L, 1, EDGES, a, "bf16x3")
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("na,nb", [(32, 1), (1, 32), (17, 9)])
def test_two_output_contraction_row_extremes(gpu_ctx, na, nb):
    """C and C_bb of the two-output launch against the FP64 GEMM of the FP64
    real shells, at the row limits (32 rows, a single row, odd counts)."""
    from cosmonpc_amd import angular as A

    ctx = gpu_ctx
    n, Lb, kmax = 48, 600.0, 0.12
    edges = np.linspace(0.01, kmax, 33)
    R = ctx.synth_radius(Lb, kmax)
    import torch

    D32 = _spectrum(ctx, 13)
    # the FP64 shells come from the same (FP32-valued) spectrum
    D64 = torch.view_as_complex(torch.view_as_real(D32).double().contiguous())
    rng = np.random.default_rng(5)
    G64 = torch.from_numpy(1.0 + 0.5 * rng.standard_normal(n ** 3)).cuda()
    G32 = G64.float().contiguous()
    p0, p2 = A.shell_components(0, 0)[0][0], A.shell_components(2, 0)[0][0]

    def real64(poly, rows):
        out = ctx.mesh(n, 0, rows)
        for b0 in range(0, rows, 8):
            c = min(8, rows - b0)
            ctx.shell_synth(D64, n, Lb, 0, edges, b0, c, poly, ctx.spectrum(n, 0, c), out[b0:b0 + c])
        return out.view(rows, n ** 3)

    def half32(poly, rows):
        H = ctx.half_shells(n, R, rows)
        work = ctx.spectrum(n, 1, 4)
        kz = np.zeros(rows, dtype=np.int32)
        for b0 in range(0, rows, 4):
            c = min(4, rows - b0)
            kz[b0:b0 + c] = ctx.shell_synth_half(D32, n, Lb, edges, b0, c, poly, work, H, b0)
        return H, kz

    HA, kza = half32(p2, na)
    HB, kzb = half32(p0, nb)
    C, Cbb = ctx.contract_zgemm(n, Lb, kmax, G32, HA, na, kza, HB, nb, kzb, with_bb=True)
    A64, B64 = real64(p2, na), real64(p0, nb)
    for got, ref, what in ((C, ctx.contract_gemm("gemm_fp64", G64, A64, B64), "C"),
                           (Cbb, ctx.contract_gemm("gemm_fp64", G64, B64), "C_bb")):
        got, ref = got.cpu().numpy(), ref.cpu().numpy()
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        _log(f"na={na} nb={nb} {what}: rel err {err:.2e}")
        assert got.shape == ref.shape and err < 1e-5, (what, err)
