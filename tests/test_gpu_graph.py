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
