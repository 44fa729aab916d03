"""The allocation-free fused contraction (cnpc_contract_zgemm_noalloc)
against cnpc_contract_zgemm on the same half-synthesised shells: bit for bit
at the same workgroup count, on two calls in a row (the second on a smaller
grid with fewer rows, after the static scratch holds the first call's
partials), with and without C_bb; both against the FP64 GEMM of the FP64
real shells at the split-bf16 bar; the refusal of a workgroup count beyond
the static scratch; and no device memory taken by the call on a fresh
context."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _shells(ctx, n, L, kmax, rows, ell, seed):
    import torch

    from cosmonpc_amd import angular as A

    rng = np.random.default_rng(seed)
    x64 = torch.from_numpy(rng.standard_normal((n, n, n))).cuda()
    D64, D32 = ctx.spectrum(n, 0), ctx.spectrum(n, 1)
    ctx.r2c(x64, D64, n, 0)
    ctx.r2c(x64.float().contiguous(), D32, n, 1)
    G64 = torch.from_numpy(1.0 + 0.5 * rng.standard_normal(n ** 3)).cuda()
    edges = np.linspace(0.01, kmax, rows + 1)
    R = ctx.synth_radius(L, kmax)
    poly = A.shell_components(ell, 0)[0][0]
    H = ctx.half_shells(n, R, rows)
    H.fill_(complex(float("nan"), float("nan")))
    work = ctx.spectrum(n, 1, 4)
    kz = np.zeros(rows, dtype=np.int32)
    for b0 in range(0, rows, 4):
        c = min(4, rows - b0)
        kz[b0:b0 + c] = ctx.shell_synth_half(D32, n, L, edges, b0, c, poly, work, H, b0)
    real = ctx.mesh(n, 0, rows)
    ctx.shell_synth(D64, n, L, 0, edges, 0, rows, poly, ctx.spectrum(n, 0, rows), real)
    return G64, H, kz, real.view(rows, n ** 3)


@pytest.mark.parametrize("n,L,kmax,na,nb,nblk", [(80, 1000.0, 0.15, 7, 6, 96), (48, 600.0, 0.13, 5, 3, 40)])
def test_noalloc_matches_allocating_entry(gpu_ctx, n, L, kmax, na, nb, nblk):
    ctx = gpu_ctx
    G64, HA, kza, A64 = _shells(ctx, n, L, kmax, na, 0, n)
    _, HB, kzb, B64 = _shells(ctx, n, L, kmax, nb, 2, n)  # same G (same seed), other weight
    G32 = G64.float().contiguous()
    ref_aa = ctx.contract_gemm("gemm_fp64", G64, A64).cpu().numpy()
    ref_ab = ctx.contract_gemm("gemm_fp64", G64, A64, B64).cpu().numpy()
    ref_bb = ctx.contract_gemm("gemm_fp64", G64, B64).cpu().numpy()
    for call in range(2):  # twice in a row: the second call finds the first one's partials
        got = ctx.contract_zgemm(n, L, kmax, G32, HA, na, kza, nblk=nblk, noalloc=True).cpu().numpy()
        exp = ctx.contract_zgemm(n, L, kmax, G32, HA, na, kza, nblk=nblk).cpu().numpy()
        assert np.array_equal(got, exp), (call, np.max(np.abs(got - exp)))
        assert np.max(np.abs(got - ref_aa)) <= 2e-6 * np.max(np.abs(ref_aa))
        g2, gbb = ctx.contract_zgemm(n, L, kmax, G32, HA, na, kza, HB, nb, kzb, nblk=nblk, with_bb=True, noalloc=True)
        e2, ebb = ctx.contract_zgemm(n, L, kmax, G32, HA, na, kza, HB, nb, kzb, nblk=nblk, with_bb=True)
        assert np.array_equal(g2.cpu().numpy(), e2.cpu().numpy()), call
        assert np.array_equal(gbb.cpu().numpy(), ebb.cpu().numpy()), call
        assert np.max(np.abs(g2.cpu().numpy() - ref_ab)) <= 2e-6 * np.max(np.abs(ref_ab))
        assert np.max(np.abs(gbb.cpu().numpy() - ref_bb)) <= 2e-6 * np.max(np.abs(ref_bb))
    # nblk <= 0: as many workgroups as the static scratch holds
    auto = ctx.contract_zgemm(n, L, kmax, G32, HA, na, kza, HB, nb, kzb, with_bb=True, noalloc=True)
    assert np.max(np.abs(auto[0].cpu().numpy() - ref_ab)) <= 2e-6 * np.max(np.abs(ref_ab))
    assert np.max(np.abs(auto[1].cpu().numpy() - ref_bb)) <= 2e-6 * np.max(np.abs(ref_bb))


def test_noalloc_refuses_oversized_and_allocates_nothing(gpu_ctx):
    import torch

    from cosmonpc_amd import _lib
    from cosmonpc_amd.device import GpuContext

    ctx = gpu_ctx
    n, L, kmax = 48, 600.0, 0.13
    G64, HA, kza, _ = _shells(ctx, n, L, kmax, 5, 0, 3)
    _, HB, kzb, _ = _shells(ctx, n, L, kmax, 3, 2, 3)
    G32 = G64.float().contiguous()
    # 4096 workgroups with C_bb need 64 MiB of partials
    with pytest.raises(_lib.CnpcError, match="static"):
        ctx.contract_zgemm(n, L, kmax, G32, HA, 5, kza, HB, 3, kzb, nblk=4096, with_bb=True, noalloc=True)
    # a fresh context has no scratch area: the call must not create one
    fresh = GpuContext(ctx.device.index)
    try:
        C = torch.empty((5, 3), dtype=torch.float64, device=ctx.device)
        Cbb = torch.empty((3, 3), dtype=torch.float64, device=ctx.device)
        exp = ctx.contract_zgemm(n, L, kmax, G32, HA, 5, kza, HB, 3, kzb, nblk=64, with_bb=True)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info(ctx.device)[0]
        fresh.contract_zgemm(n, L, kmax, G32, HA, 5, kza, HB, 3, kzb, nblk=64, with_bb=True, noalloc=True,
                             out=(C, Cbb))
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info(ctx.device)[0]
        assert free1 >= free0, (free0, free1)
        assert np.array_equal(C.cpu().numpy(), exp[0].cpu().numpy())
        assert np.array_equal(Cbb.cpu().numpy(), exp[1].cpu().numpy())
    finally:
        fresh.close()
