"""Box-geometry estimators on the GPU.

``calculate_power_spectrum_box`` mirrors ``clustering_estimator.py:566-746``
and ``calculate_bk_sugi_box`` mirrors ``clustering_estimator.py:749-1555``:
same arguments, same result keys, same normalisations.  The numeric work runs
in libcnpc_hip.so; the host keeps only O(k_bins^2) bookkeeping (symmetry
fill, magnetic-config sums, analytic S3 and normalisation).

GPU-only options (read from ``stat_attrs``; the reference ignores them):
  precision   : "fp64" (default) | "fp32"  -- carried by the painted fields
  contraction : "gemm_fp64" (fp64 default) | "pairwise" | "bf16x3"
                (fp32 default) | "fp16x3" | "fp8x6"
  shell_method: "synth" (default, pruned shell synthesis) | "filter"
                (shell filter + 3-D c2r)
  zsynth      : "fused" (default: with bf16x3 the shells' z pass runs
                inside the contraction) | "separate" (real shells in HBM)
"""
import logging
import math
import time

import numpy as np
import torch

from . import angular as A
from ._lib import is_out_of_memory
from .device import PREC, get_context

LAST_TIMINGS = {}


def _sync_time(ctx):
    ctx.sync()  # the context's stream only (a copy stream may be busy)
    return time.perf_counter()


# ----------------------------------------------------------------------- pk
def _pk_polys(even, rsd):
    return [A.make_poly(0, [1.0], A.PART_FULL, 1.0, 1.0) if p == 0 else A.legendre_poly(p, rsd) for p in even]


def pk_box_device_pass(field, stat_attrs):
    """The device work of ``calculate_power_spectrum_box`` for an
    auto-correlation (compensated, interlaced delta_k with k = 0 cleared,
    then the Legendre-weighted binning), enqueued only: returns the function
    that reads (count, ksum, re) back for ``precomputed=``.  Nothing in it
    depends on the catalog beyond the painted field, so it can be recorded in
    a graph once and replayed for every catalog."""
    ctx = get_context()
    poles = stat_attrs["poles"]
    even = [p for p in poles if not p & 1]
    rsd = np.array(stat_attrs["rsd"], dtype=float)
    k_edge = np.linspace(stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"] + 1)
    comp = 1 if stat_attrs["interlaced"] else 2
    da = field.delta_k(ctx, comp, zero_k0=True)
    out = ctx.bin_dev(da, da, field.n, field.boxsize, field.prec, k_edge, _pk_polys(even, rsd))
    nb, npoly = len(k_edge) - 1, len(even)

    def finish():
        count, ksum, re, _ = ctx.split_bins(out.cpu().numpy(), nb, npoly)
        return count, ksum, re
    return finish


def calculate_power_spectrum_box(rfield_a, rfield_b, correlation_mode, stat_attrs, comm=None, **kwargs):
    """P_ell(k) of a periodic box (``clustering_estimator.py:566-746``)."""
    ctx = get_context()
    poles = stat_attrs["poles"]
    boxsize, nmesh = np.array(stat_attrs["boxsize"]), np.array(stat_attrs["nmesh"])
    k_min, k_max, k_bins = stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"]
    sampler, interlaced = stat_attrs["sampler"], stat_attrs["interlaced"]
    rsd = np.array(stat_attrs["rsd"], dtype=float)
    NZ_a = stat_attrs["NZ_a"]
    P_shot = 1.0 / NZ_a if correlation_mode == "auto" else 0.0
    NZ_b = stat_attrs["NZ_b"] if correlation_mode == "cross" else None
    k_edge = np.linspace(k_min, k_max, k_bins + 1)
    n, L, prec = rfield_a.n, rfield_a.boxsize, rfield_a.prec

    even = [p for p in poles if not p & 1]
    precomputed = kwargs.get("precomputed")
    if precomputed is not None:
        # (count, ksum, re) of pk_box_device_pass, read back after a replay
        count, ksum, re = precomputed
    else:
        comp = 1 if interlaced else 2  # sinc^p when interlaced, else sqrt alias sums
        # the zero mode of field a is cleared (clustering_estimator.py:634-635)
        da = rfield_a.delta_k(ctx, comp, zero_k0=True)
        db = rfield_b.delta_k(ctx, comp) if correlation_mode == "cross" else da
        count, ksum, re, _ = ctx.bin(da, db, n, L, prec, k_edge, _pk_polys(even, rsd))
    k_num = count
    with np.errstate(invalid="ignore", divide="ignore"):
        results = {"k_eff": ksum / k_num, "k_num": k_num}
    for pole in poles:
        if pole & 1:
            res = np.zeros(k_bins).astype("float128")
        else:
            with np.errstate(invalid="ignore", divide="ignore"):
                res = re[even.index(pole)] / k_num
            if correlation_mode == "auto":
                res *= boxsize.prod() ** (-1) * (nmesh.prod() ** 2) / NZ_a ** 2
            else:
                res *= boxsize.prod() ** (-1) * nmesh.prod() ** 2 / (NZ_a * NZ_b)
            if pole == 0:
                res -= P_shot
            else:
                res *= 2 * pole + 1
        results.update({f"P{pole}": res, "P_shot": P_shot})
    return results


# ---------------------------------------------------------------- bispectrum
class ShellBuilder:
    """Builds real shell fields  R = c2r(w(khat) * delta * 1[k in bin])  for a
    run of bins, batching the shell filter and the inverse FFTs."""

    def __init__(self, ctx, n, boxsize, prec, edges, batch=4, method="synth"):
        self.ctx, self.n, self.L, self.prec = ctx, n, boxsize, prec
        self.edges = np.asarray(edges, dtype=np.float64)
        self.batch = batch
        if method not in ("synth", "filter"):
            raise ValueError("shell_method must be 'synth' or 'filter'")
        # the pruned synthesis needs |n_d| <= R < n/2, bins that fit int8,
        # a grid that factors into 2, 3, 5, 7 and its two intermediates
        # ((2R+1)(R+1) + (R+1) n per plane) inside the work area
        R = math.ceil(self.edges[-1] / (2 * math.pi / boxsize))
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        fits =(2 * R + 1) * (R + 1) + (R + 1) * n <= n * (n // 2 + 1)
        if method == "synth" and (len(self.edges) - 1 > 127 or R >= n // 2 or m != 1 or not fits):
            method = "filter"
        self.method = method
        self.work = ctx.spectrum(n, prec, batch)
        self._tuned = False

    def _tune_batch(self, delta, poly, bins, out, slot0):
        """Shells per x/y synthesis launch: 4, 8 or 16 (as many as the bins
        allow), each timed on the whole run of ``bins`` in launches of that
        size, on the first run of at least 8 bins (the autotuner caches the
        choice per grid, radius and bin count; the x pass reads the mode
        cube once per launch)."""
        self._tuned = True
        if len(bins) < 8:
            return
        from . import autotune

        shape = (self.n, int(math.ceil(self.edges[-1] / (2 * math.pi / self.L))), len(self.edges) - 1)
        cands = [c for c in (4, 8, 16) if c <= len(bins)]
        best = autotune.lookup("synth_half_batch", shape)
        work = None
        if best is None:
            work = self.ctx.spectrum(self.n, 1, cands[-1])

            def run(c):
                for d in range(0, len(bins), c):
                    self.ctx.shell_synth_half(delta, self.n, self.L, self.edges, bins[0] + d, min(c, len(bins) - d),
                                              poly, work, out, slot0 + d, reuse_cube=d > 0)

            best = autotune.choose("synth_half_batch", shape, cands, run)
        if best != self.batch:
            self.batch = best
            self.work = work if work is not None else self.ctx.spectrum(self.n, self.prec, best)

    def build(self, delta, poly, bins, out):
        """out[r] = shell of bins[r] (bins must be consecutive)."""
        n = self.n
        b0 = bins[0]
        nb = len(bins)
        assert list(bins) == list(range(b0, b0 + nb))
        done = 0
        while done < nb:
            c = min(self.batch, nb - done)
            dst = out[done:done + c].view(c, n, n, n)
            if self.method == "synth":
                self.ctx.shell_synth(delta, n, self.L, self.prec, self.edges, b0 + done, c, poly, self.work, dst)
            else:
                self.ctx.shell_filter(delta, n, self.L, self.prec, self.edges, b0 + done, c, poly, self.work)
                self.ctx.c2r(self.work[:c], dst, n, self.prec, c)
            done += c


    def build_half(self, delta, poly, bins, out, slot0):
        """x and y passes of the shells of ``bins`` (consecutive) into slots
        slot0 .. of the half-synthesised buffer ``out`` (FP32); returns the
        kz rows written per shell."""
        b0 = bins[0]
        nb = len(bins)
        assert list(bins) == list(range(b0, b0 + nb))
        if not self._tuned:
            self._tune_batch(delta, poly, bins, out, slot0)
        done = 0
        rows = []
        while done < nb:
            c = min(self.batch, nb - done)
            # the mode cube of this delta and weight is built by the first batch
            rows.append(self.ctx.shell_synth_half(delta, self.n, self.L, self.edges, b0 + done, c, poly, self.work,
                                                  out, slot0 + done, reuse_cube=done > 0))
            done += c
        return np.concatenate(rows)


def _gemm_mode(contraction, prec):
    if contraction in ("bf16x3", "fp16x3", "fp8x6"):
        if prec != PREC["fp32"]:
            raise ValueError(f"contraction '{contraction}' needs precision 'fp32'")
    elif contraction == "gemm_fp64":
        if prec != PREC["fp64"]:
            raise ValueError("contraction 'gemm_fp64' needs precision 'fp64'")
    elif contraction != "pairwise":
        raise ValueError("contraction must be one of 'pairwise', 'gemm_fp64', 'bf16x3', 'fp16x3', 'fp8x6'")
    return contraction


def _pow2_scale(amax, target=2.0 ** 14):
    if not amax > 0 or not math.isfinite(amax):
        return 1.0
    return 2.0 ** math.floor(math.log2(target / amax))


def _contract(ctx, mode, prec, G, R1, R2, pairs):
    """Real contraction sum_x G R1[a] R2[b]: full matrix (GEMM modes, a
    device tensor) or the listed (a, b) pairs (pairwise, host).  R2 None
    means R2 is R1."""
    if mode == "pairwise":
        ia = np.array([p[0] for p in pairs], dtype=np.int32)
        ib = np.array([p[1] for p in pairs], dtype=np.int32)
        vals = ctx.contract_pairs(prec, G, R1, R1 if R2 is None else R2, ia, ib)
        nr = R1.shape[0]
        nc = R1.shape[0] if R2 is None else R2.shape[0]
        C = np.zeros((nr, nc))
        C[ia, ib] = vals
        return C
    if mode in ("fp16x3", "fp8x6"):
        # power-of-two prescales: fp16 halves stay normal below 2^14, the fp8
        # pieces (e4m3) need |values| <= 2^7
        target = 2.0 ** 14 if mode == "fp16x3" else 2.0 ** 7
        sa = _pow2_scale(ctx.absmax(R1), target)
        sg = ctx.absmax(G)
        sb = _pow2_scale(sg * ctx.absmax(R1 if R2 is None else R2), target)
        return ctx.contract_gemm(mode, G, R1, R2, sa, sb)
    # device result, read back by the plan's finish()
    return ctx.contract_gemm(mode, G, R1, R2, nblk=_gemm_blocks(ctx, mode, G, R1, R2))


def _zgemm_blocks(ctx, zargs, with_bb=False):
    """Workgroups of the fused contraction: 0 lets the library take four per
    CU (a quarter of the tiles per workgroup keeps each wave's FP32 sums
    short; the launch time is flat from one to four per CU)."""
    return 0


class _Landing:
    """A device result copied to pinned host memory right behind the kernels
    that produce it (the copy and an event follow them on the stream), so
    reading it waits for this result only and not for device work enqueued
    after it: the next step's launches can be queued before this one is
    read."""

    def __init__(self, t):
        self.host = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
        self.host.copy_(t, non_blocking=True)
        self.event = torch.cuda.Event()
        self.event.record()

    def get(self):
        self.event.synchronize()
        return self.host.numpy()


class _ZRequest:
    """A fused contraction asked for by a configuration plan; ``value`` is
    its device result once the batch has run, ``landed`` its host copy in
    flight (None while a graph records: the reader then copies)."""

    def __init__(self, zargs):
        self.zargs = zargs
        self.value = None
        self.landed = None


class ZBatch:
    """Fused contractions of several configurations launched together.  A
    B == A request on shells X and a B != A request (Y, X) with the same G
    and the same X buffer become one launch: the B != A kernel returns
    sum_x X G X from the z transforms it already does (``with_bb``), so the
    [0,0,0] signal matrix costs no z pass beside [l,0,l]'s."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.requests = []

    def request(self, zargs):
        r = _ZRequest(zargs)
        self.requests.append(r)
        return r

    def run(self):
        same = [r for r in self.requests if r.zargs[7] is None]
        for r in self.requests:
            if r.zargs[7] is None or r.value is not None:
                continue
            z = r.zargs
            # same grid, box and k_max, the same G, X the very buffer and rows
            mate = next((s for s in same if s.value is None and s.zargs[:3] == z[:3] and s.zargs[3] is z[3]
                         and s.zargs[4] is z[7] and s.zargs[5] == z[8]
                         and np.array_equal(s.zargs[6][:z[8]], z[9][:z[8]])), None)
            if mate is not None:
                r.value, mate.value = self.ctx.contract_zgemm(*z, nblk=_zgemm_blocks(self.ctx, z, True),
                                                             with_bb=True)
        for r in self.requests:
            if r.value is None:
                r.value = self.ctx.contract_zgemm(*r.zargs, nblk=_zgemm_blocks(self.ctx, r.zargs))
        if not _RECORDING[0]:
            for r in self.requests:
                if torch.is_tensor(r.value) and r.value.is_cuda:
                    r.landed = _Landing(r.value)
        self.requests = []


def _gemm_blocks(ctx, mode, G, R1, R2):
    """Split-K workgroups of the separate-path GEMM, autotuned per shape."""
    from . import autotune

    shape = (mode, G.numel(), R1.shape[0], 0 if R2 is None else R2.shape[0])
    return autotune.choose("gemm_blocks", shape, [2048, 1024, 4096],
                           lambda c: ctx.contract_gemm(mode, G, R1, R2, nblk=c))


_FIT_SEEN = {}  # last block size per (k_bins, bytes per bin, reserve)
_RECORDING = [False]  # set while a graph records: block sizes repeat the eager run's


def _fit_block(k_bins, bytes_per_bin, ctx, reserve=0.25):
    """Largest block of bins whose shells fit in the free HBM (keeping a
    fraction in reserve for FFT work areas and spectra).  While a graph is
    being recorded the size of the eager run before it is reused, so the
    recorded launches are the ones that ran."""
    key = (k_bins, bytes_per_bin, reserve)
    if _RECORDING[0] and key in _FIT_SEEN:
        return _FIT_SEEN[key]
    free, total = torch.cuda.mem_get_info(ctx.device)
    budget = free - reserve * total
    _FIT_SEEN[key] = int(max(1, min(k_bins, budget // max(bytes_per_bin, 1))))
    return _FIT_SEEN[key]


def _weight_components(ell, m, conj_w):
    """Real shell components of Y_lm (or of conj(Y_lm) = (-1)^m Y_l,-m)."""
    if not conj_w:
        return A.shell_components(ell, m)
    return [(p, (-1) ** m * ph) for p, ph in A.shell_components(ell, -m)]


def _config_matrix(*args, **kwargs):
    """``_config_plan`` run to completion: the signal matrix (numpy)."""
    return _config_plan(*args, **kwargs)()


def _config_blocks(ctx, shells, Gs, kinfo, src1, src2, comps1, comps2, use_sym, dvm, bs, fused, kmax, mode, prec,
                   zbatch):
    """The blocks of rows and columns of ``_config_plan`` at block size
    ``bs``: (ib, bi, bj, rows2, want, [(result, phase)]) each, their device
    work enqueued."""
    n, L, _, _, k_bins = kinfo
    nc1, nc2 = len(comps1), len(comps2)
    blocks = []  # (ib, bi, bj, rows2, want, [(result, phase)]) per block of rows and columns
    for bi in range(0, k_bins, bs):
        ib = list(range(bi, min(k_bins, bi + bs)))
        R1 = shells(src1, comps1, ib)
        jstarts = [bi] if dvm == "diagonal" else list(range(bi if use_sym else 0, k_bins, bs))
        for bj in jstarts:
            jb = list(range(bj, min(k_bins, bj + bs)))
            if bj == bi and use_sym:
                R2 = None
            else:
                # off the diagonal the symmetric case rebuilds leg 1's shells
                R2 = shells(src1, comps1, jb) if use_sym else shells(src2, comps2, jb)
            # pairs this block contributes (reference loop order)
            want = []
            for i in ib:
                if dvm == "diagonal":
                    js = [i]
                elif use_sym and bj == bi:
                    js = [j for j in jb if j >= i]
                else:
                    js = jb
                for j in js:
                    want.append((i, j))
            rows2 = len(ib) if R2 is None else len(jb)
            pairs = [(c1 * len(ib) + (i - bi), c2 * rows2 + (j - bj)) for (i, j) in want
                     for c1 in range(nc1) for c2 in range(nc2)]
            parts = []
            for G, gph in Gs:
                if fused:
                    zargs = (n, L, kmax, G, R1[0], nc1 * len(ib), R1[1], None if R2 is None else R2[0],
                             nc2 * rows2, None if R2 is None else R2[1])
                    # device result, read back by finish()
                    if zbatch is not None:
                        parts.append((zbatch.request(zargs), gph))
                    else:
                        parts.append((ctx.contract_zgemm(*zargs, nblk=_zgemm_blocks(ctx, zargs)), gph))
                else:
                    parts.append((_contract(ctx, mode, prec, G, R1, R2, pairs), gph))
            blocks.append((ib, bi, bj, rows2, want, parts))
            del R2
        del R1
    return blocks


def _fused_rows(ctx, n, mode, zsynth, method, nc1, nc2, k_bins):
    """(fused, rows per block): whether a configuration with nc1 / nc2 real
    shell components per bin takes the fused z-synthesis contraction, and
    the most bins per block it allows."""
    if not (zsynth == "fused" and mode == "bf16x3" and method == "synth"):
        return False, 0
    cap = min(k_bins, 32 // max(nc1, nc2))
    return cap >= 1 and ctx.zgemm_supported(n, nc1 * cap, nc2 * cap, False), cap


def _config_plan(ctx, builder, src1, src2, Gs, kinfo, ell_1, m1, ell_2, m2, sym_eq, sym_cj, mode, block_size,
                 dvm, conj_w=False, zsynth="fused", shell_cache=None, zbatch=None):
    """3j-free signal matrix of one magnetic configuration,
    S_ij = sum_x G F_i F'_j, with the reference's upper-triangle evaluation and
    lower fill (transpose for F' = F, conjugate for F' = +-conj F;
    ``clustering_estimator.py:911-989``, ``:1773-1858``).

    F_i = c2r(w1(khat) src1 mask_i), F'_j = c2r(w2(khat) src2 mask_j) with
    w = Y_lm (or conj Y_lm), split into real shell components; src None is a
    unit spectrum.  ``Gs`` lists (real mesh, complex factor) pairs whose sum is
    the (possibly complex) G.

    With the split-bf16 GEMM and ``zsynth="fused"`` the shells stay half
    synthesised (x and y passes) and their z pass runs inside the contraction
    (``cnpc_contract_zgemm``) whenever a block of rows fits it; otherwise the
    real shells are built and contracted separately.

    The device work is enqueued here; the returned ``finish()`` reads the
    results back and assembles the matrix, and may be called again after the
    same work is replayed from a graph.  The GEMM contractions (fused or
    not) leave their results on the device until then; the pairwise path
    reads them back as it goes.  ``shell_cache`` (a dict shared by several calls on the same spectra)
    keeps the half-synthesised shells of the fused path, so a leg that
    another configuration already synthesised (same spectrum, weight and
    bins) is not synthesised again; with ``zbatch`` (a ``ZBatch``) the fused
    contractions are only requested here and launched by its ``run()``,
    before ``finish``."""
    n, L, prec, k_edge, k_bins = kinfo
    N3 = n ** 3
    use_sym = sym_eq or sym_cj
    comps1 = _weight_components(ell_1, m1, conj_w)
    if use_sym:
        sign = (-1) ** (m1 + ell_1)
        # F' = F (equal) or (-1)^(m+l) conj(F): same real shells, new phases
        comps2 = [(p, ph if sym_eq else sign * np.conj(ph)) for p, ph in comps1]
    else:
        comps2 = _weight_components(ell_2, m2, conj_w)
    nc1, nc2 = len(comps1), len(comps2)
    kmax = float(k_edge[-1])
    fused, cap = _fused_rows(ctx, n, mode, zsynth, builder.method, nc1, nc2, k_bins)
    if fused:
        R = ctx.synth_radius(L, kmax)
        shell_bytes = n * n * (R + 1) * 8  # one slot of the half-synthesised layout
    else:
        shell_bytes = N3 * (8 if prec == 0 else 4)
    # resident shells per block: nc1 * bs (+ nc2 * bs off the diagonal)
    bs_fit = _fit_block(k_bins, shell_bytes * (nc1 + nc2), ctx)
    if use_sym and _fit_block(k_bins, shell_bytes * nc1, ctx) == k_bins:
        bs_fit = k_bins  # one diagonal block: leg 2 reuses leg 1's shells
    if dvm == "full":
        bs = min(bs_fit, k_bins if block_size == "full" else int(block_size))
    else:
        bs = bs_fit
    if fused:
        bs = min(bs, cap)

    def shells(src, comps, bins):
        rows = len(comps) * len(bins)
        if fused:
            key = (0 if src is None else src.data_ptr(), tuple(bytes(p) for p, _ in comps), tuple(bins), n, R)
            if shell_cache is not None and key in shell_cache:
                return shell_cache[key]
            H = ctx.half_shells(n, R, rows)
            kz = [builder.build_half(src, poly, bins, H, c * len(bins)) for c, (poly, _) in enumerate(comps)]
            if shell_cache is not None:
                shell_cache[key] = (H, np.concatenate(kz))
            return H, np.concatenate(kz)
        out = ctx.mesh(n, prec, rows).view(rows, N3)
        for c, (poly, _) in enumerate(comps):
            builder.build(src, poly, bins, out[c * len(bins):(c + 1) * len(bins)])
        return out

    # a device allocation failure (hipMalloc OOM in the library or torch's)
    # while the blocks are built: the blocks are rebuilt with half the
    # block size, down to one bin (not while a graph records)
    nreq = len(zbatch.requests) if zbatch is not None else 0
    while True:
        try:
            blocks = _config_blocks(ctx, shells, Gs, kinfo, src1, src2, comps1, comps2, use_sym, dvm, bs, fused,
                                    kmax, mode, prec, zbatch)
            break
        except Exception as e:
            if not is_out_of_memory(e) or bs <= 1 or _RECORDING[0]:
                raise
            if zbatch is not None:
                del zbatch.requests[nreq:]
            blocks = None
            torch.cuda.empty_cache()
            logging.warning(f"device memory exhausted with block_size {bs}: retrying with {bs // 2}")
            bs //= 2

    def finish():
        sub = np.zeros((k_bins, k_bins) if dvm == "full" else k_bins, dtype=np.complex128)
        for ib, bi, bj, rows2, want, parts in blocks:
            C = 0j
            for Cm, gph in parts:
                if isinstance(Cm, _ZRequest):
                    Cm = Cm.landed.get() if Cm.landed is not None else Cm.value
                C = C + gph * (Cm.cpu().numpy() if torch.is_tensor(Cm) else Cm)
            for (i, j) in want:
                s = 0j
                for c1 in range(nc1):
                    for c2 in range(nc2):
                        s += comps1[c1][1] * comps2[c2][1] * C[c1 * len(ib) + (i - bi), c2 * rows2 + (j - bj)]
                if dvm == "full":
                    sub[i, j] = s
                else:
                    sub[i] = s
        if dvm == "full" and use_sym:
            iu = np.triu_indices(k_bins, 1)
            sub[iu[1], iu[0]] = sub[iu] if sym_eq else np.conj(sub[iu])
        return sub

    return finish


def _bk_signal(ctx, src1, src2, G, kinfo, mcfg, tjv, ell_1, ell_2, sym_allowed, mode, block_size, dvm,
               conj_w=False, shell_method="synth", zsynth="fused", defer=False, shell_cache=None, zbatch=None):
    """Sum over the box magnetic configs of 3j * sum_x G F_i F'_j with the
    reference's symmetry fill and 2 Re accumulation
    (``clustering_estimator.py:867-1061``; with ``conj_w`` and null sources,
    the S3 loop of ``:1219-1376``).  The device work of every magnetic
    config is enqueued before any result is read back; ``defer`` returns the
    function that reads them instead of the sum."""
    n, L, prec, k_edge, k_bins = kinfo
    builder = ShellBuilder(ctx, n, L, prec, k_edge, method=shell_method)
    same_ell = ell_1 == ell_2
    Gs = G if isinstance(G, list) else [(G, 1.0)]
    plans = []
    for ss, (m1, m2, M) in enumerate(mcfg):
        sym_eq = sym_allowed and same_ell and m1 == m2
        sym_cj = sym_allowed and same_ell and m1 == -m2
        plans.append((tjv[ss], (m1, m2, M), _config_plan(ctx, builder, src1, src2, Gs, kinfo, ell_1, m1, ell_2, m2,
                                                         sym_eq, sym_cj, mode, block_size, dvm, conj_w, zsynth,
                                                         shell_cache, zbatch)))

    def finish():
        total = np.zeros((k_bins, k_bins) if dvm == "full" else k_bins, dtype=np.complex128)
        for t, cfg, plan in plans:
            sub = t * plan()
            total += sub if cfg == (0, 0, 0) else 2 * np.real(sub)
        return total

    return finish if defer else finish()


def _nice_size(m):
    """Smallest m' >= m with m' % 4 == 0 and prime factors in {2, 3, 5}
    (the radices of the shell synthesis)."""
    while True:
        if m % 4 == 0:
            r = m
            for p in (2, 3, 5):
                while r % p == 0:
                    r //= p
            if r == 1:
                return m
        m += 1


def contraction_grid(n, boxsize, k_max, contraction, shell_grid="reduced"):
    """Grid for the shell contraction.  Shells hold |n_d| <= R = ceil(k_max/kf)
    and only G modes with |n_d| <= 2R meet them in a triangle, so
    sum_x G F_i F_j over the n-grid equals (n/m)^3 times the same sum over any
    m-grid with m >= 4R + 2 (no aliasing of n1 + n2 + n3 = 0).  Returns
    (m, nmax) or (n, None) when the reduced grid is not smaller."""
    if shell_grid == "full" or contraction == "pairwise":
        return n, None
    kf = 2 * math.pi / float(boxsize)
    R = int(math.ceil(k_max / kf))
    m = _nice_size(4 * R + 2)
    if m >= n:
        return n, None
    return m, 2 * R


def reduce_spectra(ctx, specs, n, prec, m, nmax):
    """Resample distinct spectra onto the m-grid (None passes through)."""
    out, seen = [], {}
    for s in specs:
        if s is None:
            out.append(None)
        elif id(s) in seen:
            out.append(seen[id(s)])
        else:
            r = ctx.resample(s, n, m, prec, nmax)
            seen[id(s)] = r
            out.append(r)
    return out


def bk_box_g00(ctx, delta, n, prec):
    """G_00 = Re c2r(delta_k) (``clustering_estimator.py:844-851``) as a flat
    real mesh; ``delta`` is left untouched."""
    tmp = delta.clone()
    G = ctx.mesh(n, prec)
    ctx.c2r(tmp, G, n, prec)
    del tmp
    return G.view(-1)


def bk_box_signal_auto(ctx, delta, G, n, boxsize, prec, k_edge, angu_config, contraction, dvm="full",
                       block_size="full", zsynth="fused", defer=False, shell_cache=None, zbatch=None,
                       shell_method="synth"):
    """Signal stage of an auto-correlation box bispectrum: from the
    compensated delta_k and G_00 resident in HBM to ``total_res``
    (``clustering_estimator.py:867-1061``) for one angular configuration.
    ``defer``: enqueue only and return the function that reads the result."""
    ell_1, ell_2, L = angu_config
    mcfg, tjv = A.get_magnetic_configs_box(ell_1, ell_2, L)
    k_bins = len(k_edge) - 1
    return _bk_signal(ctx, delta, delta, G, (n, boxsize, prec, np.asarray(k_edge), k_bins), mcfg, tjv, ell_1,
                      ell_2, True, _gemm_mode(contraction, prec), block_size, dvm, shell_method=shell_method,
                      zsynth=zsynth, defer=defer, shell_cache=shell_cache, zbatch=zbatch)


def _all_fused(ctx, n, mode, zsynth, method, angu_configs, k_bins):
    """Every magnetic configuration of every angular configuration takes the
    fused contraction (so all contractions wait for ``ZBatch.run``)."""
    for ell_1, ell_2, L in angu_configs:
        mcfg, _ = A.get_magnetic_configs_box(ell_1, ell_2, L)
        for m1, m2, _ in mcfg:
            nc1 = len(_weight_components(ell_1, m1, False))
            nc2 = nc1 if ell_1 == ell_2 and abs(m1) == abs(m2) else len(_weight_components(ell_2, m2, False))
            if not _fused_rows(ctx, n, mode, zsynth, method, nc1, nc2, k_bins)[0]:
                return False
    return True


def bk_box_signal_multi(ctx, delta, G, n, boxsize, prec, k_edge, angu_configs, contraction, dvm="full",
                        block_size="full", zsynth="fused", defer=False, shell_method="synth", consume_delta=False):
    """``bk_box_signal_auto`` for several angular configurations of the same
    field (the reference runs one ``run_stats`` per configuration).  With the
    fused contraction, shells that several configurations use are
    synthesised once (the unweighted shells of [0,0,0] are the second leg of
    every [l,0,l]) and the contractions are launched together (``ZBatch``:
    [0,0,0]'s comes out of [l,0,l]'s launch).  Returns the list of
    ``total_res`` (or, with ``defer``, the function that reads it); each
    equals the single-configuration result up to the FP32 accumulation order
    of the contraction.

    ``G`` None: G_00 is made here.  With ``consume_delta`` (the caller no
    longer needs ``delta``) and every contraction fused, the inverse FFT of
    G_00 runs on ``delta`` itself after the shells are synthesised (they
    read it first), which saves the copy of ``delta`` that the transform
    would otherwise overwrite."""
    cache = {}
    zbatch = ZBatch(ctx)
    late = False
    if G is None:
        mode = _gemm_mode(contraction, prec)
        late = consume_delta and _all_fused(ctx, n, mode, zsynth, shell_method, angu_configs, len(k_edge) - 1)
        G = ctx.mesh(n, prec).view(-1) if late else bk_box_g00(ctx, delta, n, prec)
    fins = [bk_box_signal_auto(ctx, delta, G, n, boxsize, prec, k_edge, a, contraction, dvm, block_size, zsynth,
                               defer=True, shell_cache=cache, zbatch=zbatch, shell_method=shell_method)
            for a in angu_configs]
    if late:
        ctx.c2r(delta, G.view(n, n, n), n, prec)
    zbatch.run()
    cache.clear()  # freed buffers are reused in stream order, after the work that reads them

    def finish():
        return [f() for f in fins]

    return finish if defer else finish()


SIGNAL_KEYS = ("k_min", "k_max", "k_bins", "data_vector_mode", "block_size", "contraction", "shell_method",
               "zsynth", "shell_grid", "precision")


def signal_group_key(cfg):
    """Settings that must agree for box auto-correlation bk_sugi
    configurations to share one multi-configuration signal pass (None for
    configurations that cannot: cross-correlation, survey, not bk_sugi)."""
    if cfg.get("statistic") != "bk_sugi" or cfg.get("correlation_mode") != "auto" or \
            cfg.get("geometry") != "box-like":
        return None
    dvm = cfg.get("data_vector_mode", "diagonal")
    prec = cfg.get("precision", "fp64")
    resolved = dict(cfg, data_vector_mode=dvm, precision=prec,
                    block_size=cfg.get("block_size", 1 if dvm == "diagonal" else "full"),
                    contraction=cfg.get("contraction", "gemm_fp64" if prec == "fp64" else "bf16x3"),
                    shell_method=cfg.get("shell_method", "synth"), zsynth=cfg.get("zsynth", "fused"),
                    shell_grid=cfg.get("shell_grid", "reduced"))
    return tuple(repr(resolved[k]) for k in SIGNAL_KEYS)


def bk_box_signals(field, stat_attrs_list):
    """``total_res`` of several box auto-correlation bk_sugi configurations
    of one painted field (``signal_group_key`` equal), from one
    ``bk_box_signal_multi`` pass: the delta_k, G_00 and reduced-grid spectra
    are made once, shells shared by several configurations are synthesised
    once and [0,0,0] comes out of [l,0,l]'s contraction.  Each entry is what
    ``calculate_bk_sugi_box`` computes for that configuration
    (``clustering_estimator.py:808-1061``), up to the FP32 summation order of
    the split-bf16 contraction; pass it back as ``signal=``."""
    ctx = get_context()
    sa = stat_attrs_list[0]
    dvm = sa.get("data_vector_mode", "diagonal")
    block_size = sa.get("block_size", 1 if dvm == "diagonal" else "full")
    n, Lbox, prec = field.n, field.boxsize, field.prec
    k_min, k_max, k_bins = sa["k_min"], sa["k_max"], sa["k_bins"]
    k_edge = np.linspace(k_min, k_max, k_bins + 1)
    contraction = _gemm_mode(sa.get("contraction", "gemm_fp64" if prec == 0 else "bf16x3"), prec)
    vol_per_cell = np.prod(np.array(sa["boxsize"])) / np.prod(np.array(sa["nmesh"]))
    ca = field.delta_k(ctx, 1, scale=1.0 / vol_per_cell, sub0=sa["NZ_a"])
    m, nmax = contraction_grid(n, Lbox, k_max, contraction, sa.get("shell_grid", "reduced"))
    ca_m = reduce_spectra(ctx, [ca], n, prec, m, nmax)[0] if m != n else ca
    res = bk_box_signal_multi(ctx, ca_m, None, m, Lbox, prec, k_edge,
                              [tuple(s["angu_config"]) for s in stat_attrs_list], contraction, dvm, block_size,
                              sa.get("zsynth", "fused"), shell_method=sa.get("shell_method", "synth"),
                              consume_delta=m != n)
    return [(n / m) ** 3 * r for r in res]


def _sn_poly(L):
    return A.ylm_poly(L, 0, True) if L != 0 else A.make_poly(0, [1.0], A.PART_FULL, 1.0, 1.0)


def _k3_edges(k_min, k_max, k_bins):
    return np.linspace(2 * k_min, 2 * k_max, k_bins * 2 + 1)[:-1]


def bk_box_bin_constants(stat_attrs, n, boxsize, prec):
    """The catalog-independent binnings of ``calculate_bk_sugi_box``: mode
    counts and |k| sums on the k and k3 bins, and the k3 bins of the shot
    noise template C_p/sinc^2p times the Y_L0 weight (what the binning
    subtracts per unit of SN0)."""
    ctx = get_context()
    k_min, k_max, k_bins = stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"]
    k_edge = np.linspace(k_min, k_max, k_bins + 1)
    k3_edge = _k3_edges(k_min, k_max, k_bins)
    p = {"ngp": 1, "cic": 2, "tsc": 3, "pcs": 4}[stat_attrs["sampler"]]
    count, ksum, _, _ = ctx.bin(None, None, n, boxsize, prec, k_edge)
    count3, ksum3, _, _ = ctx.bin(None, None, n, boxsize, prec, k3_edge)
    zero = ctx.spectrum(n, prec)
    zero.zero_()
    # the binning returns ab_scale a conj(b) - noise_amp * template: with a
    # zero spectrum and unit amplitude, minus the template's bins
    _, _, tmpl, _ = ctx.bin(zero, zero, n, boxsize, prec, k3_edge, [_sn_poly(stat_attrs["angu_config"][2])],
                            ab_scale=0.0, noise_window=p, noise_amp=1.0)
    return {"k": (count, ksum), "k3": (count3, ksum3), "noise": -tmpl[0]}


def bk_box_device_pass(field, stat_attrs_list, sub0_dev):
    """The device work of ``calculate_bk_sugi_box`` for the auto-correlation
    configurations of one signal group ('ana' shot noise), enqueued only:
    delta_k, the multi-configuration signal pass of ``bk_box_signals`` and
    the k3 binning of V^2 |delta_k|^2 per configuration.  NZ (``sub0_dev``, a
    device FP64 scalar) is subtracted from the zero mode on the device, so
    nothing recorded depends on the catalog and the work can be replayed
    from a graph.  Returns the function that reads, per configuration,
    {"signal", "re3"} back (combine with ``bk_box_bin_constants`` into
    ``precomputed=``)."""
    ctx = get_context()
    sa = stat_attrs_list[0]
    dvm = sa.get("data_vector_mode", "diagonal")
    block_size = sa.get("block_size", 1 if dvm == "diagonal" else "full")
    n, Lbox, prec = field.n, field.boxsize, field.prec
    k_min, k_max, k_bins = sa["k_min"], sa["k_max"], sa["k_bins"]
    k_edge = np.linspace(k_min, k_max, k_bins + 1)
    k3_edge = _k3_edges(k_min, k_max, k_bins)
    contraction = _gemm_mode(sa.get("contraction", "gemm_fp64" if prec == 0 else "bf16x3"), prec)
    V = float(np.prod(np.array(sa["boxsize"], dtype=float)))
    vol_per_cell = V / np.prod(np.array(sa["nmesh"]))
    p = {"ngp": 1, "cic": 2, "tsc": 3, "pcs": 4}[sa["sampler"]]
    ca = field.delta_k(ctx, 1, scale=1.0 / vol_per_cell, sub0=0.0)
    ca.view(-1)[0] -= sub0_dev
    m, nmax = contraction_grid(n, Lbox, k_max, contraction, sa.get("shell_grid", "reduced"))
    ca_m = reduce_spectra(ctx, [ca], n, prec, m, nmax)[0] if m != n else ca
    G = bk_box_g00(ctx, ca_m, m, prec)
    sig = bk_box_signal_multi(ctx, ca_m, G, m, Lbox, prec, k_edge, [tuple(s["angu_config"]) for s in stat_attrs_list],
                              contraction, dvm, block_size, sa.get("zsynth", "fused"), defer=True,
                              shell_method=sa.get("shell_method", "synth"))
    bins3 = [ctx.bin_dev(ca, ca, n, Lbox, prec, k3_edge, [_sn_poly(s["angu_config"][2])], ab_scale=V * V,
                         noise_window=p, noise_amp=0.0) for s in stat_attrs_list]
    nb3 = len(k3_edge) - 1

    def finish():
        res = sig()
        out = []
        for r, b3 in zip(res, bins3):
            _, _, re3, _ = ctx.split_bins(b3.cpu().numpy(), nb3, 1)
            out.append({"signal": (n / m) ** 3 * r, "re3": re3})
        return out
    return finish


def calculate_bk_sugi_box(rfield_a, rfield_b, rfield_c, correlation_mode, stat_attrs, comm=None, **kwargs):
    """Sugiyama bispectrum multipole of a periodic box
    (``clustering_estimator.py:749-1555``)."""
    ctx = get_context()
    dvm = stat_attrs.get("data_vector_mode", "diagonal")
    sn_mode = stat_attrs.get("shotnoise_mode", "ana")
    all_eff = kwargs.get("shotnoise_ana_all_eff", stat_attrs.get("shotnoise_ana_all_eff", True))
    block_size = stat_attrs.get("block_size", 1 if dvm == "diagonal" else "full")
    ell_1, ell_2, L = stat_attrs["angu_config"]
    boxsize = np.array(stat_attrs["boxsize"])
    nmesh = np.array(stat_attrs["nmesh"])
    k_min, k_max, k_bins = stat_attrs["k_min"], stat_attrs["k_max"], stat_attrs["k_bins"]
    sampler = stat_attrs["sampler"]
    tracer_type = stat_attrs["tracer_type"]
    vol_per_cell = boxsize.prod() / nmesh.prod()
    n, Lbox, prec = rfield_a.n, rfield_a.boxsize, rfield_a.prec
    contraction = _gemm_mode(stat_attrs.get("contraction", "gemm_fp64" if prec == 0 else "bf16x3"), prec)
    shell_method = stat_attrs.get("shell_method", "synth")
    zsynth = stat_attrs.get("zsynth", "fused")

    N_ells = (2 * ell_1 + 1) * (2 * ell_2 + 1) * (2 * L + 1)
    H_ells = A.wigner_3j(ell_1, ell_2, L, 0, 0, 0)
    mcfg, tjv = A.get_magnetic_configs_box(ell_1, ell_2, L)

    NZ_a, N_gal_a = stat_attrs["NZ_a"], stat_attrs["N_gal_a"]
    if correlation_mode == "cross":
        NZ_b, N_gal_b = stat_attrs["NZ_b"], stat_attrs["N_gal_b"]
        if tracer_type == "abc":
            NZ_c, N_gal_c = stat_attrs["NZ_c"], stat_attrs["N_gal_c"]
    k_edge = np.linspace(k_min, k_max, k_bins + 1)

    if dvm == "full":
        if block_size == "full":
            pass
        elif not (isinstance(block_size, int) and 1 <= block_size <= k_bins):
            raise ValueError("For full mode, block_size must be 1, 'full', or an integer in [1, k_bins].")
    elif dvm != "diagonal":
        raise ValueError("data_vector_mode must be either 'diagonal' or 'full'")

    # precomputed: the device results of bk_box_device_pass for this
    # configuration, read back after a graph replay (auto, 'ana' shot noise)
    pre = kwargs.get("precomputed")
    if pre is not None and (correlation_mode != "auto" or sn_mode != "ana"):
        raise ValueError("precomputed device results cover auto-correlation with shotnoise_mode 'ana' only")
    t0 = _sync_time(ctx) if pre is None else time.perf_counter()
    if pre is None:
        # delta_k = sinc^p-compensated r2c(n/dV - NZ)   (clustering_estimator.py:808-816)
        ca = rfield_a.delta_k(ctx, 1, scale=1.0 / vol_per_cell, sub0=NZ_a)
        cb = rfield_b.delta_k(ctx, 1, scale=1.0 / vol_per_cell, sub0=NZ_b) if correlation_mode == "cross" else None
        cc = (rfield_c.delta_k(ctx, 1, scale=1.0 / vol_per_cell, sub0=NZ_c)
              if (correlation_mode == "cross" and tracer_type == "abc") else None)
        # G_00 = Re c2r(delta_k) of the third leg (:844-851)
        gsrc = ca if correlation_mode == "auto" else (cc if tracer_type == "abc" else cb)
        count, ksum, _, _ = ctx.bin(None, None, n, Lbox, prec, k_edge)
    else:
        count, ksum = pre["k"]
    k_num = count
    with np.errstate(invalid="ignore", divide="ignore"):
        k_eff = ksum / k_num
    t1 = _sync_time(ctx) if pre is None else time.perf_counter()
    signal = kwargs.get("signal") if pre is None else pre["signal"]
    # the contraction runs on the reduced grid m (exact up to rounding)
    m, nmax = contraction_grid(n, Lbox, k_max, contraction, stat_attrs.get("shell_grid", "reduced"))
    grid_factor = (n / m) ** 3
    kinfo = (m, Lbox, prec, k_edge, k_bins)
    if signal is not None:
        # total_res of this configuration from a multi-configuration pass
        # over the same field (bk_box_signals)
        total_res = np.asarray(signal)
    else:
        # leg 2 shells come from tracer a for aaa/aab, else b (:899-902)
        src2 = ca if tracer_type in ("aaa", "aab") else cb
        same_tracer = correlation_mode == "auto" or tracer_type == "aab"
        if m != n:
            ca_m, src2_m, gsrc_m = reduce_spectra(ctx, [ca, src2, gsrc], n, prec, m, nmax)
        else:
            ca_m, src2_m, gsrc_m = ca, src2, gsrc
        G = bk_box_g00(ctx, gsrc_m, m, prec)
        total_res = grid_factor * _bk_signal(ctx, ca_m, src2_m, G, kinfo, mcfg, tjv, ell_1, ell_2, same_tracer,
                                             contraction, block_size, dvm, shell_method=shell_method, zsynth=zsynth)
        del G, ca_m, src2_m, gsrc_m
    t2 = _sync_time(ctx) if pre is None else time.perf_counter()
    LAST_TIMINGS.update({"prepare_s": t1 - t0, "signal_s": t2 - t1, "contraction": contraction, "grid": m})

    # ---------------------------------------------------------- shot noise
    k_center = 0.5 * (k_edge[1:] + k_edge[:-1])
    k3_edge = _k3_edges(k_min, k_max, k_bins)
    k3_center = 0.5 * (k3_edge[1:] + k3_edge[:-1])
    k3_bins = len(k3_center)
    if pre is None:
        k_num_k3, ksum3, _, _ = ctx.bin(None, None, n, Lbox, prec, k3_edge)
    else:
        k_num_k3, ksum3 = pre["k3"]
    with np.errstate(invalid="ignore", divide="ignore"):
        k3_eff = ksum3 / k_num_k3
    p = {"ngp": 1, "cic": 2, "tsc": 3, "pcs": 4}[sampler]

    SN3_fft = None
    if tracer_type == "abc":
        SN0 = SN1 = SN2 = SN3 = 0.0
    else:
        SN0 = N_gal_a if (correlation_mode == "auto" and [ell_1, ell_2, L] == [0, 0, 0]) else 0.0
        # P = V^2 a conj(b) - SN0 C_p/sinc^2p, times y_L0 if L != 0, on k3 bins (:1405-1434)
        if pre is None:
            bsrc = ca if correlation_mode == "auto" else cb
            _, _, re3, _ = ctx.bin(ca, bsrc, n, Lbox, prec, k3_edge, [_sn_poly(L)], ab_scale=boxsize.prod() ** 2,
                                   noise_window=p, noise_amp=SN0 if correlation_mode == "auto" else 0.0)
        else:
            # bins of V^2 |a|^2 and of the C_p/sinc^2p template, recombined
            # (the binning is linear in the noise amplitude)
            re3 = pre["re3"] - SN0 * pre["noise"]
        total_sum = re3[0].astype(np.complex128)
        SN1, SN2 = 0.0, 0.0
        if ell_1 == L and ell_2 == 0:
            with np.errstate(invalid="ignore", divide="ignore"):
                if correlation_mode == "auto":
                    SN1 = (total_sum / k_num_k3)[: len(k_center)] * (2 * L + 1)
                    if ell_1 == 0:
                        SN2 = SN1.copy()
                elif tracer_type == "abb":
                    SN1 = (total_sum / k_num_k3)[: len(k_center)] * (2 * L + 1)
        if dvm == "full":
            S1 = np.zeros((k_bins, k_bins), dtype="c16")
            S2 = np.zeros((k_bins, k_bins), dtype="c16")
            if np.ndim(SN1) == 0:
                S1[:] = SN1
            else:
                S1[:] = np.asarray(SN1)[:, None]
            if np.ndim(SN2) == 0:
                S2[:] = SN2
            else:
                S2[:] = np.asarray(SN2)[None, :]
            SN1, SN2 = S1, S2

        def s3_ana():
            if tracer_type == "abb":
                return 0.0

            # q at the effective k (default) or the bin centres (:1102-1126),
            # every (i, j) of the data vector at once
            k12, k3q = (k_eff, k3_eff) if all_eff else (k_center, A.k3_bin_centres(k_min, k_max, k_bins))
            with np.errstate(invalid="ignore", divide="ignore"):
                weighted_k_num = k_num_k3 / k3_center
            if dvm == "diagonal":
                q = A.q_ells_pairs(k12[:k_bins], k12[:k_bins], ell_1, ell_2, L, k3q)
                numer = np.sum(q * total_sum / k3_center, axis=-1)
                denom = np.array([np.sum(weighted_k_num[: 2 * i + 1]) for i in range(k_bins)])
            else:
                q = A.q_ells_pairs(k12[:k_bins, None], k12[None, :k_bins], ell_1, ell_2, L, k3q)
                valid = A.valid_k3_bins_pairs(k_center[:, None], k_center[None, :], k_min, k_max, k_bins)
                numer = np.sum(q * total_sum / k3_center, axis=-1)
                denom = np.sum(weighted_k_num * valid, axis=-1)
            s3 = np.where(denom > 0, numer / np.where(denom > 0, denom, 1.0), 0.0).astype("complex128")
            return s3 * (H_ells * N_ells)

        def s3_fft():
            """S3 from shells of conj(Y) * mask contracted with the source
            field c2r(V^2 a conj(b) - SN0 C_p/sinc^2p)  (:1182-1384)."""
            if tracer_type == "abb":
                return np.zeros((k_bins, k_bins) if dvm == "full" else k_bins, dtype="complex128")
            spec = ctx.spectrum(n, prec)
            ctx.cross_spectrum(ca, ca if correlation_mode == "auto" else cb, spec, n, prec, boxsize.prod() ** 2, p,
                               SN0 if correlation_mode == "auto" else 0.0)
            if m != n:
                spec = ctx.resample(spec, n, m, prec, nmax)
            S = ctx.mesh(m, prec)
            ctx.c2r(spec, S, m, prec)
            del spec
            tot = grid_factor * _bk_signal(ctx, None, None, S.view(-1), kinfo, mcfg, tjv, ell_1, ell_2, True,
                                           contraction, block_size, dvm, conj_w=True,
                                           shell_method=shell_method, zsynth=zsynth)
            del S
            kk = k_num.astype(np.float64)
            with np.errstate(invalid="ignore", divide="ignore"):
                norm = np.outer(kk, kk) if dvm == "full" else kk ** 2
                tot = np.where(norm > 0, tot / np.where(norm > 0, norm, 1.0), 0.0)
            return tot * (H_ells * N_ells) * (1 / nmesh.prod())

        if sn_mode == "ana":
            SN3 = s3_ana()
        elif sn_mode == "fft":
            SN3 = s3_fft()
        elif sn_mode == "both":
            SN3 = s3_ana()
            SN3_fft = s3_fft()
        else:
            raise ValueError("shotnoise_mode must be either 'ana', 'fft', or 'both'.")

    # ---------------------------------------------------------- normalisation
    V = boxsize.prod()
    if correlation_mode == "auto":
        I_norm = (N_gal_a ** 3) / V ** 2
    elif tracer_type == "aab":
        I_norm = (N_gal_a ** 2 * N_gal_b) / V ** 2
    elif tracer_type == "abb":
        I_norm = (N_gal_a * N_gal_b ** 2) / V ** 2
    else:
        I_norm = (N_gal_a * N_gal_b * N_gal_c) / V ** 2
    total_shot_noise = SN0 + SN1 + SN2 + SN3
    total_res = total_res * (N_ells * H_ells / I_norm)
    with np.errstate(invalid="ignore", divide="ignore"):
        if dvm == "full":
            total_res *= V ** 2 / np.outer(k_num[:k_bins], k_num[:k_bins])
        else:
            total_res *= V ** 2 / (k_num[: len(total_res)]) ** 2
    total_res *= vol_per_cell
    total_shot_noise = total_shot_noise * (1 / I_norm)
    sn_terms = {"SN0": SN0 / I_norm, "SN1": SN1 / I_norm, "SN2": SN2 / I_norm, "SN3": SN3 / I_norm}
    if sn_mode == "both":
        sn_terms["SN3_FFT"] = SN3_fft / I_norm
    return {
        "B_sugi": total_res - total_shot_noise,
        "SN_terms": sn_terms,
        "I_norm": I_norm,
        "Shot_noise": total_shot_noise,
        "Bk_raw": total_res,
        "k_eff": k_eff,
        "nmodes": k_num,
    }
