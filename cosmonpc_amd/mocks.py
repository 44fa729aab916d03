"""Batched box mocks: one measurement configuration over many catalogs.

A mock campaign runs the same box-like auto-correlation tasks (pk and/or
bk_sugi) on many catalogs, one ``run_stats`` call per catalog in the
reference (``task_executor.py:330-343``).  ``run_mocks`` shares the GPU work
of a chunk of catalogs instead:

* all catalogs of the chunk are painted in one launch (``cnpc_paint_batch``,
  one catalog per grid row, uneven sizes welcome);
* one batched rocFFT r2c per mesh copy transforms the whole chunk;
* each catalog's spectra then go through the same estimators and writers as
  ``run_stats`` (``run_task`` with the painted field), so every output file
  and result dict is the one ``run_stats`` would write for that catalog;
* bk_sugi configurations that share their binning and contraction settings
  (``estimators.signal_group_key``) get their signal matrices from one
  multi-configuration pass per catalog (``estimators.bk_box_signals``:
  shared shell legs, [0,0,0] out of [l,0,l]'s contraction), equal to the
  one-at-a-time results up to FP32 summation order.

Catalogs whose outputs all exist already are skipped, so an interrupted
campaign resumes where it stopped.  The next chunk's catalogs are read and
copied to the device by a prefetch thread on a copy stream while the
current chunk is measured.
"""
import logging
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from .catalog import read_box_catalog
from .device import PREC, get_context
from .estimators import (bk_box_bin_constants, bk_box_device_pass, bk_box_signals, pk_box_device_pass,
                         signal_group_key)
from .graphs import StepGraph
from .mesh import DeviceField
from .param_helper import catalog_check, validate_config
from .task_executor import check_task, result_path, run_task

MESH_KEYS = ("nmesh", "boxsize", "sampler", "interlaced", "column_names", "precision", "apply_rsd", "cosmology",
             "redshift_box", "rsd")

LAST_TIMINGS = {}


def _task_kwargs(cfg):
    return {k: v for k, v in cfg.items() if k not in ("statistic", "correlation_mode", "geometry", "catalogs")}


def _stage(chunk, base, arrays, device, stream):
    """Reads the catalogs of ``chunk`` and copies each into its slice of one
    device array pair on ``stream`` (prefetch thread).  Returns the host
    arrays, the offsets, the device arrays and the copy's completion event."""
    pos_l, w_l = [], []
    for path in chunk:
        if arrays is not None and path in arrays:
            pos, w = arrays[path]
        else:
            pos, w = read_box_catalog(path, base["column_names"], apply_rsd=base.get("apply_rsd", False),
                                      para_cosmo=base.get("cosmology"), boxsize=base["boxsize"],
                                      redshift_box=base.get("redshift_box"), los=base.get("rsd"))
        # writable C-ordered FP64 (a copy only for read-only or other
        # inputs): torch.from_numpy of a read-only array is undefined
        pos_l.append(np.require(pos, dtype=np.float64, requirements=["C", "W"]))
        w_l.append(np.require(w, dtype=np.float64, requirements=["C", "W"]))
    off = np.concatenate([[0], np.cumsum([len(w) for w in w_l])]).astype(np.int64)
    with torch.cuda.device(device), torch.cuda.stream(stream):
        pos_d = torch.empty((int(off[-1]), 3), dtype=torch.float64, device=device)
        w_d = torch.empty(int(off[-1]), dtype=torch.float64, device=device)
        for b in range(len(chunk)):
            if off[b + 1] > off[b]:
                pos_d[off[b]:off[b + 1]].copy_(torch.from_numpy(pos_l[b]))
                w_d[off[b]:off[b + 1]].copy_(torch.from_numpy(w_l[b]))
        done = torch.cuda.Event()
        done.record(stream)
    return pos_l, w_l, off, pos_d, w_d, done


def _graphable(configs):
    """The chunk pipeline can be recorded when every configuration's device
    work is catalog-independent given the painted field: pk, and bk_sugi
    with the analytic shot noise ('ana')."""
    for c in configs:
        if c["statistic"] == "bk_sugi" and c.get("shotnoise_mode", "ana") != "ana":
            return False
        if c["statistic"] not in ("pk", "bk_sugi"):
            return False
    return True


class _ChunkGraph:
    """The device work of a chunk of ``nb`` catalogs recorded once as a
    hipGraph (``StepGraph``) and replayed for every chunk: the batch paint
    from fixed capacity buffers (``cap_total`` particles, catalogs of at
    most ``cap_max``; the device offsets say which particles belong to which
    catalog), the batched r2c of both mesh copies, and each catalog's
    estimator device passes (``pk_box_device_pass``; ``bk_box_device_pass``
    per signal group, NZ read from a device array).  Before each replay the
    chunk's particles, offsets and NZ values are copied into the buffers;
    after it each catalog's results are read back and finished by the
    estimators' host code (``run_task(..., precomputed=...)``)."""

    def __init__(self, ctx, configs, groups, n, L, prec, sampler, interlaced, nb, cap_total, cap_max):
        dev = ctx.device
        self.ctx, self.configs, self.groups = ctx, configs, groups
        self.n, self.L, self.prec, self.sampler, self.interlaced = n, L, prec, sampler, interlaced
        self.nb, self.cap_total, self.cap_max = nb, cap_total, cap_max
        self.pos = torch.zeros((cap_total, 3), dtype=torch.float64, device=dev)
        self.w = torch.zeros(cap_total, dtype=torch.float64, device=dev)
        self.off = torch.zeros(nb + 1, dtype=torch.int64, device=dev)
        self.nz = torch.zeros(nb, dtype=torch.float64 if prec == PREC["fp64"] else torch.float32, device=dev)
        self.host_off = np.zeros(nb + 1, dtype=np.int64)
        self.graph = StepGraph(ctx)
        self.pending = None
        # catalog-independent binnings of each bk configuration
        self.consts = {i: bk_box_bin_constants(c, n, L, prec) for i, c in enumerate(configs)
                       if c["statistic"] == "bk_sugi"}

    def fits(self, off):
        return len(off) - 1 <= self.nb and off[-1] <= self.cap_total and np.diff(off).max(initial=0) <= self.cap_max

    def load(self, pos_d, w_d, off, nzs):
        """The chunk's particles, offsets (the slots past the chunk hold empty
        catalogs) and NZ values into the recorded buffers."""
        tot = int(off[-1])
        if tot:
            self.pos[:tot].copy_(pos_d[:tot])
            self.w[:tot].copy_(w_d[:tot])
        full = np.full(self.nb + 1, tot, dtype=np.int64)
        full[:len(off)] = off
        self.host_off[:] = full
        self.off.copy_(torch.from_numpy(full))
        nz = np.zeros(self.nb)
        nz[:len(nzs)] = nzs
        self.nz.copy_(torch.from_numpy(nz))

    def _enqueue(self):
        ctx, n, prec, nb = self.ctx, self.n, self.prec, self.nb
        m0 = ctx.mesh(n, prec, nb)
        m1 = ctx.mesh(n, prec, nb) if self.interlaced else None
        ctx.paint_batch(self.pos, self.w, self.host_off, n, self.L, self.sampler, self.interlaced, prec, m0, m1,
                        off_dev=self.off, bounds=(self.cap_total, self.cap_max))
        s1 = ctx.spectrum(n, prec, nb)
        ctx.r2c(m0, s1, n, prec, nb)
        del m0
        s2 = None
        if self.interlaced:
            s2 = ctx.spectrum(n, prec, nb)
            ctx.r2c(m1, s2, n, prec, nb)
            del m1
        slots = []
        for b in range(nb):
            field = DeviceField(s1[b], s2[b] if self.interlaced else None, n, self.L, prec, self.sampler,
                                self.interlaced)
            pk = {i: pk_box_device_pass(field, c) for i, c in enumerate(self.configs) if c["statistic"] == "pk"}
            bk = [(g, bk_box_device_pass(field, [self.configs[i] for i in g], self.nz[b])) for g in self.groups]
            slots.append((field, pk, bk))
        return slots

    def record(self, first):
        """Record the chunk pipeline; the warm-up run before the recording
        measures the loaded chunk, whose results ``first(slots)`` takes."""
        self.pending = self.graph.record(self._enqueue, first)
        return self.pending

    def replay(self):
        self.graph.replay()
        return self.pending

    def close(self):
        self.graph.close()

    def finish_slot(self, slots, b):
        """{config index: precomputed} of slot b after its device work ran."""
        field, pk, bk = slots[b]
        pre = {i: fin() for i, fin in pk.items()}
        for g, fin in bk:
            for i, r in zip(g, fin()):
                pre[i] = dict(self.consts[i], signal=r["signal"], re3=r["re3"])
        return field, pre


def run_mocks(configs, catalogs, batch=16, skip_existing=True, arrays=None, graph=True):
    """Run every config of ``configs`` (box-like, auto-correlation run_stats
    configs that agree on the mesh keys) on every catalog path in
    ``catalogs``.  ``arrays`` may map a path to (Position, WEIGHT) instead of
    reading the file.  Returns {path: [result dict per config]} for the
    catalogs that ran."""
    configs = [dict(c) for c in configs]
    catalogs = list(catalogs)
    if not configs:
        raise ValueError("run_mocks needs at least one config")
    if batch < 1:
        raise ValueError("batch must be at least 1")
    for c in configs:
        if c.get("geometry") != "box-like" or c.get("correlation_mode") != "auto":
            raise ValueError("run_mocks supports box-like auto-correlation configs only")
        c["catalogs"] = {"data_a": catalogs[0] if catalogs else "mock"}
        validate_config(c)
        catalog_check(c["catalogs"], c["geometry"], c["correlation_mode"], c["statistic"])
        check_task(c["statistic"], c["correlation_mode"], c["geometry"], **_task_kwargs(c))
    base = configs[0]
    for c in configs[1:]:
        for k in MESH_KEYS:
            if c.get(k) != base.get(k):
                raise ValueError(f"all configs of a mock run must share '{k}'")
    todo = []
    for path in catalogs:
        outs = [result_path(c["statistic"], "auto", {"data_a": path}, **_task_kwargs(c)) for c in configs]
        if skip_existing and all(os.path.exists(o) for o in outs):
            logging.info(f"mock {path}: outputs exist, skipped")
            continue
        todo.append(path)
    LAST_TIMINGS.clear()
    LAST_TIMINGS.update({"paint_fft_s": 0.0, "estimators_s": 0.0, "mocks": len(todo)})
    if not todo:
        return {}
    ctx = get_context()
    n = int(base["nmesh"][0])
    L = float(base["boxsize"][0])
    V = float(np.prod(np.asarray(base["boxsize"], dtype=float)))
    prec = PREC[base.get("precision", "fp64")]
    sampler, interlaced = base["sampler"], bool(base["interlaced"])
    results = {}
    t_paint = t_est = t_rec = t_rep = t_fin = 0.0
    chunks = [todo[c0:c0 + batch] for c0 in range(0, len(todo), batch)]
    main = torch.cuda.current_stream(ctx.device)  # the context's stream
    copy_stream = torch.cuda.Stream(ctx.device)
    pool = ThreadPoolExecutor(max_workers=1)
    groups = _signal_groups(configs, all_groups=True)
    use_graph = graph and _graphable(configs)
    cg = None
    try:
        nxt = pool.submit(_stage, chunks[0], base, arrays, ctx.device, copy_stream)
        for ci, chunk in enumerate(chunks):
            t0 = time.perf_counter()
            pos_l, w_l, off, pos_d, w_d, done = nxt.result()
            if ci + 1 < len(chunks):  # read and copy the next chunk meanwhile
                nxt = pool.submit(_stage, chunks[ci + 1], base, arrays, ctx.device, copy_stream)
            main.wait_event(done)
            # allocated on the copy stream, used on the context stream: the
            # allocator must not reuse them before the paint has read them
            pos_d.record_stream(main)
            w_d.record_stream(main)
            if not use_graph:
                t_p, t_e = _measure_chunk(ctx, configs, chunk, pos_l, w_l, off, pos_d, w_d, n, L, V, prec, sampler,
                                          interlaced, results, t0)
                t_paint += t_p
                t_est += t_e
            else:
                nzs = [np.sum(w) / V for w in w_l]
                if cg is None or not cg.fits(off):
                    if cg is not None:
                        cg.close()
                    # capacity with headroom for the chunks to come
                    cap_total = max(1, int(1.25 * off[-1]))
                    cap_max = max(1, int(1.25 * np.diff(off).max(initial=0)))
                    cg = _ChunkGraph(ctx, configs, groups, n, L, prec, sampler, interlaced, len(chunk),
                                     max(cap_total, cap_max), cap_max)
                    cg.load(pos_d, w_d, off, nzs)
                    cg.record(lambda slots: _finish_chunk(cg, slots, configs, chunk, w_l, V, interlaced, sampler,
                                                          results))
                    t_rec += time.perf_counter() - t0
                else:
                    cg.load(pos_d, w_d, off, nzs)
                    slots = cg.replay()
                    main.synchronize()
                    t1 = time.perf_counter()
                    _finish_chunk(cg, slots, configs, chunk, w_l, V, interlaced, sampler, results)
                    t2 = time.perf_counter()
                    t_rep += t1 - t0
                    t_fin += t2 - t1
            del pos_d, w_d
    finally:
        pool.shutdown(wait=True)
        if cg is not None:
            cg.close()
    if use_graph:
        LAST_TIMINGS.update({"graph": True, "first_chunk_and_recording_s": t_rec, "replays_s": t_rep,
                             "replayed_chunks_host_s": t_fin, "replayed_chunks_s": t_rep + t_fin})
    else:
        LAST_TIMINGS.update({"graph": False, "paint_fft_s": t_paint, "estimators_s": t_est})
    return results


def _signal_groups(configs, all_groups=False):
    """Indices of bk_sugi configurations that share a multi-configuration
    signal pass (``signal_group_key`` equal); with ``all_groups`` a lone
    configuration forms its own group."""
    groups = {}
    for i, c in enumerate(configs):
        key = signal_group_key(c)
        if key is not None:
            groups.setdefault(key, []).append(i)
    return [g for g in groups.values() if all_groups or len(g) > 1]


def _finish_chunk(cg, slots, configs, chunk, w_l, V, interlaced, sampler, results):
    """Each catalog's results from the chunk's device results: read back,
    finished by the estimators' host code, written like run_stats."""
    for b, path in enumerate(chunk):
        field, pre = cg.finish_slot(slots, b)
        attrs = {"NZ_a": np.sum(w_l[b]) / V, "N_gal_a": len(w_l[b]), "interlaced": interlaced, "sampler": sampler}
        res = []
        for i, c in enumerate(configs):
            cfg = dict(c)
            cfg["catalogs"] = {"data_a": path}
            cfg["fields"] = (attrs, field)
            cfg["gc_collect"] = False
            cfg["precomputed"] = pre[i]
            res.append(run_task(**cfg))
        results[path] = res


def _measure_chunk(ctx, configs, chunk, pos_l, w_l, off, pos_d, w_d, n, L, V, prec, sampler, interlaced,
                   results, t0):
    """Paint + batched r2c of one chunk, then every config on every catalog
    of it; returns (paint + FFT seconds, estimator seconds).  Waits on the
    context stream only (the copy stream may be filling the next chunk)."""
    main = torch.cuda.current_stream(ctx.device)
    nb = len(chunk)
    m0 = ctx.mesh(n, prec, nb)
    m1 = ctx.mesh(n, prec, nb) if interlaced else None
    ctx.paint_batch(pos_d, w_d, off, n, L, sampler, interlaced, prec, m0, m1)
    s1 = ctx.spectrum(n, prec, nb)
    ctx.r2c(m0, s1, n, prec, nb)
    del m0
    s2 = None
    if interlaced:
        s2 = ctx.spectrum(n, prec, nb)
        ctx.r2c(m1, s2, n, prec, nb)
        del m1
    main.synchronize()
    t1 = time.perf_counter()
    shared = _signal_groups(configs)
    for b, path in enumerate(chunk):
        field = DeviceField(s1[b], s2[b] if interlaced else None, n, L, prec, sampler, interlaced)
        attrs = {"NZ_a": np.sum(w_l[b]) / V, "N_gal_a": len(w_l[b]), "interlaced": interlaced, "sampler": sampler}
        signals = {}
        for g in shared:
            for i, sig in zip(g, bk_box_signals(field, [dict(configs[i], **attrs) for i in g])):
                signals[i] = sig
        res = []
        for i, c in enumerate(configs):
            cfg = dict(c)
            cfg["catalogs"] = {"data_a": path}
            cfg["fields"] = (attrs, field)
            cfg["gc_collect"] = False
            if i in signals:
                cfg["signal"] = signals[i]
            res.append(run_task(**cfg))
        results[path] = res
    main.synchronize()
    return t1 - t0, time.perf_counter() - t1
