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
from .estimators import bk_box_signals, signal_group_key
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


def run_mocks(configs, catalogs, batch=16, skip_existing=True, arrays=None):
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
    t_paint = t_est = 0.0
    chunks = [todo[c0:c0 + batch] for c0 in range(0, len(todo), batch)]
    main = torch.cuda.current_stream(ctx.device)  # the context's stream
    copy_stream = torch.cuda.Stream(ctx.device)
    pool = ThreadPoolExecutor(max_workers=1)
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
            t_p, t_e = _measure_chunk(ctx, configs, chunk, pos_l, w_l, off, pos_d, w_d, n, L, V, prec, sampler,
                                      interlaced, results, t0)
            del pos_d, w_d
            t_paint += t_p
            t_est += t_e
    finally:
        pool.shutdown(wait=True)
    LAST_TIMINGS.update({"paint_fft_s": t_paint, "estimators_s": t_est})
    return results


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
    groups = {}
    for i, c in enumerate(configs):
        key = signal_group_key(c)
        if key is not None:
            groups.setdefault(key, []).append(i)
    shared = [g for g in groups.values() if len(g) > 1]
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
