"""Mesh generation on the GPU (box geometry).

``get_mesh_box`` mirrors ``mesh_generator.get_mesh_box`` (``mesh_generator.py:15-169``):
it reads the catalogs, forms N_gal and NZ = sum(W)/V per tracer (``:56-60``)
and paints one field per tracer (``pm_painter``, ``:172-224``).

A painted field is kept as the unnormalised half spectra of the plain paint
and, when interlaced, of the +0.5-cell shifted paint; the interlace combine
(``:207-218``) is fused into each estimator's spectrum preparation, which is
exactly ``r2c`` of the reference's ``rfield`` (``c2r`` of the combined
spectrum) with the estimator's own scaling and compensation applied.
"""
import logging

import numpy as np
import torch

from .catalog import read_box_catalog
from .device import PREC, get_context


class DeviceField:
    """Half spectra of one painted tracer resident in HBM."""

    def __init__(self, spec1, spec2, n, boxsize, prec, sampler, interlaced):
        self.spec1 = spec1
        self.spec2 = spec2
        self.n = n
        self.boxsize = float(boxsize)
        self.prec = prec
        self.sampler = sampler
        self.interlaced = bool(interlaced)

    def delta_k(self, ctx, comp, scale=1.0, sub0=0.0, zero_k0=False, out=None):
        """comp(k) * (scale * c(k) - sub0 [k=0]) with c = pmesh r2c(rfield)
        (forward FFT normalised by 1/N^3) of the (interlaced) painted mesh."""
        if out is None:
            out = ctx.spectrum(self.n, self.prec)
        ctx.prepare(self.spec1, self.spec2, out, self.n, self.boxsize, self.prec, self.interlaced, comp,
                    self.sampler, scale / float(self.n) ** 3, sub0, zero_k0)
        return out


def _host_f64(a):
    """A host array as a CPU FP64 tensor (a copy only when the array is not
    writable C-ordered FP64: torch.from_numpy of read-only memory is
    undefined)."""
    return torch.from_numpy(np.require(a, dtype=np.float64, requirements=["C", "W"]))


def paint_field(ctx, pos, w, n, boxsize, sampler, interlaced, prec, boxcenter=None):
    """Paint host or device particles and return a DeviceField."""
    dev = ctx.device
    pos_d = pos if torch.is_tensor(pos) else _host_f64(pos).to(dev)
    w_d = w if torch.is_tensor(w) else _host_f64(w).to(dev)
    m0 = ctx.mesh(n, prec)
    m1 = ctx.mesh(n, prec) if interlaced else None
    ctx.paint(pos_d, w_d, n, boxsize, sampler, interlaced, prec, m0, m1, boxcenter)
    s1 = ctx.spectrum(n, prec)
    ctx.r2c(m0, s1, n, prec)
    s2 = None
    if interlaced:
        s2 = ctx.spectrum(n, prec)
        ctx.r2c(m1, s2, n, prec)
    del m0, m1
    return DeviceField(s1, s2, n, boxsize, prec, sampler, interlaced)


def get_mesh_box(catalogs, correlation_mode, nmesh, geometry, boxsize, sampler, interlaced, column_names,
                 comm=None, statistic="pk", tracer_type="ab", apply_rsd=False, para_cosmo=None,
                 redshift_box=None, los=None, precision="fp64", arrays=None):
    """Returns (mesh_attrs, field_a, field_b[, field_c]) like the reference.

    ``arrays`` may map 'data_a'/'data_b'/'data_c' to (Position, WEIGHT)
    arrays instead of reading ``catalogs`` from disk."""
    ctx = get_context()
    prec = PREC[precision]
    n = int(nmesh[0])
    L = float(boxsize[0])
    V = float(np.prod(np.asarray(boxsize, dtype=float)))
    tracers = ["a"]
    if correlation_mode == "cross":
        tracers.append("b")
    if tracer_type == "abc":
        tracers.append("c")
    attrs = {}
    fields = {}
    for t in tracers:
        key = "data_" + t
        if arrays is not None and key in arrays:
            pos, w = arrays[key]
        else:
            pos, w = read_box_catalog(catalogs[key], column_names, apply_rsd=apply_rsd, para_cosmo=para_cosmo,
                                      boxsize=boxsize, redshift_box=redshift_box, los=los)
        attrs["N_gal_" + t] = len(w)
        attrs["NZ_" + t] = np.sum(w) / V
        fields[t] = paint_field(ctx, pos, w, n, L, sampler, interlaced, prec)
        logging.info(f"Mesh for tracer_{t} generated.")
    ordered = {}
    for k in ("NZ_a", "NZ_b", "N_gal_a", "N_gal_b", "NZ_c", "N_gal_c"):
        if k in attrs:
            ordered[k] = attrs[k]
    ordered.update({"interlaced": interlaced, "sampler": sampler})
    fa, fb, fc = fields.get("a"), fields.get("b"), fields.get("c")
    if statistic == "pk":
        return ordered, fa, fb
    return ordered, fa, fb, fc
