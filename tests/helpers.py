"""Shared test inputs and comparison helpers."""
import functools

import numpy as np


@functools.lru_cache(maxsize=None)
def catalog(nparticles, boxsize, seed, ngrid=64):
    from cosmonpc_amd import synth

    pos = synth.lognormal_box(nparticles, boxsize=boxsize, ngrid=ngrid, seed=seed)
    pos.setflags(write=False)
    return pos


def weights(n, seed=None):
    if seed is None:
        return np.ones(n)
    rng = np.random.default_rng(seed)
    return 0.5 + rng.random(n)


def rel_err(a, b):
    """max_i |a_i - b_i| / |b_i| over bins with b_i != 0."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    m = np.abs(b) > 0
    return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m])))


def rel_to_max(a, b):
    """max_i |a_i - b_i| / max_i |b_i| (for bins that cross zero)."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def pk_attrs(nmesh, boxsize, k_min, k_max, k_bins, poles, sampler, interlaced, rsd=(0, 0, 1)):
    return dict(poles=list(poles), boxsize=[boxsize] * 3, nmesh=[nmesh] * 3, k_min=k_min, k_max=k_max,
                k_bins=k_bins, rsd=list(rsd), sampler=sampler, interlaced=interlaced)


def bk_attrs(nmesh, boxsize, k_min, k_max, k_bins, angu, dvm, sampler, interlaced, sn="ana", **extra):
    d = dict(boxsize=[boxsize] * 3, nmesh=[nmesh] * 3, k_min=k_min, k_max=k_max, k_bins=k_bins,
             angu_config=list(angu), data_vector_mode=dvm, tracer_type=extra.pop("tracer_type", "aaa"),
             shotnoise_mode=sn, sampler=sampler, interlaced=interlaced,
             block_size=extra.pop("block_size", "full" if dvm == "full" else 1))
    d.update(extra)
    return d
