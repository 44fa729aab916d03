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


# ---------------------------------------------------------------- survey inputs
SURVEY_PATCH = dict(ra=(170.0, 180.0), dec=(20.0, 30.0), zr=(0.45, 0.5))
COSMO = {"h": 0.6736, "Omega0": 0.3153}
NO_COMP = {"scheme": None, "name_alias": None}


@functools.lru_cache(maxsize=None)
def survey_rows(ndata, nrandoms, seed, extra_cols=False):
    """Small seeded survey catalogs (FITS-like rows) on a 10 x 10 deg patch,
    0.45 < z < 0.5; optionally with BOSS completeness columns."""
    from cosmonpc_amd import synth

    d, r = synth.lognormal_survey(ndata, nrandoms, seed=seed, omega_m=COSMO["Omega0"], ngrid=32,
                                  nz_mean=0.475, nz_sigma=0.05, **SURVEY_PATCH)
    if extra_cols:
        rng = np.random.default_rng(seed + 1)
        out = []
        for a in (d, r):
            b = np.zeros(len(a), dtype=a.dtype.descr + [("WEIGHT_SYSTOT", "f8"), ("WEIGHT_NOZ", "f8"),
                                                          ("WEIGHT_CP", "f8")])
            for nm in a.dtype.names:
                b[nm] = a[nm]
            b["WEIGHT_SYSTOT"] = 0.8 + 0.4 * rng.random(len(a))
            b["WEIGHT_NOZ"] = 1.0 + (rng.random(len(a)) < 0.1)
            b["WEIGHT_CP"] = 1.0 + (rng.random(len(a)) < 0.05)
            out.append(b)
        d, r = out
    d.setflags(write=False)
    r.setflags(write=False)
    return d, r


def oracle_survey_cats(rows, z_range=None, plan=NO_COMP, cosmo=COSMO, ns="particle"):
    """Oracle-side read catalogs for a dict key -> rows."""
    from oracle.cnpc_ref import catalog as OC

    return {k: OC.survey_catalog(v, z_range, plan, cosmo, "randoms" if k.startswith("randoms") else "data", ns)
            for k, v in rows.items()}
