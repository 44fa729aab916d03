"""``run_stats`` / ``run_task``: the drop-in boundary (``task_executor.py:11-343``).

Same config dict, same validation and errors, same output file names and the
same pickled ``.npy`` result dict (results merged with ``stat_attrs``).  The
numeric work runs on one MI355X through libcnpc_hip.so; there are no MPI
ranks (the reference's rank-0 logic is the only rank here).
"""
import gc
import logging
import os
import time

import numpy as np

from .estimators import calculate_bk_sugi_box, calculate_power_spectrum_box
from .mesh import get_mesh_box
from .survey import (calculate_bk_sugi_survey, calculate_power_spectrum_survey, get_mesh_bk_survey,
                     get_mesh_pk_survey)
from .param_helper import (catalog_check, validate_config, validate_poles, validate_sugi_poles, validate_tracer)


def _save(path, res):
    np.save(path, res, allow_pickle=True)
    logging.info(f"result saved to {path}")


def result_path(statistic, correlation_mode, catalogs, **kwargs):
    """Output file of a task (``task_executor.py:130-145``, ``:298-314``)."""
    out_dir = kwargs.get("output_dir")
    name_a = os.path.splitext(os.path.basename(catalogs["data_a"]))[0]
    if statistic == "pk":
        if correlation_mode == "auto":
            return os.path.join(out_dir, f"pk_res_{correlation_mode}_{name_a}.npy")
        name_b = os.path.splitext(os.path.basename(catalogs["data_b"]))[0]
        return os.path.join(out_dir, f"pk_res_{correlation_mode}_{name_a}_{name_b}.npy")
    parent = os.path.basename(os.path.dirname(catalogs["data_a"]))
    if kwargs.get("use_parent_dir", True):
        name_a = parent + "_" + name_a
    l1, l2, L = kwargs["angu_config"]
    dvm = kwargs.get("data_vector_mode", "diagonal")
    return os.path.join(out_dir, f"bk_sugi_res_{dvm}_{l1}{l2}{L}_{correlation_mode}_{kwargs['tracer_type']}_{name_a}.npy")


def check_task(statistic, correlation_mode, geometry, **kwargs):
    """The checks ``run_task`` makes before any catalog or mesh work, in the
    reference's order and with its messages: pk ``task_executor.py:34-42``,
    ``:68-72``; bk_sugi ``:160-165``, ``:189-201``, ``:222``."""
    k_max, nmesh, boxsize = kwargs["k_max"], kwargs["nmesh"], kwargs["boxsize"]
    nyquist_freq = np.pi * nmesh[0] / boxsize[0]
    if statistic == "pk":
        if k_max >= nyquist_freq:
            raise ValueError(
                f"k_max {k_max} is larger than the Nyquist frequency {nyquist_freq}. Please choose a smaller k_max.")
        validate_poles(kwargs["poles"])
        poles = kwargs["poles"]
        if geometry != "survey-like" and all(p % 2 == 1 for p in poles):
            raise ValueError(
                f"All poles in {poles} are odd. Please choose at least one even pole, since odd poles are "
                "defaultly set to 0 for periodic boxes.")
    elif statistic == "bk_sugi":
        dvm = kwargs.get("data_vector_mode", "diagonal")
        sn_mode = kwargs.get("shotnoise_mode", "ana")
        if dvm not in ("diagonal", "full"):
            raise ValueError("data_vector_mode must be either 'diagonal' or 'full'")
        if sn_mode not in ("ana", "fft", "both"):
            raise ValueError("shotnoise_mode must be either 'ana', 'fft', or 'both'")
        validate_tracer(kwargs["tracer_type"], correlation_mode)
        if 2 * k_max >= nyquist_freq:
            raise ValueError(
                f"2 * k_max = {2 * k_max} is larger than the Nyquist frequency {nyquist_freq}. "
                "Please choose a smaller k_max.")
        validate_sugi_poles(kwargs["angu_config"], geometry)


def run_task(statistic, correlation_mode, geometry, catalogs, **kwargs):
    """``task_executor.py:11-328``.  ``kwargs`` may carry ``arrays`` (catalog
    arrays instead of files), the GPU options ``precision`` /
    ``contraction``, ``fields`` = (mesh attrs, field_a) of an already
    painted box catalog (auto-correlation; used by the batched mock driver),
    ``signal`` = the bk box ``total_res`` already computed by a
    multi-configuration pass (``estimators.bk_box_signals``),
    ``precomputed`` = the device results of the estimator's device pass
    (``pk_box_device_pass`` / ``bk_box_device_pass``, replayed from a graph
    by the batched mock driver) and ``gc_collect`` (default True: the reference's gc pass after the pk
    estimator, ``task_executor.py:115-116``)."""
    logging.basicConfig(level=logging.INFO, format="%(asctime)s - %(module)s - %(levelname)s - %(message)s")
    logging.info(f"Running {correlation_mode} {statistic} task with catalogs:{catalogs},\n")
    arrays = kwargs.pop("arrays", None)
    painted = kwargs.pop("fields", None)
    signal = kwargs.pop("signal", None)  # bk box total_res from a multi-configuration pass
    # device results recorded and replayed by the batched mock driver's graph
    precomputed = kwargs.pop("precomputed", None)
    collect = kwargs.pop("gc_collect", True)  # the batched mock driver skips the per-task gc pass
    if painted is not None and (geometry != "box-like" or correlation_mode != "auto"):
        raise ValueError("pre-painted fields are supported for box-like auto-correlation only")
    precision = kwargs.get("precision", "fp64")
    nmesh, boxsize = kwargs["nmesh"], kwargs["boxsize"]
    nyquist_freq = np.pi * nmesh[0] / boxsize[0]
    logging.info(f"Nyquist frequency: {nyquist_freq}")
    t0 = time.time()

    check_task(statistic, correlation_mode, geometry, **kwargs)

    if statistic == "pk":
        if geometry == "survey-like":
            logging.info("Using survey-like geometry...")
            stat_attrs, fa, fb = get_mesh_pk_survey(
                catalogs, correlation_mode, nmesh=nmesh, geometry=geometry, column_names=kwargs["column_names"],
                boxsize=boxsize, sampler=kwargs["sampler"], interlaced=kwargs["interlaced"],
                z_range=kwargs["z_range"], comp_weight_plan=kwargs["comp_weight_plan"],
                para_cosmo=kwargs["cosmology"], normalization_scheme=kwargs["normalization_scheme"],
                alpha_scheme=kwargs.get("alpha_scheme", "pypower"), precision=precision, arrays=arrays)
        else:
            logging.info("Using box geometry...")
            if painted is not None:
                stat_attrs, fa, fb = dict(painted[0]), painted[1], None
            else:
                stat_attrs, fa, fb = get_mesh_box(
                    catalogs, correlation_mode, nmesh=kwargs["nmesh"], geometry=geometry, boxsize=kwargs["boxsize"],
                    sampler=kwargs["sampler"], interlaced=kwargs["interlaced"], column_names=kwargs["column_names"],
                    statistic=statistic, apply_rsd=kwargs["apply_rsd"], para_cosmo=kwargs["cosmology"],
                    redshift_box=kwargs["redshift_box"], los=kwargs["rsd"], precision=precision, arrays=arrays)
        stat_attrs.update(kwargs)
        stat_attrs["nyquist_freq"] = nyquist_freq
        t1 = time.time()
        logging.info(f"Time to create (FKP) overdensity field(s): {t1 - t0:.2f} seconds")
        if geometry == "survey-like":
            pk_res = calculate_power_spectrum_survey(stat_attrs, fa, fb, correlation_mode, **kwargs)
        else:
            pk_res = calculate_power_spectrum_box(fa, fb, correlation_mode, stat_attrs, precomputed=precomputed,
                                                  **kwargs)
        del fa, fb
        if collect:
            gc.collect()
        pk_res.update(stat_attrs)
        if geometry == "box-like":
            for key in ("z_range", "scheme"):
                pk_res.pop(key, None)
        os.makedirs(kwargs.get("output_dir"), exist_ok=True)
        _save(result_path(statistic, correlation_mode, catalogs, **kwargs), pk_res)
        logging.info(f"Time to compute power spectrum: {time.time() - t1:.2f} seconds")
        return pk_res

    if statistic == "bk_sugi":
        tracer_type = kwargs["tracer_type"]
        angu_config = kwargs["angu_config"]
        dvm = kwargs.get("data_vector_mode", "diagonal")
        kwargs.setdefault("block_size", 1 if dvm == "diagonal" else "full")
        if geometry == "box-like":
            logging.info("Using box geometry...")
            if painted is not None:
                stat_attrs, fa, fb, fc = dict(painted[0]), painted[1], None, None
            else:
                stat_attrs, fa, fb, fc = get_mesh_box(
                    catalogs, correlation_mode, statistic=statistic, nmesh=kwargs["nmesh"], geometry=geometry,
                    boxsize=kwargs["boxsize"], sampler=kwargs["sampler"], interlaced=kwargs["interlaced"],
                    column_names=kwargs["column_names"], tracer_type=tracer_type, apply_rsd=kwargs["apply_rsd"],
                    para_cosmo=kwargs["cosmology"], redshift_box=kwargs["redshift_box"], los=kwargs["rsd"],
                    precision=precision, arrays=arrays)
        else:
            logging.info("Using survey-like geometry...")
            stat_attrs, fa, fb, fc = get_mesh_bk_survey(
                catalogs, correlation_mode, tracer_type=tracer_type, angu_config=angu_config, nmesh=nmesh,
                geometry=geometry, column_names=kwargs["column_names"], boxsize=boxsize,
                sampler=kwargs["sampler"], interlaced=kwargs["interlaced"], z_range=kwargs["z_range"],
                comp_weight_plan=kwargs["comp_weight_plan"], para_cosmo=kwargs["cosmology"],
                normalization_scheme=kwargs.get("normalization_scheme", "particle"),
                alpha_scheme=kwargs.get("alpha_scheme", "pypower"), precision=precision, arrays=arrays)
        stat_attrs.update(kwargs)
        stat_attrs["nyquist_freq"] = nyquist_freq
        t1 = time.time()
        logging.info(f"Time to create overdensity field(s): {t1 - t0:.2f} seconds")
        if geometry == "box-like":
            bk_res = calculate_bk_sugi_box(fa, fb, fc, correlation_mode, stat_attrs, signal=signal,
                                           precomputed=precomputed, **kwargs)
        else:
            bk_res = calculate_bk_sugi_survey(fa, fb, fc, correlation_mode, stat_attrs, **kwargs)
        del fa, fb, fc
        bk_res.update(stat_attrs)
        if geometry == "box-like":
            drop = ("z_range", "comp_weight_plan", "scheme", "alpha_scheme", "normalization_scheme")
        else:
            drop = ("redshift_box", "los", "apply_rsd")
        for key in drop:
            bk_res.pop(key, None)
        os.makedirs(kwargs.get("output_dir"), exist_ok=True)
        _save(result_path(statistic, correlation_mode, catalogs, **kwargs), bk_res)
        logging.info(f"Time to compute bispectrum: {time.time() - t1:.2f} seconds")
        return bk_res
    raise ValueError(f"unsupported statistic {statistic!r}")


def run_stats(config):
    """``task_executor.py:330-343``: validate, create output_dir, run."""
    validate_config(config)
    catalog_check(config["catalogs"], config["geometry"], config["correlation_mode"], config["statistic"])
    os.makedirs(config["output_dir"], exist_ok=True)
    return run_task(**config)
