"""Config validation of the ``run_stats`` boundary.

Same checks, exception types and messages as the reference's
``param_helper.py`` (``validate_tracer`` :4-16, ``validate_poles`` :19-34,
``validate_sugi_poles`` :37-59, ``Cubic_Check`` :62-69,
``validate_boolean_fields`` :72-87, ``validate_config`` :90-136,
``catalog_check`` :139-165), so a caller sees identical failures before any
GPU work starts.
"""
import numpy as np

TRACER_TYPES = ("aaa", "aab", "abb", "abc")
SAMPLERS = ("ngp", "cic", "pcs", "tsc")
SHOTNOISE_MODES = ("ana", "fft", "both")
HIGH_ORDER_MODES = ("default", "fast", "compare")
HIGH_ORDER_POLES = [[0, 2, 4], [0, 2, 4, 6], [0, 2, 4, 6, 8]]
BOOL_FIELDS = ("interlaced", "compensation", "apply_rsd", "use_parent_dir")


def validate_tracer(tracer_type, correlation_mode):
    assert correlation_mode in ("auto", "cross"), "correlation_mode must be either 'auto' or 'cross'"
    assert tracer_type in TRACER_TYPES, "tracer_type must be one of 'aaa', 'aab', 'abb', or 'abc'"
    if correlation_mode == "auto":
        assert tracer_type == "aaa", "For auto-correlation, tracer_type must be 'aaa'"


def validate_poles(poles):
    """Non-negative ints, no duplicates, ascending."""
    if not all(isinstance(p, int) and p >= 0 for p in poles):
        raise ValueError("All elements in 'poles' must be non-negative integers.")
    if len(set(poles)) != len(poles):
        raise ValueError("'poles' contains duplicate values.")
    if len(poles) > 1 and list(poles) != sorted(poles):
        raise ValueError("'poles' must be sorted in ascending order if it contains more than one value.")


def validate_sugi_poles(angu_config, geometry):
    """ell1 >= ell2, even sum, even L, triangle rule, known geometry."""
    l1, l2, L = angu_config
    if l1 < l2:
        raise ValueError("In 'angu_config', ell1 must be greater than or equal to ell2.")
    if (l1 + l2 + L) % 2:
        raise ValueError("The sum of ell1, ell2, and L in 'angu_config' must be even.")
    if L % 2:
        raise ValueError("L must be even.")
    if L < l1 - l2 or L > l1 + l2:
        raise ValueError("The values in 'angu_config' do not satisfy the triangle condition.")
    if geometry not in ("box-like", "survey-like"):
        raise ValueError("geometry must be either 'box-like' or 'survey-like'.")


def Cubic_Check(value, value_name, value_type):
    assert isinstance(value, (list, tuple)) and len(value) == 3, \
        f"{value_name} must be a list or tuple of three elements"
    # the message is only formed on failure; for a tuple of types this raises
    # AttributeError, as the reference does (param_helper.py:66-68)
    assert all(isinstance(x, value_type) and x > 0 for x in value), \
        f"All elements in {value_name} must be positive {value_type.__name__}s"
    assert len(set(value)) == 1, f"All elements in {value_name} must be equal"


def validate_boolean_fields(config):
    for field in BOOL_FIELDS:
        if field in config and not isinstance(config[field], (bool, np.bool_)):
            raise TypeError(
                f"Config field '{field}' must be a boolean (True/False), got "
                f"{type(config[field]).__name__}: {config[field]!r}")


def validate_config(config):
    """Normalises aliases in place, then validates (``param_helper.py:90-136``)."""
    if "shotnoise-mode" in config and "shotnoise_mode" not in config:
        config["shotnoise_mode"] = config.pop("shotnoise-mode")
    if "high_order_mode" in config:
        config.pop("use_fast_mode", None)
    elif "use_fast_mode" in config:
        config["high_order_mode"] = "fast" if config.pop("use_fast_mode") else "default"
    else:
        config["high_order_mode"] = "default"

    assert config["sampler"] in SAMPLERS, "sampler should be 'ngp', 'cic', 'pcs', or 'tsc'"
    Cubic_Check(config["nmesh"], "nmesh", int)
    Cubic_Check(config["boxsize"], "boxsize", (float, int))

    if config.setdefault("shotnoise_mode", "ana") not in SHOTNOISE_MODES:
        raise ValueError("shotnoise_mode must be either 'ana', 'fft', or 'both'")
    hom = config["high_order_mode"]
    if hom not in HIGH_ORDER_MODES:
        raise ValueError("high_order_mode must be one of 'default', 'fast', or 'compare'")
    if hom in ("fast", "compare"):
        if config.get("statistic") != "pk" or config.get("geometry") != "survey-like":
            raise ValueError("high_order_mode 'fast' and 'compare' are only supported for survey-like pk measurements")
        if config.get("correlation_mode") != "auto":
            raise ValueError("high_order_mode 'fast' and 'compare' are only supported for auto-correlation")
        if config.get("poles") not in HIGH_ORDER_POLES:
            raise ValueError(f"high_order_mode '{hom}' only supports poles {HIGH_ORDER_POLES}")
    validate_boolean_fields(config)


def catalog_check(catalogs, geometry, correlation_mode, statistic, tracer_type=None):
    assert catalogs["data_a"] is not None, "data_a catalog must be provided"
    if statistic in ("pk", "bk_sco", "bk_sugi"):
        survey = geometry == "survey-like"
        if correlation_mode == "auto":
            if survey:
                assert catalogs["randoms_a"] is not None, \
                    "randoms_a catalog must be provided for survey-like auto-correlation"
        elif correlation_mode == "cross":
            assert catalogs["data_b"] is not None, "data_b catalog must be provided for cross-correlation"
            if survey:
                assert catalogs["randoms_a"] is not None, \
                    "randoms_a catalog must be provided for survey-like cross-correlation"
                assert catalogs["randoms_b"] is not None, \
                    "randoms_b catalog must be provided for survey-like cross-correlation"
    if tracer_type == "abc" and statistic in ("bk_sco", "bk_sugi"):
        assert catalogs.get("data_c") is not None, "data_c catalog must be provided for tracer_type 'abc' in bispectrum"
        if geometry == "survey-like":
            assert catalogs.get("randoms_c") is not None, \
                "randoms_c catalog must be provided for tracer_type 'abc' in survey-like bispectrum"
