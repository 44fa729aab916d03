"""cosmonpc_amd -- MI355X-native mesh-estimator hot path of CosmoNPC.

``run_stats(config)`` is the drop-in entry point (same config dict, same
outputs as ``cosmonpc.run_stats``); painting, FFTs, compensation, binning,
shell filtering and the bispectrum contractions run as gfx950 HIP kernels and
rocFFT plans in ``libcnpc_hip.so``.
"""

__version__ = "0.1.0"


def run_stats(config):
    from .task_executor import run_stats as _run

    return _run(config)


def run_task(statistic, correlation_mode, geometry, catalogs, **kwargs):
    from .task_executor import run_task as _run

    return _run(statistic, correlation_mode, geometry, catalogs, **kwargs)
