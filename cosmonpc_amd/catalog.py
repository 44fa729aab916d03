"""Catalog ingestion for the box geometry (host side).

Restates the box branch of ``catalog_reader`` (``catalog_processor.py:293-449``)
for single-process use: a ``.npy`` 2-D array whose columns are named by
``column_names`` (``x``, ``y``, ``z`` required, ``w`` optional weight,
``:354-386``).  The HDF5 / FITS readers of the reference need h5py / fitsio,
which are not part of this framework; those extensions raise with the
reference's message for unsupported types.
"""
import os

import numpy as np

SUPPORTED = {"npy", "fits", "h5", "hdf5"}


def _ext(path):
    p = path[0] if isinstance(path, list) else path
    return p.split(".")[-1].lower() if p else None


def read_box_catalog(path, column_names):
    """Returns (Position (Np,3) float64, WEIGHT (Np,) float64)."""
    ext = _ext(path)
    if ext not in SUPPORTED:
        raise ValueError(f"Unsupported data file type: {ext}")
    if ext != "npy":
        raise NotImplementedError(f"box-like catalogs in .{ext} format need a reader that is not part of cosmonpc_amd")
    assert column_names is not None and len(column_names) >= 3, (
        "For box-like geometry with .npy file, column_names must be provided "
        "with at least 3 elements for x,y,z.")
    arr = np.load(path, mmap_mode="r", allow_pickle=False)
    idx = [column_names.index(a) for a in ("x", "y", "z")]
    pos = np.ascontiguousarray(np.column_stack([arr[:, i] for i in idx]), dtype=np.float64)
    if "w" in column_names:
        w = np.ascontiguousarray(arr[:, column_names.index("w")], dtype=np.float64)
    else:
        w = np.ones(len(arr), dtype=np.float64)
    return pos, w


def catalog_stem(path):
    return os.path.splitext(os.path.basename(path))[0]
