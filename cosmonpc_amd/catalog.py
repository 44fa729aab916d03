"""Catalog ingestion (``catalog_processor.py``), single process.

Box geometry: a ``.npy`` 2-D array whose columns are named by
``column_names`` (``x``, ``y``, ``z`` required, ``w`` optional weight;
``catalog_reader`` box branch, ``catalog_processor.py:293-449``).

Survey geometry: FITS binary tables (``:481-563``) read by ``cosmonpc_amd.fits``
-- redshift cut, completeness weights (``add_completeness_weight``,
``:179-282``), RA/DEC/Z -> comoving Cartesian positions on the GPU
(``ra_dec_z_to_xyz``, ``:589-645``) and the box centre from the first randoms
catalog (``:565-574``).  HDF5 needs h5py, which is not part of this framework.
"""
import logging
import os

import numpy as np

from . import fits

SUPPORTED = {"npy", "fits", "h5", "hdf5"}


def _ext(path):
    p = path[0] if isinstance(path, list) else path
    return p.split(".")[-1].lower() if p else None


def _efunc(z, omega_m):
    """E(z) of flat LCDM without radiation (astropy FlatLambdaCDM, Tcmb0 = 0)."""
    return float(np.sqrt(omega_m * (1.0 + z) ** 3 + (1.0 - omega_m)))


def add_rsd(pos, vel, h, om0, redshift_box, los, box_length):
    """Box branch of ``add_rsd`` (``catalog_processor.py:648-693``): shift by
    (1+z)/(100 E(z)) * v along the unit LOS, then wrap into the box."""
    rsd_factor = (1 + redshift_box) / (_efunc(redshift_box, om0) * 100)
    los = np.array(los)
    if np.linalg.norm(los) != 1:
        raise ValueError("LOS must be a 3d unit vector")
    pos = pos + rsd_factor * los * vel
    if box_length is None:
        raise ValueError("For box-like geometry, box_length must be provided")
    return pos % box_length


def read_box_catalog(path, column_names, apply_rsd=False, para_cosmo=None, boxsize=None, redshift_box=None,
                     los=None):
    """Returns (Position (Np,3) float64, WEIGHT (Np,) float64)."""
    ext = _ext(path)
    if ext not in SUPPORTED:
        raise ValueError(f"Unsupported data file type: {ext}")
    if ext in ("h5", "hdf5"):
        raise NotImplementedError(f"box-like catalogs in .{ext} format need h5py, which is not part of cosmonpc_amd")
    if ext == "fits":
        # catalog_processor.py:388-449: X/Y/Z (or x/y/z), optional RSD from
        # VX/VY/VZ (or vx/vy/vz), weight column 'w'
        arr = read_fits_rows(path, column_names)
        names = arr.dtype.names
        pos = np.zeros((len(arr), 3))
        for axes in (("X", "Y", "Z"), ("x", "y", "z")):
            if all(a in names for a in axes):
                pos = np.column_stack([np.asarray(arr[a], dtype=np.float64) for a in axes])
                logging.info(f"Using {axes} as the position columns")
                break
        if apply_rsd:
            vel = None
            for vaxes in (("VX", "VY", "VZ"), ("vx", "vy", "vz")):
                if all(a in names for a in vaxes):
                    vel = np.column_stack([np.asarray(arr[a], dtype=np.float64) for a in vaxes])
                    break
            if vel is None:
                raise ValueError("apply_rsd needs velocity columns VX/VY/VZ or vx/vy/vz")
            pos = add_rsd(pos, vel, para_cosmo["h"], para_cosmo["Omega0"], redshift_box, los,
                          None if boxsize is None else boxsize[0])
        w = np.asarray(arr["w"], dtype=np.float64) if "w" in names else np.ones(len(arr))
        return np.ascontiguousarray(pos, dtype=np.float64), np.ascontiguousarray(w)
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


# ---------------------------------------------------------------------- survey
def add_completeness_weight(arr, comp_weight_plan, catalog_type):
    """``add_completeness_weight`` (``catalog_processor.py:179-282``)."""
    scheme = comp_weight_plan["scheme"]
    alias = comp_weight_plan["name_alias"]
    assert catalog_type in ["data", "randoms"], "catalog_type must be 'data' or 'randoms'"
    assert scheme in [None, "boss", "eboss", "desi"]
    names = arr.dtype.names
    if "WEIGHT" in names:
        logging.info("Using existing WEIGHT column")
        return arr, np.asarray(arr["WEIGHT"], dtype=np.float64)
    if alias is None and scheme is None:
        logging.info("No completeness weight scheme or alias provided. Setting completeness weight to 1.0")
        return arr, np.ones(len(arr))
    if alias is not None and scheme is None:
        if alias in names:
            logging.info(f"Using {alias} as the completeness weight")
            return arr, np.asarray(arr[alias], dtype=np.float64)
        raise ValueError(f"Alias {alias} not found in catalog columns.")
    if alias is None and scheme is not None:
        if scheme == "boss":
            if catalog_type == "data":
                w = arr["WEIGHT_SYSTOT"] * (arr["WEIGHT_NOZ"] + arr["WEIGHT_CP"] - 1.0)
            else:
                w = np.ones(len(arr))
            logging.info(f"Using BOSS-like completeness weight for {catalog_type}")
            return arr, np.asarray(w, dtype=np.float64)
        if scheme == "eboss":
            if catalog_type == "data":
                w = arr["WEIGHT_SYSTOT"] * arr["WEIGHT_NOZ"] * arr["WEIGHT_CP"]
            else:
                w = np.ones(len(arr))
            logging.info(f"Using eBOSS-like completeness weight for {catalog_type}")
            return arr, np.asarray(w, dtype=np.float64)
        raise ValueError("DESI completeness weight scheme not yet implemented.")
    raise ValueError("Cannot specify both a scheme and an alias for completeness weight.")


class SurveyCatalog:
    """One survey catalog resident in HBM: Position (Np, 3), WEIGHT,
    WEIGHT_FKP and NZ (float64 device tensors)."""

    def __init__(self, pos, weight, weight_fkp, nz):
        self.pos = pos
        self.weight = weight
        self.weight_fkp = weight_fkp
        self.nz = nz

    def __len__(self):
        return int(self.weight.shape[0])

    def take(self, index):
        """Sub-catalog of the rows ``index`` (device index tensor or slice)."""
        return SurveyCatalog(self.pos[index].contiguous(), self.weight[index].contiguous(),
                             self.weight_fkp[index].contiguous(), self.nz[index].contiguous())


def read_fits_rows(path, column_names):
    """``fits_reader`` (``catalog_processor.py:130-176``) for one process:
    the requested columns that exist, concatenated over a list of files."""
    files = [path] if isinstance(path, str) else list(path)
    parts = [fits.read_table(f, column_names) for f in files]
    return parts[0] if len(parts) == 1 else np.concatenate(parts)


def read_survey_catalog(ctx, catalog, column_names, z_range, comp_weight_plan, para_cosmo, catalog_type,
                        normalization_scheme="particle", boxcenter=None, rows=None):
    """Survey branch of ``catalog_reader`` (``catalog_processor.py:481-563``).

    ``rows`` may hold the structured array a FITS read would return (then
    ``catalog`` is only used for its extension).  Returns the SurveyCatalog,
    plus the box centre when ``boxcenter`` is None."""
    import torch

    ext = _ext(catalog) if catalog is not None else "fits"
    if ext not in SUPPORTED:
        raise ValueError(f"Unsupported data file type: {ext}")
    if ext == "npy":
        raise NotImplementedError("Survey-like geometry with .npy file is not yet implemented.")
    if ext != "fits":
        raise NotImplementedError(f"survey-like catalogs in .{ext} format need a reader that is not part of cosmonpc_amd")
    assert "Z" in column_names, \
        "For survey-like geometry with .fits file, column_names must include 'Z' for redshift slice."
    arr = rows if rows is not None else read_fits_rows(catalog, column_names)
    if z_range is not None:
        z_min, z_max = z_range
        arr = arr[(arr["Z"] > z_min) & (arr["Z"] < z_max)]
    logging.info(f"Total objects after applying z_range cut: {len(arr)}")
    arr, weight = add_completeness_weight(arr, comp_weight_plan, catalog_type)
    if not (para_cosmo is not None and "h" in para_cosmo and "Omega0" in para_cosmo):
        raise NotImplementedError("the default Planck18 cosmology is not available: give cosmology 'h' and 'Omega0'")
    names = arr.dtype.names
    if "WEIGHT_FKP" not in names:
        raise ValueError("WEIGHT_FKP column is missing in the catalog.")
    if "NZ" in names:
        nz = arr["NZ"]
    elif "NX" in names:
        nz = arr["NX"]
        logging.info("Using NX column as NZ for DESI-like catalog.")
    else:
        if normalization_scheme == "particle":
            raise ValueError("NZ column is missing in the catalog, which is required for particle normalization.")
        nz = np.ones(len(arr))
        logging.info("NZ column does not exist in the catalog, Setting NZ to 1.0")
    dev = ctx.device

    def col(a):
        return torch.from_numpy(np.require(a, dtype=np.float64, requirements=["C", "W"])).to(dev)

    pos = ctx.radecz_to_xyz(col(arr["RA"]), col(arr["DEC"]), col(arr["Z"]), float(para_cosmo["Omega0"]))
    cat = SurveyCatalog(pos, col(weight), col(arr["WEIGHT_FKP"]), col(nz))
    if boxcenter is None:
        if len(cat) == 0:
            raise ValueError("empty catalog: the box centre is undefined")
        lo = torch.amin(pos, dim=0).cpu().numpy()
        hi = torch.amax(pos, dim=0).cpu().numpy()
        boxcenter = 0.5 * (lo + hi)
        logging.info(f"Calculated boxcenter: {boxcenter}")
        return cat, boxcenter
    logging.info(f"Using provided boxcenter: {boxcenter}")
    return cat
