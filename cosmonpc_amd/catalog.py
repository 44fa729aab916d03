"""Catalog ingestion (``catalog_processor.py``), single process.

Box geometry: a ``.npy`` 2-D array whose columns are named by
``column_names`` (``x``, ``y``, ``z`` required, ``w`` optional weight), or
FITS / HDF5 tables of named columns (X/Y/Z or x/y/z, optional velocities
for RSD, ``w``; ``catalog_reader`` box branch, ``catalog_processor.py:293-449``;
HDF5 through ``cosmonpc_amd.hdf5``, no h5py).

Survey geometry: FITS binary tables (``:481-563``) read by ``cosmonpc_amd.fits``
-- redshift cut, completeness weights (``add_completeness_weight``,
``:179-282``), RA/DEC/Z -> comoving Cartesian positions on the GPU
(``ra_dec_z_to_xyz``, ``:589-645``: flat LCDM from the config's h and
Omega0, else the default Planck18 with photons and neutrinos) and the box
centre from the first randoms catalog (``:565-574``).  Survey catalogs in
HDF5 are not read by the reference either (its survey branch handles FITS
only).
"""
import logging
import os

import numpy as np

from . import fits, hdf5

SUPPORTED = {"npy", "fits", "h5", "hdf5"}


def _ext(path):
    p = path[0] if isinstance(path, list) else path
    return p.split(".")[-1].lower() if p else None


# Planck18 as astropy defines it (Planck 2018 results VI, Table 2,
# TT,TE,EE+lowE+lensing+BAO): flat, H0 = 67.66 km/s/Mpc, Om0 = 0.30966,
# Tcmb0 = 2.7255 K, Neff = 3.046, neutrino masses (0, 0, 0.06) eV
PLANCK18 = {"H0": 67.66, "Om0": 0.30966, "Tcmb0": 2.7255, "Neff": 3.046, "m_nu": (0.0, 0.0, 0.06)}


def flrw_params(H0, Om0, Tcmb0, Neff, m_nu):
    """(cosmo[9], c/H0 in Mpc) of a flat FLRW cosmology with photons and
    neutrinos for ``cnpc_radecz_to_xyz_flrw``.  Omega_gamma = 4 sigma_SB
    Tcmb0^4 / c^3 / rho_crit; neutrinos: floor(Neff) species sharing Neff,
    T_nu0 = (4/11)^(1/3) Tcmb0, the massive ones through the fitting
    function of Komatsu et al. 2011 (eq. 26, with astropy's constants),
    Omega_de = 1 - Omega_m - Omega_r(z = 0).  SI constants: CODATA 2018 and
    the IAU 2015 parsec."""
    c, G = 299792458.0, 6.67430e-11
    sigma_sb, k_b, ev = 5.670374419e-8, 1.380649e-23, 1.602176634e-19
    mpc = 3.0856775814913673e22
    h0 = H0 * 1e3 / mpc
    rho_crit = 3.0 * h0 ** 2 / (8.0 * np.pi * G)
    og = 4.0 * sigma_sb * Tcmb0 ** 4 / c ** 3 / rho_crit
    nspecies = int(np.floor(Neff))
    nupre = 0.22710731766 * Neff / nspecies  # 7/8 (4/11)^(4/3) per species
    massive = [m for m in m_nu if m > 0]
    if len(m_nu) != nspecies or len(massive) > 3:
        raise ValueError("one neutrino mass per species (at most 3 massive)")
    t_nu = 0.7137658555036082 * Tcmb0  # (4/11)^(1/3)
    y = [m * ev / (k_b * t_nu) for m in massive]
    nless = nspecies - len(massive)
    nu0 = nupre * (sum((1.0 + (0.3173 * v) ** 1.83) ** 0.54644808743 for v in y) + nless)
    ode = 1.0 - Om0 - og * (1.0 + nu0)
    cosmo = [Om0, og, ode, nupre, float(nless), float(len(massive))] + y + [0.0] * (3 - len(y))
    return cosmo, c / 1e3 / H0


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
    if ext in ("fits", "h5", "hdf5"):
        # catalog_processor.py:388-449: X/Y/Z (or x/y/z), optional RSD from
        # VX/VY/VZ (or vx/vy/vz), weight column 'w'.  HDF5 rows hold every
        # requested column (h5_reader, :16-88), present in the file or not
        arr = read_fits_rows(path, column_names) if ext == "fits" else hdf5.read_columns(path, column_names)
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
        raise NotImplementedError(f"survey-like catalogs in .{ext} format are not supported (the reference's "
                                  "survey branch reads FITS only)")
    assert "Z" in column_names, \
        "For survey-like geometry with .fits file, column_names must include 'Z' for redshift slice."
    arr = rows if rows is not None else read_fits_rows(catalog, column_names)
    if z_range is not None:
        z_min, z_max = z_range
        arr = arr[(arr["Z"] > z_min) & (arr["Z"] < z_max)]
    logging.info(f"Total objects after applying z_range cut: {len(arr)}")
    arr, weight = add_completeness_weight(arr, comp_weight_plan, catalog_type)
    # 'h' and 'Omega0' given: FlatLambdaCDM(H0 = 100 h, Om0) without
    # radiation; otherwise the default Planck18, its distances in Mpc times
    # the config's h (catalog_processor.py:609-632: a missing config or 'h'
    # fails there as here, on para_cosmo["h"])
    planck = not (para_cosmo is not None and "h" in para_cosmo and "Omega0" in para_cosmo)
    if planck:
        h_cfg = para_cosmo["h"]
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

    if planck:
        logging.info("Using default Planck18 cosmology")
        p18 = PLANCK18
        cosmo, dh = flrw_params(p18["H0"], p18["Om0"], p18["Tcmb0"], p18["Neff"], p18["m_nu"])
        pos = ctx.radecz_to_xyz_flrw(col(arr["RA"]), col(arr["DEC"]), col(arr["Z"]), cosmo, dh * float(h_cfg))
    else:
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
