"""Survey catalog processing restated in numpy/scipy (oracle; test
infrastructure only).

Follows ``catalog_reader``'s survey branch (``catalog_processor.py:481-563``),
``add_completeness_weight`` (``:179-282``) and ``ra_dec_z_to_xyz``
(``:589-645``) on an already-read structured array (the rows ``fits_reader``
returns).  The comoving distance of astropy's ``FlatLambdaCDM(H0=100h, Om0)``
(no radiation: Tcmb0 defaults to 0) is restated as
chi(z) = (c / 100 km/s/Mpc) * int_0^z dz' / sqrt(Om (1+z')^3 + 1 - Om) in Mpc/h,
integrated with ``scipy.integrate.quad``; astropy itself is not importable
here.  Pinned in ``tests/test_oracle_survey.py`` by the Einstein-de Sitter and
empty-universe closed forms.
"""
import numpy as np
from scipy import integrate

C_KMS = 299792.458


def comoving_distance(z, omega_m):
    """chi(z) in Mpc/h for flat LCDM without radiation."""
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))

    def f(zp):
        return 1.0 / np.sqrt(omega_m * (1.0 + zp) ** 3 + (1.0 - omega_m))

    out = np.array([integrate.quad(f, 0.0, zz, epsabs=0.0, epsrel=1e-13, limit=200)[0] for zz in z])
    return out * (C_KMS / 100.0)


def ra_dec_z_to_xyz(arr, para_cosmo):
    """``catalog_processor.py:589-645`` with a FlatLambdaCDM cosmology."""
    ra = np.deg2rad(arr["RA"])
    dec = np.deg2rad(arr["DEC"])
    chi = comoving_distance(arr["Z"], para_cosmo["Omega0"])
    return np.column_stack([chi * np.cos(dec) * np.cos(ra), chi * np.cos(dec) * np.sin(ra), chi * np.sin(dec)])


def completeness_weight(arr, comp_weight_plan, catalog_type):
    """``add_completeness_weight`` (``catalog_processor.py:179-282``)."""
    scheme, alias = comp_weight_plan["scheme"], comp_weight_plan["name_alias"]
    names = arr.dtype.names
    if "WEIGHT" in names:
        return np.asarray(arr["WEIGHT"], dtype=np.float64)
    if alias is None and scheme is None:
        return np.ones(len(arr))
    if alias is not None and scheme is None:
        if alias in names:
            return np.asarray(arr[alias], dtype=np.float64)
        raise ValueError(f"Alias {alias} not found in catalog columns.")
    if alias is None and scheme == "boss":
        if catalog_type == "data":
            return arr["WEIGHT_SYSTOT"] * (arr["WEIGHT_NOZ"] + arr["WEIGHT_CP"] - 1.0)
        return np.ones(len(arr))
    if alias is None and scheme == "eboss":
        if catalog_type == "data":
            return arr["WEIGHT_SYSTOT"] * arr["WEIGHT_NOZ"] * arr["WEIGHT_CP"]
        return np.ones(len(arr))
    if alias is None and scheme == "desi":
        raise ValueError("DESI completeness weight scheme not yet implemented.")
    raise ValueError("Cannot specify both a scheme and an alias for completeness weight.")


def survey_catalog(rows, z_range, comp_weight_plan, para_cosmo, catalog_type, normalization_scheme="particle"):
    """Survey ``catalog_reader`` on read rows: returns the structured array
    with Position, WEIGHT, WEIGHT_FKP, NZ (``catalog_processor.py:481-563``)."""
    arr = rows
    if z_range is not None:
        arr = arr[(arr["Z"] > z_range[0]) & (arr["Z"] < z_range[1])]
    w = completeness_weight(arr, comp_weight_plan, catalog_type)
    pos = ra_dec_z_to_xyz(arr, para_cosmo)
    out = np.zeros(len(arr), dtype=[("Position", "f8", (3,)), ("WEIGHT", "f8"), ("WEIGHT_FKP", "f8"), ("NZ", "f8")])
    out["Position"] = pos
    out["WEIGHT"] = w
    if "WEIGHT_FKP" not in arr.dtype.names:
        raise ValueError("WEIGHT_FKP column is missing in the catalog.")
    out["WEIGHT_FKP"] = arr["WEIGHT_FKP"]
    if "NZ" in arr.dtype.names:
        out["NZ"] = arr["NZ"]
    elif "NX" in arr.dtype.names:
        out["NZ"] = arr["NX"]
    elif normalization_scheme == "particle":
        raise ValueError("NZ column is missing in the catalog, which is required for particle normalization.")
    else:
        out["NZ"] = 1.0
    return out
