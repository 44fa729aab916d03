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
empty-universe closed forms.  The default Planck18 branch (photons, two
massless and one 0.06 eV neutrino, astropy's published FLRW algorithm) is
pinned by E(0) = 1, the no-radiation limit, the massless-neutrino ratio
7/8 (4/11)^(4/3) N_eff, the photon density 2.47e-5 h^-2 (T/2.7255 K)^4 and
the non-relativistic neutrino limit sum(m) / 93.14 eV; against astropy's own
numbers it is parity unpinned (astropy is absent).
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


# astropy's Planck18 (Planck 2018 VI, Table 2, TT,TE,EE+lowE+lensing+BAO)
P18 = dict(H0=67.66, Om0=0.30966, Tcmb0=2.7255, Neff=3.046, m_nu=(0.0, 0.0, 0.06))


def omega_gamma(H0, Tcmb0):
    """Photon density today: u = a_rad T^4 with a_rad = 4 sigma_SB / c (J/m^3),
    over the critical energy density 3 H0^2 c^2 / (8 pi G)."""
    sigma_sb, c, G, mpc = 5.670374419e-8, 299792458.0, 6.67430e-11, 3.0856775814913673e22
    H = H0 * 1e3 / mpc
    return (4.0 * sigma_sb / c * Tcmb0 ** 4) / (3.0 * H * H * c * c / (8.0 * np.pi * G))


def nu_density_ratio(z, Tcmb0, Neff, m_nu):
    """rho_nu(z) / rho_gamma(z) as astropy's FLRW.nu_relative_density
    computes it: 7/8 (4/11)^(4/3) (Neff / N) [sum over massive species of
    (1 + (k y)^p)^(1/p) + N_massless], y = m c^2 / (k_B T_nu0 (1 + z)),
    T_nu0 = (4/11)^(1/3) Tcmb0, k = 0.3173, p = 1.83 (Komatsu et al. 2011,
    ApJS 192, 18, eq. 26), N = floor(Neff)."""
    k_b, ev = 1.380649e-23, 1.602176634e-19
    n = int(np.floor(Neff))
    t_nu = (4.0 / 11.0) ** (1.0 / 3.0) * Tcmb0
    massive = [m for m in m_nu if m > 0]
    rel = float(n - len(massive))
    for m in massive:
        y = m * ev / (k_b * t_nu) / (1.0 + z)
        rel += (1.0 + (0.3173 * y) ** 1.83) ** 0.54644808743
    return 7.0 / 8.0 * (4.0 / 11.0) ** (4.0 / 3.0) * (Neff / n) * rel


def efunc_flrw(z, H0, Om0, Tcmb0, Neff, m_nu):
    """H(z)/H0 of a flat FLRW cosmology with matter, photons, neutrinos and
    a cosmological constant (Omega_de = 1 - Om0 - Omega_r(0))."""
    og = omega_gamma(H0, Tcmb0)
    ode = 1.0 - Om0 - og * (1.0 + nu_density_ratio(0.0, Tcmb0, Neff, m_nu))
    q = 1.0 + z
    return np.sqrt(Om0 * q ** 3 + og * (1.0 + nu_density_ratio(z, Tcmb0, Neff, m_nu)) * q ** 4 + ode)


def comoving_distance_flrw(z, h, H0=P18["H0"], Om0=P18["Om0"], Tcmb0=P18["Tcmb0"], Neff=P18["Neff"],
                           m_nu=P18["m_nu"]):
    """chi(z) = c/H0 int_0^z dz'/E(z') in Mpc, times ``h`` (the Planck18
    branch of ``catalog_processor.py:609-632``), by scipy.integrate.quad."""
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))
    out = np.array([integrate.quad(lambda zp: 1.0 / efunc_flrw(zp, H0, Om0, Tcmb0, Neff, m_nu), 0.0, zz,
                                   epsabs=0.0, epsrel=1e-13, limit=200)[0] for zz in z])
    return out * (C_KMS / H0) * h


def ra_dec_z_to_xyz(arr, para_cosmo):
    """``catalog_processor.py:589-645``: FlatLambdaCDM(H0 = 100 h, Om0)
    when the config gives 'h' and 'Omega0', else the Planck18 distances
    times the config's h."""
    ra = np.deg2rad(arr["RA"])
    dec = np.deg2rad(arr["DEC"])
    if para_cosmo is not None and "h" in para_cosmo and "Omega0" in para_cosmo:
        chi = comoving_distance(arr["Z"], para_cosmo["Omega0"])
    else:
        chi = comoving_distance_flrw(arr["Z"], para_cosmo["h"])
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
