"""Seeded synthetic box catalogs: lognormal field + Poisson-like sampling +
linear (Zel'dovich) redshift-space displacement along z.

Recipe (the workloads' catalog definition):
* Gaussian field g on an ``ngrid^3`` grid in a box of side ``boxsize`` whose
  lognormal transform 1 + delta = exp(g - sigma^2/2) has power
  P(k) = amp * (k / 0.1)^slope (Mpc/h)^3: xi_LN from P, xi_G = ln(1 + xi_LN),
  P_G = FFT(xi_G) clipped at 0;
* exactly ``nparticles`` objects are distributed over cells with
  probabilities proportional to 1 + delta (multinomial), positions uniform
  within their cell, float64;
* displacement f * Psi_z with Psi_k = i k delta_G(k) / k^2 (linear field),
  f = 0.8, periodic wrap.
"""
import numpy as np


def _kgrid(ngrid, boxsize):
    k1 = np.fft.fftfreq(ngrid, d=boxsize / ngrid) * 2 * np.pi
    kz = np.fft.rfftfreq(ngrid, d=boxsize / ngrid) * 2 * np.pi
    return k1[:, None, None], k1[None, :, None], kz[None, None, :]


def _lognormal_density(n, L, rng, amp, slope):
    """Gaussian field g and lognormal density 1 + delta = exp(g - sigma^2/2)
    on an n^3 grid of side L whose lognormal power is amp (k/0.1)^slope."""
    V = L ** 3
    dV = V / n ** 3
    kx, ky, kz = _kgrid(n, L)
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    with np.errstate(divide="ignore"):
        P_ln = np.where(k > 0, amp * (k / 0.1) ** slope, 0.0)
    # xi(r) = (1/V) sum_k P(k) e^{ikr}
    xi_ln = np.fft.irfftn(P_ln, s=(n, n, n), axes=(0, 1, 2)) * (n ** 3) / V
    xi_g = np.log1p(np.maximum(xi_ln, -0.99))
    P_g = np.fft.rfftn(xi_g) * dV
    P_g = np.maximum(P_g.real, 0.0)
    P_g[0, 0, 0] = 0.0
    white = rng.standard_normal((n, n, n))
    gk = np.fft.rfftn(white) * np.sqrt(P_g / dV)
    g = np.fft.irfftn(gk, s=(n, n, n), axes=(0, 1, 2))
    sigma2 = g.var()
    return g, np.exp(g - 0.5 * sigma2)


def lognormal_box(nparticles, boxsize=1000.0, ngrid=256, seed=0, amp=2.0e4, slope=-1.5, f_growth=0.8,
                  return_velocity=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(ngrid)
    L = float(boxsize)
    kx, ky, kz = _kgrid(n, L)
    g, rho = _lognormal_density(n, L, rng, amp, slope)
    prob = (rho / rho.sum()).ravel()
    counts = rng.multinomial(int(nparticles), prob)
    cells = np.repeat(np.arange(n ** 3, dtype=np.int64), counts)
    ix, rem = np.divmod(cells, n * n)
    iy, iz = np.divmod(rem, n)
    H = L / n
    pos = np.empty((len(cells), 3), dtype=np.float64)
    pos[:, 0] = (ix + rng.random(len(cells))) * H
    pos[:, 1] = (iy + rng.random(len(cells))) * H
    pos[:, 2] = (iz + rng.random(len(cells))) * H
    # linear displacement along z from the Gaussian (linear) field
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        psi_k = np.where(k2 > 0, 1j * kz / k2, 0.0) * np.fft.rfftn(g)
    psi_z = np.fft.irfftn(psi_k, s=(n, n, n), axes=(0, 1, 2)).ravel()[cells]
    pos[:, 2] = np.mod(pos[:, 2] + f_growth * psi_z, L)
    pos = np.mod(pos, L)
    if return_velocity:
        return pos, psi_z
    return pos


def uniform_box(nparticles, boxsize=1000.0, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.random((int(nparticles), 3)) * float(boxsize)


def write_npy_catalog(path, pos, weight=None):
    """Box catalog file for run_stats: columns x, y, z[, w]."""
    cols = [pos[:, 0], pos[:, 1], pos[:, 2]]
    if weight is not None:
        cols.append(weight)
    np.save(path, np.column_stack(cols).astype(np.float64))
    return ["x", "y", "z"] + (["w"] if weight is not None else [])


# ---------------------------------------------------------------------- survey
def _chi_table(omega_m, zmax, npts=4097):
    """Comoving distance chi(z) (Mpc/h) on a grid, flat LCDM, trapezoid rule."""
    z = np.linspace(0.0, zmax, npts)
    e = 1.0 / np.sqrt(omega_m * (1 + z) ** 3 + 1 - omega_m)
    chi = np.concatenate([[0.0], np.cumsum(0.5 * (e[1:] + e[:-1]) * np.diff(z))]) * 2997.92458
    return z, chi


def _sky_draw(rng, count, ra, dec, zr, nz_fn, omega_m):
    """Points uniform on the sky patch with redshifts distributed as
    n(z) dV/dz; returns RA, DEC (deg), Z and Cartesian positions."""
    zt, chi = _chi_table(omega_m, zr[1] * 1.05)
    m = (zt >= zr[0]) & (zt <= zr[1])
    zg = zt[m]
    pdf = nz_fn(zg) * chi[m] ** 2 * np.gradient(chi, zt)[m]
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(zg))])
    cdf /= cdf[-1]
    z = np.interp(rng.random(count), cdf, zg)
    r = rng.uniform(ra[0], ra[1], count)
    sd = rng.uniform(np.sin(np.deg2rad(dec[0])), np.sin(np.deg2rad(dec[1])), count)
    d = np.rad2deg(np.arcsin(sd))
    c = np.interp(z, zt, chi)
    rr, dd = np.deg2rad(r), np.deg2rad(d)
    pos = np.column_stack([c * np.cos(dd) * np.cos(rr), c * np.cos(dd) * np.sin(rr), c * np.sin(dd)])
    return r, d, z, pos


def lognormal_survey(ndata, nrandoms, seed=0, ra=(150.0, 250.0), dec=(0.0, 60.0), zr=(0.4, 0.7), nz_mean=0.55,
                     nz_sigma=0.08, nz_peak=3.0e-4, omega_m=0.31, ngrid=128, amp=2.0e4, slope=-1.5, fkp_p0=1.0e4):
    """Seeded survey-like catalogs (structured arrays with the FITS columns
    RA, DEC, Z, WEIGHT_FKP, NZ): randoms uniform on the sky patch with
    redshifts following n(z) dV/dz, n(z) Gaussian; data drawn from the same
    distribution thinned by a lognormal density (acceptance (1+delta)/max);
    WEIGHT_FKP = 1 / (1 + NZ * fkp_p0); no completeness columns."""
    rng = np.random.Generator(np.random.PCG64(seed))

    def nz_fn(z):
        return nz_peak * np.exp(-0.5 * ((z - nz_mean) / nz_sigma) ** 2)

    def rows(r, d, z):
        out = np.zeros(len(z), dtype=[("RA", "f8"), ("DEC", "f8"), ("Z", "f8"), ("WEIGHT_FKP", "f8"), ("NZ", "f8")])
        out["RA"], out["DEC"], out["Z"] = r, d, z
        out["NZ"] = nz_fn(z)
        out["WEIGHT_FKP"] = 1.0 / (1.0 + out["NZ"] * fkp_p0)
        return out

    r, d, z, pos = _sky_draw(rng, int(nrandoms), ra, dec, zr, nz_fn, omega_m)
    randoms = rows(r, d, z)
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    L = float(np.max(hi - lo)) * 1.02
    center = 0.5 * (lo + hi)
    _, rho = _lognormal_density(int(ngrid), L, rng, amp, slope)
    cmax = rho.max()
    n = int(ngrid)
    got_r, got_d, got_z = [], [], []
    need = int(ndata)
    while need > 0:
        m = max(1024, int(need * cmax * 1.2))
        r, d, z, pos = _sky_draw(rng, m, ra, dec, zr, nz_fn, omega_m)
        idx = np.clip(((pos - center) / L + 0.5) * n, 0, n - 1).astype(np.int64)
        acc = rng.random(m) < rho[idx[:, 0], idx[:, 1], idx[:, 2]] / cmax
        sel = np.nonzero(acc)[0][:need]
        got_r.append(r[sel])
        got_d.append(d[sel])
        got_z.append(z[sel])
        need -= len(sel)
    data = rows(np.concatenate(got_r), np.concatenate(got_d), np.concatenate(got_z))
    return data, randoms
