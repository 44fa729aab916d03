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


def lognormal_box(nparticles, boxsize=1000.0, ngrid=256, seed=0, amp=2.0e4, slope=-1.5, f_growth=0.8,
                  return_velocity=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(ngrid)
    L = float(boxsize)
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
    rho = np.exp(g - 0.5 * sigma2)
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
