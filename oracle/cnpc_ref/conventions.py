"""pmesh conventions restated in numpy (oracle; test infrastructure only).

The reference paints and transforms through ``pmesh.ParticleMesh`` with
``dtype="complex128"`` (``mesh_generator.py:189-219``).  pmesh is not vendored,
so the conventions the reference depends on are restated here:

1. mesh points sit at integer mesh coordinates ``u = x * N / L`` (+0.5 for the
   shifted interlacing paint, ``pm.affine.shift(0.5)`` at ``mesh_generator.py:201``),
   periodic in every axis;
2. windows are the B-splines of order 1..4 (NGP/CIC/TSC/PCS) whose Fourier
   transforms are the ``sinc**p`` compensations at ``math_evaluator.py:206-248``;
3. ``r2c`` is the forward c2c transform ``e^{-ik.x}`` normalised by ``1/N^3``
   (so ``c[0] = mean``), ``c2r`` is the unnormalised inverse (implied by the
   normalisations at ``clustering_estimator.py:696-700`` and ``:342-343``);
4. ``cfield.x`` gives ``k_d = 2*pi/L * n_d`` with ``n_d`` wrapped to
   ``[-N/2, N/2)``; ``kind="circular"`` gives ``w_d = k_d * H`` in ``[-pi, pi)``;
5. ``rfield.x`` gives ``x_d = i_d * H`` wrapped to ``[-L/2, L/2)``.

Pinned by ``tests/test_oracle_conventions.py`` (mass conservation, partition
of unity, single-particle known answers, sinc**p Fourier transforms of the
windows, interlacing alias cancellation, FFT round trips).
"""
import numpy as np

WINDOW_ORDER = {"ngp": 1, "cic": 2, "tsc": 3, "pcs": 4}


def window_weights(u, order):
    """Return (anchor index, list of per-offset weights) for 1-D B-spline of
    the given order evaluated at mesh coordinate ``u`` (convention 1-2)."""
    if order == 1:  # nearest grid point
        i0 = np.floor(u + 0.5)
        return i0.astype(np.int64), [np.ones_like(u)]
    if order == 2:  # CIC: W = 1 - |s|
        i0 = np.floor(u)
        f = u - i0
        return i0.astype(np.int64), [1.0 - f, f]
    if order == 3:  # TSC: nearest point and its two neighbours
        i1 = np.floor(u + 0.5)
        d = u - i1
        return (i1 - 1).astype(np.int64), [
            0.5 * (0.5 - d) ** 2,
            0.75 - d * d,
            0.5 * (0.5 + d) ** 2,
        ]
    if order == 4:  # PCS: cubic B-spline
        i1 = np.floor(u)
        f = u - i1
        g = 1.0 - f
        return (i1 - 1).astype(np.int64), [
            g ** 3 / 6.0,
            (4.0 - 6.0 * f * f + 3.0 * f ** 3) / 6.0,
            (4.0 - 6.0 * g * g + 3.0 * g ** 3) / 6.0,
            f ** 3 / 6.0,
        ]
    raise ValueError(order)


def paint(position, weight, nmesh, boxsize, sampler, shift=0.0, chunk=1 << 20):
    """Mass assignment onto a periodic N^3 mesh (pmesh ``pm.paint``, used at
    ``mesh_generator.py:199,204-205``).  Returns a complex128 mesh whose real
    part holds the painted mass per cell, like pmesh's complex RealField."""
    n = int(nmesh)
    L = float(boxsize)
    order = WINDOW_ORDER[sampler]
    mesh = np.zeros(n ** 3, dtype=np.float64)
    pos = np.asarray(position, dtype=np.float64)
    w = np.asarray(weight, dtype=np.float64)
    for s in range(0, len(w), chunk):
        p = pos[s:s + chunk]
        ww = w[s:s + chunk]
        anchors, weights = [], []
        for d in range(3):
            u = p[:, d] * (n / L) + shift
            a, wt = window_weights(u, order)
            anchors.append(a)
            weights.append(wt)
        for dx in range(order):
            ix = np.mod(anchors[0] + dx, n)
            for dy in range(order):
                iy = np.mod(anchors[1] + dy, n)
                wxy = ww * weights[0][dx] * weights[1][dy]
                for dz in range(order):
                    iz = np.mod(anchors[2] + dz, n)
                    idx = (ix * n + iy) * n + iz
                    mesh += np.bincount(idx, weights=wxy * weights[2][dz], minlength=n ** 3)
    return mesh.reshape(n, n, n).astype(np.complex128)


def r2c(field):
    """pmesh forward transform: c2c, normalised by 1/N^3 (convention 3)."""
    return np.fft.fftn(field) / field.size


def c2r(cfield):
    """pmesh inverse transform: c2c, unnormalised (convention 3)."""
    return np.fft.ifftn(cfield) * cfield.size


def int_axis(n):
    """Wrapped integer frequencies [0..N/2-1, -N/2..-1]."""
    i = np.arange(n)
    return np.where(i >= n // 2, i - n, i)


def k_axes(nmesh, boxsize):
    """``cfield.x``: broadcastable k_d = 2*pi/L * n_d (convention 4)."""
    n = int(nmesh)
    kf = 2 * np.pi / float(boxsize)
    ax = int_axis(n) * kf
    return [ax.reshape(-1, 1, 1), ax.reshape(1, -1, 1), ax.reshape(1, 1, -1)]


def circular_axes(nmesh):
    """``kind="circular"``: w_d = k_d * H = 2*pi*n_d/N (convention 4)."""
    n = int(nmesh)
    ax = int_axis(n) * (2 * np.pi / n)
    return [ax.reshape(-1, 1, 1), ax.reshape(1, -1, 1), ax.reshape(1, 1, -1)]


def x_axes(nmesh, boxsize):
    """``rfield.x``: x_d = i_d*H wrapped to [-L/2, L/2) (convention 5)."""
    n = int(nmesh)
    H = float(boxsize) / n
    ax = int_axis(n) * H
    return [ax.reshape(-1, 1, 1), ax.reshape(1, -1, 1), ax.reshape(1, 1, -1)]


def pm_painter(position, weight, nmesh, boxsize, sampler, interlaced, boxcenter=None):
    """``pm_painter`` (``mesh_generator.py:172-224``): paint, and with
    interlacing combine ``c = c1/2 + c2/2 * exp(+i/2 sum_d k_d H_d)`` of the
    plain and +0.5-cell-shifted paints (``:212-218``) before ``c2r`` (``:219``)."""
    pos = np.asarray(position, dtype=np.float64)
    if boxcenter is not None:
        pos = pos - np.asarray(boxcenter)
    if not interlaced:
        return paint(pos, weight, nmesh, boxsize, sampler)
    c1 = r2c(paint(pos, weight, nmesh, boxsize, sampler))
    c2 = r2c(paint(pos, weight, nmesh, boxsize, sampler, shift=0.5))
    H = float(boxsize) / int(nmesh)
    k = k_axes(nmesh, boxsize)
    kH = k[0] * H + k[1] * H + k[2] * H
    c = c1 * 0.5 + c2 * 0.5 * np.exp(0.5 * 1j * kH)
    return c2r(c)
