"""CPU FP64 oracle for the tabulated-pseudopotential set-up -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module, as the checker of the engine's
``ttdft_spline_parabolic``, ``ttdft_psp_local`` and ``ttdft_psp_projectors``.
The engine never routes through it.

Restates, in numpy:
  * spline_parabolic: alglib spline1dbuildcubic with no end condition (its
    default: parabolically terminated, S''' = 0 on both end segments), as
    atoms/NonLocalPSPData.cc:382-502 builds every radial table.  alglib is a
    third-party library absent here (free edition, version unpinned); this is
    its published definition, solved here in the slope (Hermite) form.
  * real_harmonics: the real combinations of Boost's spherical_harmonic_r/_i
    that NonLocalPSPData::computePhi_lm (:527-604) forms (sqrt2 Im Y_l^|m| for
    m < 0, sqrt2 Re Y_l^m for m > 0, Y_l^0), Condon-Shortley phase included.
  * local_part: NonLocalPSPData::computeLocalPart (:214-275).
  * projector_fields / c_lm: the constructor (:33-212) up to its Tucker step
    and computeC_lm (:609-645), the integral summed exactly on the grid.

Parity pinning (independent of these formulas):
  * spline_parabolic: the global 4(n-1)-unknown system of interpolation, C1,
    C2 and zero third derivative on the end segments, solved densely; exact
    reproduction of quadratics.
  * real_harmonics: scipy.special.sph_harm_y.
  * c_lm: the four 1 / C_lm values the reference printed for the example's
    tables (example/output:48-51).
  * local_part: the -Z/d tail beyond 75.001 and the table's own values at
    knot distances.
"""
from __future__ import annotations

import numpy as np


def spline_parabolic(x, y):
    """Segment coefficients (n-1, 4): c0 + c1 t + c2 t^2 + c3 t^3, t = x - x_s.
    Slopes d_i from the C2 conditions
      h_i d_{i-1} + 2 (h_{i-1} + h_i) d_i + h_{i-1} d_{i+1} = 3 (h_i s_{i-1} + h_{i-1} s_i)
    and the parabolic ends d_0 + d_1 = 2 s_0, d_{n-2} + d_{n-1} = 2 s_{n-2}."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = len(x)
    h = np.diff(x)
    s = np.diff(y) / h
    if n == 2:
        return np.array([[y[0], s[0], 0.0, 0.0]])
    sub = np.zeros(n)
    dia = np.zeros(n)
    sup = np.zeros(n)
    rhs = np.zeros(n)
    dia[0], sup[0], rhs[0] = 1.0, 1.0, 2.0 * s[0]
    for i in range(1, n - 1):
        sub[i] = h[i]
        dia[i] = 2.0 * (h[i - 1] + h[i])
        sup[i] = h[i - 1]
        rhs[i] = 3.0 * (h[i] * s[i - 1] + h[i - 1] * s[i])
    sub[n - 1], dia[n - 1], rhs[n - 1] = 1.0, 1.0, 2.0 * s[n - 2]
    for i in range(1, n):
        w = sub[i] / dia[i - 1]
        dia[i] -= w * sup[i - 1]
        rhs[i] -= w * rhs[i - 1]
    d = np.zeros(n)
    d[n - 1] = rhs[n - 1] / dia[n - 1]
    for i in range(n - 2, -1, -1):
        d[i] = (rhs[i] - sup[i] * d[i + 1]) / dia[i]
    c = np.zeros((n - 1, 4))
    c[:, 0] = y[:-1]
    c[:, 1] = d[:-1]
    c[:, 2] = (3.0 * s - 2.0 * d[:-1] - d[1:]) / h
    c[:, 3] = (d[:-1] + d[1:] - 2.0 * s) / (h * h)
    return c


def spline_eval(x, coef, r):
    """Segment search as alglib spline1dcalc: the end segments extend."""
    x = np.asarray(x, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    seg = np.clip(np.searchsorted(x, r, side="left") - 1, 0, len(x) - 2)
    t = r - x[seg]
    c = coef[seg]
    return c[..., 0] + t * (c[..., 1] + t * (c[..., 2] + t * c[..., 3]))


def _assoc_legendre(l, m, u):
    """P_l^m(u), m >= 0, with the Condon-Shortley phase, by the upward
    recurrence in l."""
    u = np.asarray(u, dtype=np.float64)
    pmm = np.ones_like(u)
    sq = np.sqrt(np.maximum(0.0, 1.0 - u * u))
    for i in range(1, m + 1):
        pmm = -pmm * (2 * i - 1) * sq
    if l == m:
        return pmm
    p1 = u * (2 * m + 1) * pmm
    if l == m + 1:
        return p1
    p0 = pmm
    for ll in range(m + 2, l + 1):
        p0, p1 = p1, ((2 * ll - 1) * u * p1 - (ll + m - 1) * p0) / (ll - m)
    return p1


def real_harmonics(lmax, u, phi):
    """(lmax^2, ...) real harmonics in the order (l, m = -l..l) at cos(polar)
    u and azimuth phi."""
    from math import factorial, pi, sqrt

    u = np.asarray(u, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    out = np.zeros((lmax * lmax,) + u.shape)
    for l in range(lmax):
        for m in range(-l, l + 1):
            am = abs(m)
            nrm = sqrt((2 * l + 1) / (4 * pi) * factorial(l - am) / factorial(l + am))
            p = _assoc_legendre(l, am, u)
            if m == 0:
                v = nrm * p
            elif m > 0:
                v = sqrt(2.0) * nrm * p * np.cos(am * phi)
            else:
                v = sqrt(2.0) * nrm * p * np.sin(am * phi)
            out[l * l + l + m] = v
    return out


def local_part(atoms, r, vloc, gx, gy, gz):
    """V[qx, qy, qz] = sum_a (d > 75.001 ? -Z/d : spline(|q - R_a|)),
    d = |q - R_a| + 1e-12 (atoms/NonLocalPSPData.cc:240-272)."""
    coef = spline_parabolic(r, vloc)
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    V = np.zeros(X.shape)
    for a in np.atleast_2d(atoms):
        dist = np.sqrt((X - a[1]) ** 2 + (Y - a[2]) ** 2 + (Z - a[3]) ** 2)
        d = dist + 1.0e-12
        V += np.where(d > 75.001, -a[0] / d, spline_eval(r, coef, dist))
    return V


def projector_fields(r, vloc, vl, rl, gx, gy, gz):
    """(dV_l phi_lm, phi_lm, dV per field) on the grid, arrays indexed
    [f, qx, qy, qz] with f = l^2 + l + m (atoms/NonLocalPSPData.cc:127-164,
    527-604; dV_l = V_l - V_loc on the table, :441-463)."""
    vl = np.atleast_2d(vl)
    rl = np.atleast_2d(rl)
    lmax = vl.shape[0]
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    dist = np.sqrt(X * X + Y * Y + Z * Z)
    u = Z / (dist + 1.0e-12)
    phi_az = np.arctan2(Y, X)
    Ylm = real_harmonics(lmax, u, phi_az)
    F = lmax * lmax
    field = np.zeros((F,) + X.shape)
    phi = np.zeros((F,) + X.shape)
    dvf = np.zeros((F,) + X.shape)
    for l in range(lmax):
        dv = spline_eval(r, spline_parabolic(r, vl[l] - vloc), dist)
        R = spline_eval(r, spline_parabolic(r, rl[l]), dist)
        for m in range(-l, l + 1):
            f = l * l + l + m
            phi[f] = R * Ylm[f]
            field[f] = dv * phi[f]
            dvf[f] = dv
    return field, phi, dvf


def c_lm(r, vloc, vl, rl, gx, gy, gz, wx, wy, wz):
    """C_lm = sum phi_lm^2 dV_l wx wy wz (computeC_lm, :609-645, exact on the
    grid)."""
    _, phi, dvf = projector_fields(r, vloc, vl, rl, gx, gy, gz)
    W = (np.asarray(wx)[:, None, None] * np.asarray(wy)[None, :, None] *
         np.asarray(wz)[None, None, :])
    return np.array([np.sum(phi[f] * phi[f] * dvf[f] * W) for f in range(phi.shape[0])])
