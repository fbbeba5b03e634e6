"""Seeded divergence-free initial conditions for the plane-Couette runs.

The reference has no seeded generator (its only initial condition is the
analytic field of ``examples/pcf_fluctuations_jax.py:77-97``), so this module
defines one, in spectral space:

* ``numpy.random.default_rng(seed)``; standard-normal real and imaginary parts
  for u0 (clamped Shen index < K) and g (Dirichlet index < K) at streamwise
  wavenumbers |ky| < K and spanwise 0 <= kz < K, plus real mean profiles v00,
  w00 (Dirichlet index < K); draws happen in that order, over arrays of shape
  (K, 2K-1, K) [u0 re, u0 im, g re, g im] then (K,) [v00, w00], with the
  streamwise axis ordered ky = -(K-1) .. K-1;
* the kz = 0 plane is made Hermitian, Nyquist planes are zero, and the mean
  (0, 0) mode of u0 and g is zero;
* u1, u2 follow from incompressibility exactly as the solver reconstructs them
  (``examples/channelflow_kmm.py:855-870``), so div u = 0 to round-off;
* the field is scaled so that the volume-averaged kinetic energy
  E = (1/V) int |u|^2 / 2 dV equals ``energy``.  The reference's ``Epert``
  diagnostic is int |u|^2 dV = 2 V E.

Host-side setup only; the wall-normal matrices come from the library.
"""

from __future__ import annotations

import math

import numpy as np


def fejer_weights(n: int) -> np.ndarray:
    theta = np.pi + (2 * np.arange(n) + 1) * np.pi / (2 * n)
    m = np.arange(1, n // 2 + 1)
    corr = np.sum(2.0 * np.cos(2.0 * theta[:, None] * m[None, :]) / (4.0 * m[None, :] ** 2 - 1.0), axis=1)
    return (2.0 / n) * (1.0 - corr)


def _hermitian_kz0(c: np.ndarray) -> np.ndarray:
    """Make the kz = 0 plane satisfy c(-ky, 0) = conj(c(ky, 0))."""
    ny = c.shape[1]
    plane = c[:, :, 0]
    neg = plane[:, (-np.arange(ny)) % ny]
    c[:, :, 0] = 0.5 * (plane + np.conj(neg))
    return c


def _full_from_half(h: np.ndarray, nz: int) -> np.ndarray:
    """(n, ny, nz/2) half spectrum -> (n, ny, nz) full spectrum (Nyquist = 0)."""
    n, ny, kzs = h.shape
    full = np.zeros((n, ny, nz), complex)
    full[:, :, :kzs] = h
    iy = (-np.arange(ny)) % ny
    for kz in range(1, kzs):
        full[:, :, nz - kz] = np.conj(h[:, iy, kz])
    return full


def perturbation_energy(solver, u) -> float:
    """int |u|^2 dV (the reference's Epert) from spectral coefficients."""
    VB, VD = solver.matrix("VB"), solver.matrix("VD")
    w = fejer_weights(solver.N[0])
    tot = 0.0
    for c, V in ((u[0], VB), (u[1], VD), (u[2], VD)):
        vals = np.tensordot(V, c, axes=(1, 0))  # (n0, ny, nz) nodal values per mode
        tot += solver.Ly * solver.Lz * float(np.sum(w[:, None, None] * np.abs(vals) ** 2))
    return tot


def seeded_state(solver, seed: int, energy: float, K: int):
    """Seeded divergence-free perturbation state for ``solver`` (see module doc)."""
    n0, ny, nz = solver.N
    nb, nd = n0 - 4, n0 - 2
    if K > min(ny, nz) // 2 or K > nb:
        raise ValueError("K exceeds the resolved modes")
    rng = np.random.default_rng(seed)
    shp = (K, 2 * K - 1, K)
    u0r, u0i, gr, gi = (rng.standard_normal(shp) for _ in range(4))
    v00, w00 = rng.standard_normal(K), rng.standard_normal(K)
    kzs = nz // 2
    u0h = np.zeros((nb, ny, kzs), complex)
    gh = np.zeros((nd, ny, kzs), complex)
    kys = np.arange(-(K - 1), K)
    for j, ky in enumerate(kys):
        iy = ky % ny
        u0h[:K, iy, :K] = u0r[:, j, :] + 1j * u0i[:, j, :]
        gh[:K, iy, :K] = gr[:, j, :] + 1j * gi[:, j, :]
    u0h[:, 0, 0] = 0
    gh[:, 0, 0] = 0
    u0h = _hermitian_kz0(u0h)
    gh = _hermitian_kz0(gh)
    u0 = solver.mask_nyquist(_full_from_half(u0h, nz))
    g = solver.mask_nyquist(_full_from_half(gh, nz))
    # incompressible reconstruction (channelflow_kmm.py:855-870)
    R = solver.matrix("R")
    f = np.tensordot(R, u0, axes=(1, 0))
    ky, kz = solver.wavenumbers()
    k2 = ky**2 + kz**2
    den = np.where(k2 == 0, 1.0, k2)
    kyo, kzo = ky / den, kz / den
    u1 = 1j * (kyo * f + kzo * g)
    u2 = 1j * (kzo * f - kyo * g)
    vz = np.zeros(nd)
    wz = np.zeros(nd)
    vz[:K] = v00
    wz[:K] = w00
    u1[:, 0, 0] = vz
    u2[:, 0, 0] = wz
    u1 = solver.mask_nyquist(u1)
    u2 = solver.mask_nyquist(u2)
    e_raw = perturbation_energy(solver, (u0, u1, u2))
    volume = 2.0 * solver.Ly * solver.Lz
    target = 2.0 * volume * energy  # Epert = int |u|^2 = 2 V E
    s = math.sqrt(target / e_raw)
    u0, u1, u2, g = u0 * s, u1 * s, u2 * s, g * s
    return solver._with_zero_history((u0, u1, u2), g)


def magnetic_energy(solver, A) -> float:
    """int |b|^2 dV of b = curl A from spectral coefficients (host setup)."""
    VD, VD1 = solver.matrix("VD"), solver.matrix("VD1")
    if A[1].shape[0] == solver.N[0]:  # Ay, Az in TC
        VY, VY1 = solver.matrix("VC"), solver.matrix("VC1")
    else:
        VY, VY1 = VD, VD1
    ky, kz = solver.wavenumbers()
    ev = lambda V, c: np.tensordot(V, c, axes=(1, 0))  # noqa: E731
    DX, DY, DZ, EY, EZ = ev(VD, A[0]), ev(VY, A[1]), ev(VY, A[2]), ev(VY1, A[1]), ev(VY1, A[2])
    b = (1j * ky * DZ - 1j * kz * DY, 1j * kz * DX - EZ, EY - 1j * ky * DX)
    w = fejer_weights(solver.N[0])
    return sum(solver.Ly * solver.Lz * float(np.sum(w[:, None, None] * np.abs(c) ** 2)) for c in b)


def seeded_mhd_state(solver, seed: int, energy_u: float, energy_b: float, K: int):
    """Seeded MHD state: the velocity of ``seeded_state(solver, seed, energy_u, K)``
    and a vector potential drawn from ``numpy.random.default_rng([seed, 1])``
    (standard-normal real and imaginary parts, components x, y, z in turn, for
    wall-normal index < K at |ky| < K, 0 <= kz < K; kz = 0 plane Hermitian,
    Nyquist planes and the (0, 0) mode zero).  For insulating walls the
    tangential potentials are made to satisfy the vacuum-matching rows: per mode
    and parity, the highest drawn coefficient of P = cy Ay + cz Az and
    Q = cz Ay - cy Az is set so that sum_j beta_j x_j = 0 (beta_j = j^2 for P,
    j^2 + k for Q).  B = curl A is solenoidal to round-off; A is scaled so that
    (1/V) int |b|^2 / 2 dV equals ``energy_b``."""
    flow = seeded_state(solver, seed, energy_u, K)
    n0, ny, nz = solver.N
    rows = [solver._a_rows(i) for i in range(3)]
    if K > min(ny, nz) // 2 or K > min(rows):
        raise ValueError("K exceeds the resolved modes")
    rng = np.random.default_rng([seed, 1])
    shp = (K, 2 * K - 1, K)
    kzs = nz // 2
    kys = np.arange(-(K - 1), K)
    A = []
    for c in range(3):
        re, im = rng.standard_normal(shp), rng.standard_normal(shp)
        h = np.zeros((rows[c], ny, kzs), complex)
        for j, kyv in enumerate(kys):
            h[:K, kyv % ny, :K] = re[:, j, :] + 1j * im[:, j, :]
        h[:, 0, 0] = 0
        A.append(h)
    if rows[1] == n0:  # insulating walls
        ky, kz = solver.wavenumbers()
        kyh, kzh = np.broadcast_to(ky, (ny, nz))[:, :kzs], np.broadcast_to(kz, (ny, nz))[:, :kzs]
        k = np.sqrt(kyh**2 + kzh**2)
        kn = np.where(k > 0, k, 1.0)
        cy, cz = np.where(k > 0, kyh / kn, 1.0), np.where(k > 0, kzh / kn, 0.0)
        P = cy * A[1] + cz * A[2]
        Q = cz * A[1] - cy * A[2]
        j = np.arange(n0)[:, None, None]
        for X, beta in ((P, j * j + 0.0 * k), (Q, j * j + k)):
            for par in (0, 1):
                top = (K - 1 - par) // 2 * 2 + par  # highest drawn row of this parity
                sel = np.arange(par, top, 2)
                X[top] = -np.sum(beta[sel] * X[sel], axis=0) / beta[top]
        A[1] = cy * P + cz * Q
        A[2] = cz * P - cy * Q
    A = [solver.mask_nyquist(_full_from_half(_hermitian_kz0(h), nz)) for h in A]
    e_raw = magnetic_energy(solver, A)
    volume = 2.0 * solver.Ly * solver.Lz
    s = math.sqrt(2.0 * volume * energy_b / e_raw)
    return solver._new_state(flow, [a * s for a in A])
