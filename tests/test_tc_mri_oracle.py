"""Pins of the conducting-wall Taylor-Couette MHD oracle (``oracle.tc.TCMRIOracle``), CPU only.

* The published golden run ``production/goldens/taylor_couette_mhd_dns_v1``
  (shenfun reference ``AxisymmetricMRIDNS``: Re = Rm = 1000, B0 = 0.1,
  R1 = 1, R2 = 2, Omega2 = 0.5^1.5 (Keplerian), Nr = 40, Nz = 8, Lz = pi / 3,
  CNAB2, dt = 2e-3, 250 steps from the amp = 1e-7 leading MRI eigenmode): the
  linear growth rate, the DNS growth rate and both energies at t = 0 and
  t = 0.5.  As for the hydrodynamic golden, the reference port writes half
  the eigenvector at +k and at -k, so the published energies are 4x these.
  These pin the Robin basis of d(r b_theta)/dr = 0, the Neumann basis of b_z,
  every induction and Lorentz coupling, the total-pressure saddle point and
  the conducting MRI pencil that seeds the run.
* The two wall bases against their boundary conditions and the closed form
  of the Neumann stencil.
* The nonlinear rows (advection minus Lorentz force; minus the curl of the
  EMF) against direct series evaluations with explicit sums, with a planted
  mistake in the EMF curl that fails.
* The zero state stays zero under every integrator.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from oracle.tc import CircularCouette, TCMRIOracle, stencil_neumann, stencil_robin, stencil_two_point

KEPLER = 0.5**1.5
GOLDEN_LZ = 1.0471975511965976


def test_golden_mri_dns_growth_and_energies():
    o = TCMRIOracle(CircularCouette(1.0, 2.0, 1.0, KEPLER), B0=0.1, nu=1e-3, eta_mag=1e-3, Nr=40, Nz=8, Lz=GOLDEN_LZ,
                    dt=2e-3, dealias=1.0, time_integrator="CNAB2")
    s0, eig = o.seed_linear_eigenmode(kz_mode=1, amp=1e-7)
    assert abs(eig.real - 0.3403972051440628) <= 1e-10
    ek0, em0 = o.energy_parts(s0)
    assert abs(4.0 * ek0 / 5.222194430323697e-15 - 1.0) <= 1e-9
    assert abs(4.0 * em0 / 7.396547400914273e-15 - 1.0) <= 1e-9
    s1 = o.solve(s0, 250)
    ek1, em1 = o.energy_parts(s1)
    assert abs(4.0 * ek1 / 7.339824415880832e-15 - 1.0) <= 1e-9
    assert abs(4.0 * em1 / 1.039589006705832e-14 - 1.0) <= 1e-9
    growth = 0.5 * math.log((ek1 + em1) / (ek0 + em0)) / 0.5
    assert abs(growth - 0.3403972182913979) <= 1e-9
    d = o.diagnostics(s1)
    assert d["divu"] < 1e-9 and d["divb"] < 1e-9 and d["wall_u"] == 0.0


def test_wall_bases():
    N, R1, R2 = 16, 1.0, 2.5
    df = 2.0 / (R2 - R1)
    S = stencil_robin(N, R1 * df, R2 * df)
    ends = np.array([-1.0, 1.0])
    for row in S:
        v = npcheb.chebval(ends, row)
        dv = df * npcheb.chebval(ends, npcheb.chebder(row))
        # d(r b)/dr = b + r db/dr = 0 at r = R1, R2
        assert np.max(np.abs(v + np.array([R1, R2]) * dv)) <= 1e-12 * np.max(np.abs(row))
    Z = stencil_neumann(N)
    # the closed form against the general two-condition solve
    G = stencil_two_point(N, lambda j: (-1.0) ** (j + 1) * j * j, lambda j: float(j * j))
    assert np.max(np.abs(Z - G)) <= 1e-15
    for row in Z:
        assert np.max(np.abs(npcheb.chebval(ends, npcheb.chebder(row)))) <= 1e-12 * max(1.0, N**2)


def _direct_rows(o: TCMRIOracle, x, emf_sign=1.0):
    """Both row groups by explicit sums (no FFT or DCT)."""
    Mz, Mr, Nz, Nr = o.Mz, o.Mr, o.Nz, o.Nr
    df = o.rf.df

    def nodes(M):
        return np.cos(np.pi + (2 * np.arange(M) + 1) * np.pi / (2 * M))

    xp, xs = nodes(Mr), nodes(Nr)
    rp = o.base.R1 + (xp + 1.0) / df
    rs = o.base.R1 + (xs + 1.0) / df
    kint = np.fft.fftfreq(Nz, 1.0 / Nz)
    kz = 2 * np.pi * kint / o.Lz
    kz_odd = kz.copy()
    kz_odd[Nz // 2] = 0.0
    zp = np.arange(Mz) * o.Lz / Mz
    zs = np.arange(Nz) * o.Lz / Nz
    Ep = np.exp(1j * np.outer(zp, kz))
    Es = np.exp(1j * np.outer(zs, kz))
    fields = []
    for f, c in enumerate(x):
        S = o.bases[f]
        val = np.stack([npcheb.chebval(xp, s) for s in S], axis=1)
        der = df * np.stack([npcheb.chebval(xp, npcheb.chebder(s)) for s in S], axis=1)
        fields.append((Ep @ c @ val.T, Ep @ c @ der.T, Ep @ (1j * kz_odd[:, None] * c) @ val.T))
    (ur, urr, urz), (ut, utr, utz), (uz, uzr, uzz) = fields[:3]
    (br, brr, brz), (bt, btr, btz), (bz, bzr, bzz) = fields[3:]
    inv = 1.0 / rp[None, :]
    mom = (ur * urr + uz * urz - ut * ut * inv - (br * brr + bz * brz - bt * bt * inv),
           ur * utr + uz * utz + ur * ut * inv - (br * btr + bz * btz + br * bt * inv),
           ur * uzr + uz * uzz - (br * bzr + bz * bzz))
    emf = (ut * bz - uz * bt, uz * br - ur * bz, ur * bt - ut * br)
    Tp = npcheb.chebvander(xp, Nr - 1)  # (Mr, Nr)
    Ts = npcheb.chebvander(xs, Nr - 1)

    def coeffs(f):
        """Fourier then Chebyshev coefficients of padded values, truncated."""
        fk = (np.conj(Ep).T @ f) / Mz  # (Nz, Mr)
        c = (2.0 / Mr) * fk @ Tp
        c[:, 0] *= 0.5
        return c

    def proj(vals_s, S):
        """Lz (pi / Nr) / df sum_q psi_i(x_q) f(z, x_q), then the axial analysis."""
        psi = np.stack([npcheb.chebval(xs, s) for s in S], axis=1)
        fk = o.Lz * (np.conj(Es).T @ vals_s) / Nz
        rows = (np.pi / Nr) / df * fk @ psi
        rows[Nz // 2] = 0.0
        return rows

    def std(c, deriv=0, dz=False):
        cc = c.copy()
        if deriv:
            cc = df * np.stack([npcheb.chebder(row, 1) for row in cc])
            cc = np.concatenate([cc, np.zeros((Nz, 1))], axis=1)
        if dz:
            cc = 1j * kz_odd[:, None] * cc
        return Es @ cc @ Ts.T

    SD = o.rf.SD
    out = [proj(std(coeffs(f)), SD) for f in mom]
    er, et, ez = (coeffs(f) for f in emf)
    eps_t_std = std(et)
    nb_r = std(et, dz=True)
    nb_t = -std(er, dz=True) + std(ez, deriv=1)
    nb_z = -std(et, deriv=1) - emf_sign * eps_t_std / rs[None, :]
    out += [proj(nb_r, SD), proj(nb_t, o.SBT), proj(nb_z, o.SBZ)]
    return out


@pytest.mark.parametrize("dealias", [1.5, 1.0])
def test_nonlinear_rows_match_direct_series(dealias):
    o = TCMRIOracle(CircularCouette(1.0, 2.0, 1.0, KEPLER), B0=0.1, nu=1e-3, eta_mag=1e-3, Nr=12, Nz=8, dt=1e-3,
                    dealias=dealias)
    rng = np.random.default_rng(11)
    x = []
    for _ in range(6):
        c = np.zeros((o.Nz, o.Nd), dtype=complex)
        for k in range(1, 3):
            blk = rng.standard_normal(5) + 1j * rng.standard_normal(5)
            c[k, :5] = blk
            c[-k, :5] = np.conj(blk)
        c[0, :5] = rng.standard_normal(5)
        x.append(c)
    got = o.nonlinear(tuple(x))
    ref = _direct_rows(o, x)
    scale = max(np.max(np.abs(r)) for r in ref)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g - r)) <= 1e-12 * scale
    bad = _direct_rows(o, x, emf_sign=-1.0)
    assert np.max(np.abs(got[5] - bad[5])) > 1e-3 * scale


def test_zero_state_stays_zero():
    """``tests/couette/test_taylor_couette_dns_jax.py:287-311``."""
    for ti in ("CNAB2", "SBDF3", "IMEXRK3"):
        o = TCMRIOracle(CircularCouette(1.0, 2.0, 1.0, KEPLER), B0=0.1, nu=2e-3, eta_mag=2e-3, Nr=10, Nz=8, dt=1e-3,
                        time_integrator=ti)
        z = o.solve(o.zero_state(), 3)
        assert all(np.max(np.abs(c)) == 0.0 for c in z.u) and np.max(np.abs(z.p)) == 0.0
