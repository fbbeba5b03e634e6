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


# ---------------------------------------------------------------------------
# the 3D class (TaylorCouetteMRIDNSJax)
# ---------------------------------------------------------------------------


def test_golden_3d_mri_sbdf3_run():
    """production/goldens/taylor_couette_mhd_3d_v1 (written by the reference's
    3D class itself: Nr = 12, Ntheta = 4, Nz = 4, dealias 3/2, SBDF3,
    dt = 2e-3, three steps from the amp = 1e-7 m = 1 MRI eigenmode).  The
    divergences are maxima over the 4 x 4 (theta, z) grid of a travelling
    wave whose phase is the eigensolver's arbitrary eigenvector phase (no
    pivot phase for m != 0), so they agree in size only."""
    from oracle.tc import TC3DMRIOracle

    o = TC3DMRIOracle(CircularCouette(1.0, 2.0, 1.0, KEPLER), B0=0.1, nu=1e-3, eta_mag=1e-3, Nr=12, Ntheta=4, Nz=4,
                      Lz=GOLDEN_LZ, dt=2e-3, dealias=1.5, time_integrator="SBDF3")
    s0, eig = o.seed_linear_eigenmode(m=1, kz_mode=1, amp=1e-7)
    assert abs(eig.real - 0.2912937448361517) <= 1e-10
    ek0, em0 = o.energy_parts(s0)
    assert abs(ek0 / 7.85609816898076e-15 - 1.0) <= 1e-9 and abs(em0 / 1.1298011285580755e-14 - 1.0) <= 1e-9
    s1 = o.solve(s0, 3)
    ek1, em1 = o.energy_parts(s1)
    assert abs(ek1 / 7.883607407650011e-15 - 1.0) <= 1e-9 and abs(em1 / 1.1337572869748595e-14 - 1.0) <= 1e-9
    growth = 0.5 * math.log((ek1 + em1) / (ek0 + em0)) / 6e-3
    assert abs(growth - 0.29129374483210074) <= 1e-7
    d = o.diagnostics(s1)
    assert abs(d["divu"] / 1.486388530128091e-08 - 1.0) < 0.05
    assert abs(d["divb"] / 1.0541502761126154e-08 - 1.0) < 0.05


def _direct_rows_3d(o, x, drop_azimuthal=False):
    """Both row groups of the 3D class by explicit sums (no FFT or DCT)."""
    Mth, Mz, Mr, Nth, Nz, Nr = o.Mth, o.Mz, o.Mr, o.Ntheta, o.Nz, o.Nr
    df = o.rf.df

    def nodes(M):
        return np.cos(np.pi + (2 * np.arange(M) + 1) * np.pi / (2 * M))

    xp, xs = nodes(Mr), nodes(Nr)
    rp = o.base.R1 + (xp + 1.0) / df
    rs = o.base.R1 + (xs + 1.0) / df
    mint = np.fft.fftfreq(Nth, 1.0 / Nth)
    kz = 2 * np.pi * np.fft.fftfreq(Nz, 1.0 / Nz) / o.Lz
    m_odd, kz_odd = mint.copy(), kz.copy()
    m_odd[Nth // 2] = 0.0
    kz_odd[Nz // 2] = 0.0
    Etp = np.exp(1j * np.outer(np.arange(Mth) * 2 * np.pi / Mth, mint))
    Ezp = np.exp(1j * np.outer(np.arange(Mz) * o.Lz / Mz, kz))
    Ets = np.exp(1j * np.outer(np.arange(Nth) * 2 * np.pi / Nth, mint))
    Ezs = np.exp(1j * np.outer(np.arange(Nz) * o.Lz / Nz, kz))
    fields = []
    for f, c in enumerate(x):
        S = o.bases[f]
        val = np.stack([npcheb.chebval(xp, s) for s in S], axis=1)
        der = df * np.stack([npcheb.chebval(xp, npcheb.chebder(s)) for s in S], axis=1)
        syn = lambda cc, R: np.einsum("am,bk,mki,ri->abr", Etp, Ezp, cc, R)  # noqa: E731
        fields.append((syn(c, val), syn(c, der), syn(1j * m_odd[:, None, None] * c, val),
                       syn(1j * kz_odd[None, :, None] * c, val)))
    (ur, urr, urt, urz), (ut, utr, utt, utz), (uz, uzr, uzt, uzz) = fields[:3]
    (br, brr, brt, brz), (bt, btr, btt, btz), (bz, bzr, bzt, bzz) = fields[3:]
    inv = 1.0 / rp[None, None, :]
    au = 0.0 if drop_azimuthal else ut * inv
    ab = 0.0 if drop_azimuthal else bt * inv
    mom = (ur * urr + au * urt + uz * urz - ut * ut * inv - (br * brr + ab * brt + bz * brz - bt * bt * inv),
           ur * utr + au * utt + uz * utz + ur * ut * inv - (br * btr + ab * btt + bz * btz + br * bt * inv),
           ur * uzr + au * uzt + uz * uzz - (br * bzr + ab * bzt + bz * bzz))
    emf = (ut * bz - uz * bt, uz * br - ur * bz, ur * bt - ut * br)
    Tp = npcheb.chebvander(xp, Nr - 1)
    Ts = npcheb.chebvander(xs, Nr - 1)

    def coeffs(f):
        c = np.einsum("am,bk,abr->mkr", np.conj(Etp), np.conj(Ezp), f) / (Mth * Mz)
        c = (2.0 / Mr) * c @ Tp
        c[..., 0] *= 0.5
        return c

    def std(c, kind=None):
        cc = c.copy()
        if kind == "r":
            cc = df * np.apply_along_axis(lambda row: np.append(npcheb.chebder(row), 0.0), -1, cc)
        elif kind == "t":
            cc = 1j * m_odd[:, None, None] * cc
        elif kind == "z":
            cc = 1j * kz_odd[None, :, None] * cc
        return np.einsum("am,bk,mkn,qn->abq", Ets, Ezs, cc, Ts)

    def proj(vals, S):
        psi = np.stack([npcheb.chebval(xs, s) for s in S], axis=1)
        fk = 2 * np.pi * o.Lz * np.einsum("am,bk,abr->mkr", np.conj(Ets), np.conj(Ezs), vals) / (Nth * Nz)
        rows = (np.pi / Nr) / df * fk @ psi
        rows[Nth // 2] = 0.0
        rows[:, Nz // 2] = 0.0
        return rows

    SD = o.rf.SD
    out = [proj(std(coeffs(f)), SD) for f in mom]
    er, et, ez = (coeffs(f) for f in emf)
    inv_s = 1.0 / rs[None, None, :]
    nb_r = -inv_s * std(ez, "t") + std(et, "z")
    nb_t = -std(er, "z") + std(ez, "r")
    nb_z = -std(et, "r") - std(et) * inv_s + inv_s * std(er, "t")
    out += [proj(nb_r, SD), proj(nb_t, o.SBT), proj(nb_z, o.SBZ)]
    return out


def test_3d_nonlinear_rows_match_direct_series():
    from oracle.tc import TC3DMRIOracle

    o = TC3DMRIOracle(CircularCouette(1.0, 2.0, 1.0, KEPLER), B0=0.1, nu=1e-3, eta_mag=1e-3, Nr=10, Ntheta=6, Nz=4,
                      dt=1e-3, dealias=1.5)
    rng = np.random.default_rng(17)
    x = []
    for _ in range(6):
        c = np.zeros((o.Ntheta, o.Nz, o.Nd), dtype=complex)
        c[0, 0, :4] = rng.standard_normal(4)
        for m, k in ((1, 0), (0, 1), (1, 1), (-2, 1)):
            blk = rng.standard_normal(4) + 1j * rng.standard_normal(4)
            c[m % o.Ntheta, k % o.Nz, :4] = blk
            c[-m % o.Ntheta, -k % o.Nz, :4] = np.conj(blk)
        x.append(c)
    got = o.nonlinear(tuple(x))
    ref = _direct_rows_3d(o, x)
    scale = max(np.max(np.abs(r)) for r in ref)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g - r)) <= 1e-12 * scale
    bad = _direct_rows_3d(o, x, drop_azimuthal=True)
    assert max(np.max(np.abs(g - b)) for g, b in zip(got, bad)) > 1e-3 * scale


def test_3d_axisymmetric_state_steps_as_the_axisymmetric_oracle():
    from oracle.tc import TC3DMRIOracle

    base = CircularCouette(1.0, 2.0, 1.0, KEPLER)
    kw = dict(B0=0.1, nu=2e-3, eta_mag=3e-3, Nr=12, Nz=8, dt=2e-3, dealias=1.5, time_integrator="SBDF3")
    o2, o3 = TCMRIOracle(base, **kw), TC3DMRIOracle(base, Ntheta=4, **kw)
    s2, _ = o2.seed_linear_eigenmode(kz_mode=1, amp=3e-2)
    s3 = o3.zero_state()
    s3 = type(s3)(u=tuple(np.stack([c] + [np.zeros_like(c)] * 3) for c in s2.u), p=s3.p,
                  nonlinear_old=s3.nonlinear_old)
    a, b = o2.solve(s2, 4), o3.solve(s3, 4)
    for c2, c3 in zip(a.u, b.u):
        assert np.max(np.abs(c3[0] - c2)) <= 1e-12 * np.max(np.abs(c2))
        assert np.max(np.abs(c3[1:])) <= 1e-14 * np.max(np.abs(c2))
