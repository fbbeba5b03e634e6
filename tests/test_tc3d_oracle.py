"""Pins of the 3D Taylor-Couette oracle (``oracle.tc.TC3DOracle``), CPU only.

* The published golden run ``production/goldens/taylor_couette_hydro_3d_v1``
  (Re = 100, R1 = 1, R2 = 2, Omega2 = 0, Nr = 12, Ntheta = 4, Nz = 4,
  Lz = 2 pi / 3.13, dealias 3/2, SBDF3, dt = 1e-3, three steps from the
  amp = 1e-6 leading m = 1 eigenmode): the linear growth rate, the seeded and
  final kinetic energies, the DNS growth rate and the divergence.  These pin
  the per-(m, kz) blocks (base-flow advection -i m Omega, the 2/r^2 azimuthal
  couplings, the 3D Laplacian), the pressure pair, the pinned solve, the
  ARS(4,4,3) start-up and the 2 pi Lz Parseval energy.
* The 3D advection rows against a direct evaluation of the series at every
  padded node, with a planted mistake (the azimuthal advection dropped).
* The axisymmetric reduction: a state with only m = 0 content steps exactly
  as the axisymmetric oracle steps it.
* The reference's own 3D cases (``tests/couette/test_taylor_couette_dns_jax.py:91-175,239-285``):
  the zero state stays zero, the m = 0 seed equals the axisymmetric one, the
  m = 1 eigenmode grows at its linear rate with the pinned pressure at zero.
"""

from __future__ import annotations

import math

import numpy as np
from numpy.polynomial import chebyshev as npcheb

from oracle.tc import CircularCouette, TC3DOracle, TCOracle

GOLDEN_LZ = 2.0074074463832545


def test_golden_3d_linear_and_dns_growth():
    o = TC3DOracle(CircularCouette(1.0, 2.0, 1.0, 0.0), nu=0.01, Nr=12, Ntheta=4, Nz=4, Lz=GOLDEN_LZ, dt=1e-3,
                   dealias=1.5, time_integrator="SBDF3")
    s0, eig = o.seed_linear_eigenmode(m=1, kz_mode=1, amp=1e-6)
    assert abs(eig.real - 0.08387873243612082) <= 1e-10
    e0 = o.energy(s0)
    assert abs(e0 / 7.951614264231089e-12 - 1.0) <= 1e-9
    s1 = o.solve(s0, 3)
    e1 = o.energy(s1)
    assert abs(e1 / 7.955617099356648e-12 - 1.0) <= 1e-9
    growth = 0.5 * math.log(e1 / e0) / 3e-3
    assert abs(growth - 0.08387873243632032) <= 1e-7
    d = o.diagnostics(s1)
    assert abs(d["div_linf"] / 7.255456995739012e-10 - 1.0) <= 1e-3


def _direct_rows_3d(o: TC3DOracle, u, drop_azimuthal=False):
    """inner(v_i, N_i) by direct series evaluation on the (Mth, Mz, Mr) grid."""
    Mth, Mz, Mr, Nth, Nz = o.Mth, o.Mz, o.Mr, o.Ntheta, o.Nz
    xp = np.cos(np.pi + (2 * np.arange(Mr) + 1) * np.pi / (2 * Mr))
    rp = o.base.R1 + (xp + 1.0) * (o.base.R2 - o.base.R1) / 2.0
    df = 2.0 / (o.base.R2 - o.base.R1)
    th = np.arange(Mth) * 2 * np.pi / Mth
    zp = np.arange(Mz) * o.Lz / Mz
    mint = np.fft.fftfreq(Nth, 1.0 / Nth)
    kint = np.fft.fftfreq(Nz, 1.0 / Nz)
    kz = 2 * np.pi * kint / o.Lz
    m_odd, kz_odd = mint.copy(), kz.copy()
    m_odd[Nth // 2] = 0.0
    kz_odd[Nz // 2] = 0.0
    Et = np.exp(1j * np.outer(th, mint))  # (Mth, Nth)
    Ez = np.exp(1j * np.outer(zp, kz))  # (Mz, Nz)
    phi = np.zeros((o.Nd, o.Nr))
    phi[np.arange(o.Nd), np.arange(o.Nd)] = 1.0
    phi[np.arange(o.Nd), np.arange(o.Nd) + 2] = -1.0
    val = np.stack([npcheb.chebval(xp, c) for c in phi], axis=1)  # (Mr, Nd)
    der = df * np.stack([npcheb.chebval(xp, npcheb.chebder(c)) for c in phi], axis=1)

    def syn(c, R):
        return np.einsum("am,bk,mki,ri->abr", Et, Ez, c, R)

    fields = []
    for c in u:
        fields.append((syn(c, val), syn(c, der), syn(1j * m_odd[:, None, None] * c, val),
                       syn(1j * kz_odd[None, :, None] * c, val)))
    (ur, urr, urt, urz), (ut, utr, utt, utz), (uz, uzr, uzt, uzz) = fields
    inv = 1.0 / rp[None, None, :]
    at = 0.0 if drop_azimuthal else ut * inv
    prods = (ur * urr + at * urt + uz * urz - ut * ut * inv, ur * utr + at * utt + uz * utz + ur * ut * inv,
             ur * uzr + at * uzt + uz * uzz)
    out = []
    for f in prods:
        fk = (2 * np.pi / Mth) * (o.Lz / Mz) * np.einsum("am,bk,abr->mkr", np.conj(Et), np.conj(Ez), f)
        rows = (np.pi / Mr) / df * fk @ val
        rows[Nth // 2] = 0.0
        rows[:, Nz // 2] = 0.0
        out.append(rows)
    return out


def _random_state(o, rng, modes=((0, 1), (1, 0), (1, 1), (-1, 2))):
    u = []
    for _ in range(3):
        c = np.zeros((o.Ntheta, o.Nz, o.Nd), dtype=complex)
        c[0, 0, :6] = rng.standard_normal(6)
        for m, k in modes:
            blk = rng.standard_normal(6) + 1j * rng.standard_normal(6)
            c[m % o.Ntheta, k % o.Nz, :6] = blk
            c[-m % o.Ntheta, -k % o.Nz, :6] = np.conj(blk)
        u.append(c)
    return tuple(u)


def test_3d_nonlinear_rows_match_direct_series():
    o = TC3DOracle(CircularCouette(1.0, 2.0, 1.0, -0.3), nu=0.01, Nr=12, Ntheta=6, Nz=8, dt=1e-3, dealias=1.5)
    u = _random_state(o, np.random.default_rng(5))
    got = o.nonlinear(u)
    ref = _direct_rows_3d(o, u)
    scale = max(np.max(np.abs(r)) for r in ref)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g - r)) <= 1e-12 * scale
    bad = _direct_rows_3d(o, u, drop_azimuthal=True)
    assert max(np.max(np.abs(g - b)) for g, b in zip(got, bad)) > 1e-3 * scale


def test_axisymmetric_state_steps_as_the_axisymmetric_oracle():
    base = CircularCouette(1.0, 2.0, 1.0, 0.0)
    kw = dict(nu=0.01, Nr=14, Nz=8, dt=2e-3, dealias=1.5, time_integrator="SBDF3")
    o3 = TC3DOracle(base, Ntheta=4, **kw)
    o2 = TCOracle(base, **kw)
    s2 = o2.initial_state(amp=0.2, kz_mode=1)
    s3 = o3.zero_state()
    s3 = type(s3)(u=tuple(np.stack([c] + [np.zeros_like(c)] * 3) for c in s2.u), p=s3.p,
                  nonlinear_old=s3.nonlinear_old)
    a = o2.solve(s2, 4)
    b = o3.solve(s3, 4)
    for c2, c3 in zip(a.u, b.u):
        assert np.max(np.abs(c3[0] - c2)) <= 1e-12 * np.max(np.abs(c2))
        assert np.max(np.abs(c3[1:])) <= 1e-14 * np.max(np.abs(c2))


def test_reference_3d_cases():
    base = CircularCouette()
    o = TC3DOracle(base, nu=0.002, Nr=8, Ntheta=4, Nz=6, dt=1e-3, dealias=1.0, time_integrator="CNAB2")
    z = o.step(o.zero_state())
    assert all(np.max(np.abs(c)) == 0.0 for c in z.u) and np.max(np.abs(z.p)) == 0.0
    o2 = TCOracle(base, nu=0.002, Nr=8, Nz=6, dt=1e-3, dealias=1.0)
    sa, ea = o2.seed_linear_eigenmode(kz_mode=1, amp=1e-5)
    s3, e3 = o.seed_linear_eigenmode(m=0, kz_mode=1, amp=1e-5)
    assert abs(e3 - ea) <= 1e-12 * abs(ea)
    for va, v3 in zip(o2.velocity_physical(sa), o.velocity_physical(s3)):
        assert np.allclose(v3[0], va, rtol=1e-10, atol=1e-12)
    s, eig = o.seed_linear_eigenmode(m=1, kz_mode=1, amp=1e-8)
    e0 = o.energy(s)
    out = o.solve(s, 100)
    rate = 0.5 * math.log(o.energy(out) / e0) / 0.1
    assert abs(rate - eig.real) <= 1e-6 * max(1.0, abs(eig.real))
    assert o.continuity_l2(out) < 1e-15
    assert abs(out.p[0, 0, 0]) < 1e-18
