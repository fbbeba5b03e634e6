"""Pins of the insulating-wall MHD oracle (oracle/mhd_insulating.py).

The reference publishes no trajectory for this class, so the oracle is pinned
by what the mathematics fixes and by the reference's own invariant thresholds:

* the vacuum-matching residual stays below 1e-10 and div B at rounding
  (``tests/couette/test_pcf_sbdf3_jax.py:164-167`` asserts the same bounds);
* with zero field, S = -1 and Omega = 0 the flow equals the hydrodynamic
  oracle bit for bit (same base flow, no Lorentz force, no Coriolis force);
* with no flow and no shear the EMF vanishes, and a P or Q eigenmode of the
  bordered resistive operator decays by the scalar IMEXRK3/SBDF3 recurrence of
  its eigenvalue, which pins the bordered systems, their wall rows, the P/Q
  rotation, the stage and history coefficients;
* B = curl A and the TC forward transform against nodal values.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from oracle.kmm import IMEXRK3_A, IMEXRK3_B, PCFOracle
from oracle.mhd import MHDOracleState
from oracle.mhd_insulating import InsulatingMHDOracle


def make(**kw):
    args = dict(N=(17, 16, 16), Re=400.0, Rm=300.0, omega=0.4, shear_rate=1.5, background_b=(0.0, 0.02, 0.1),
                dt=2e-3, perturbation_amplitude=0.05, magnetic_amplitude=0.3)
    args.update(kw)
    return InsulatingMHDOracle(**args)


def _tc_field(o, seed, scale):
    """Real-field TC coefficients (Hermitian, Nyquist-masked) that need not meet the walls' rows."""
    rng = np.random.default_rng(seed)
    x = o.xb
    y = np.linspace(0, o.L1, o.N[1], endpoint=False)[None, :, None]
    z = np.linspace(0, o.L2, o.N[2], endpoint=False)[None, None, :]
    f = np.zeros((o.N0, o.N[1], o.N[2]))
    for _ in range(5):
        a, b = rng.integers(0, 3, 2)
        f = f + rng.normal() * (1 + x + x**3) * np.cos(2 * np.pi * a * y / o.L1 + rng.uniform(0, 6)) * np.cos(
            2 * np.pi * b * z / o.L2 + rng.uniform(0, 6))
    return scale * o.mask(o.tp_forward(f.astype(complex), np.eye(o.N0), np.diag(o.h)))


def test_vacuum_matching_and_solenoidal_field():
    o = make()
    s = o.initial_state()
    s = o.reset_mhd_history(MHDOracleState(flow=s.flow, A=(s.A[0], _tc_field(o, 1, 0.01), _tc_field(o, 2, 0.01))))
    assert o.diagnostics(s)["insulating_bc_residual"] > 1e-3  # the seed does not meet the wall rows
    for k in range(5):
        s = o.step(s)
        d = o.diagnostics(s)
        assert d["insulating_bc_residual"] < 1e-10, (k, d)
        assert d["divB_L2"] < 1e-13 and d["divL2"] < 1e-13
    assert np.max(np.abs(s.A[1])) > 1e-3 and np.max(np.abs(s.A[2])) > 1e-3  # the tangential potentials are live


def test_zero_field_reproduces_hydro_oracle():
    o = make(omega=0.0, shear_rate=-1.0, background_b=(0.0, 0.0, 0.0), magnetic_amplitude=0.0)
    h = PCFOracle(N=o.N, Re=o.Re, U_wall=1.0, dt=o.dt)
    so = o.initial_state()
    sh = h.reset_history(type(so.flow)(u=so.flow.u, g=so.flow.g))
    so, sh = o.solve(so, 5), h.solve(sh, 5)
    for a, b in zip(so.flow.leaves(), sh.leaves()):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def _bordered_eigenmode(o, kind, k):
    n = o.N0
    K = o.eta * (o.D2C - k * k * o.M1)
    rows = np.stack([o.d1[0], o.d1[1]]) if kind == "P" else np.stack([o.d1[0] - k * o.d0[0], o.d1[1] + k * o.d0[1]])
    Z = sla.null_space(rows)
    lam, w = sla.eig((K @ Z)[: n - 2], (o.M1 @ Z)[: n - 2])
    i = int(np.argmax(lam.real))
    x = Z @ np.real(w[:, i])
    return float(lam[i].real), x / np.max(np.abs(x))


def test_bordered_eigenmodes_decay_by_scalar_recurrence():
    dt = 0.01
    o = make(N=(17, 8, 8), Rm=20.0, omega=0.0, shear_rate=0.0, background_b=(0.0, 0.0, 0.0), dt=dt,
             perturbation_amplitude=0.0, magnetic_amplitude=0.0)
    iy, iz = 1, 2
    k = float(np.hypot(o.K1[0, iy, 0], o.K2[0, 0, iz]))
    cy, cz = o.cy[iy, iz], o.cz[iy, iz]
    eps = 1e-7
    for kind in ("P", "Q"):
        lam, x = _bordered_eigenmode(o, kind, k)
        Ay = np.zeros((o.N0, o.N[1], o.N[2]), complex)
        Az = np.zeros_like(Ay)
        # P = cy Ay + cz Az, Q = cz Ay - cy Az; the conjugate mode carries conj(A)
        fy, fz = (cy, cz) if kind == "P" else (cz, -cy)
        Ay[:, iy, iz], Az[:, iy, iz] = eps * fy * x, eps * fz * x
        Ay[:, -iy, -iz], Az[:, -iy, -iz] = np.conj(Ay[:, iy, iz]), np.conj(Az[:, iy, iz])
        s0 = o.initial_state()
        s = o.reset_mhd_history(MHDOracleState(flow=s0.flow, A=(s0.A[0], Ay, Az)))
        ys, y = [1.0], 1.0
        for _ in range(2):
            for rk in range(3):
                g = (IMEXRK3_A[rk] + IMEXRK3_B[rk]) * dt / 2.0
                y = (1 + g * lam) * y / (1 - g * lam)
            ys.append(y)
        for _ in range(5):
            ys.append((18 * ys[-1] - 9 * ys[-2] + 2 * ys[-3]) / 11 / (1 - 6 * dt / 11 * lam))
        for n in range(1, len(ys)):
            s = o.step(s)
            for got, f in ((s.A[1], fy), (s.A[2], fz)):
                err = np.max(np.abs(got[:, iy, iz] - ys[n] * eps * f * x)) / eps
                assert err < 1e-11, (kind, n, err)
            assert np.max(np.abs(s.A[0])) < 1e-15 * eps


def test_curl_and_tc_forward():
    o = make()
    rng = np.random.default_rng(2)
    s = o.solve(o.initial_state(), 3)
    B = o.B_from_A(s.A)
    bp = o.backward_B(B, padded=False)
    dD = lambda c, **k: o.tp_backward(c, o.SD, padded=False, **k)  # noqa: E731
    dC = lambda c, **k: o.tp_backward(c, None, padded=False, **k)  # noqa: E731
    A = s.A
    ref = (dC(A[2], k1=1) - dC(A[1], k2=1), dD(A[0], k2=1) - dC(A[2], kx=1), dC(A[1], kx=1) - dD(A[0], k1=1))
    for got, want in zip(bp, ref):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    # TC forward of band-limited nodal data on the padded grid is exact
    c = o.mask(rng.standard_normal(A[1].shape) + 1j * rng.standard_normal(A[1].shape))
    c[:, :, 0] = c[:, :, 0].real
    phys = o.tp_backward(c, None, padded=True)
    back = o.tp_forward_C(np.real(phys).astype(complex))
    herm = o.mask(0.5 * (c + np.conj(np.roll(np.flip(c, (1, 2)), 1, (1, 2)))))
    assert np.max(np.abs(back - herm)) < 1e-12 * np.max(np.abs(herm))
