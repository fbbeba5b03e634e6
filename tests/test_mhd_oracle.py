"""Pins of the MHD oracle (oracle/mhd.py) against the reference's published
FP64 results and against identities the mathematics fixes.

* Published goldens (slow): diagnostics and 38-leaf state checksums of
  ``PlaneCouetteMHDJax`` after 325 SBDF3 steps at 65x64x64 and 50 at 97x96x96
  (``benchmark-results/mhd-sbdf3-rotational-*.json``; fixture
  ``tests/golden/reference_mhd_published.json``).
* B = curl A against derivatives of the backward transforms of A at the nodes.
* The TD projection inside J = curl B against forward(backward(.)) by
  Gauss quadrature (the identity ``project_from_orthogonal`` states,
  ``tensorproductspace.py:706-713``).
* A = 0 reproduces the hydrodynamic oracle exactly.
* A decaying Ax mode whose EMF vanishes identically (ky = 0) follows the
  scalar IMEXRK3/SBDF3 recurrence of its discrete eigenvalue of
  eta M_A^{-1} L_A, which pins the A operators, the stage and history
  coefficients and the level order of the A history.
"""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg as sla

from oracle.kmm import IMEXRK3_A, IMEXRK3_B, PCFOracle
from oracle.mhd import MHDOracle, MHDOracleState, RotatingShearMHDOracle

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_mhd_published.json").read_text())


def _rand_field(o, rows, seed, K=4):
    """Hermitian-symmetric, Nyquist-masked random coefficients (real physical field)."""
    rng = np.random.default_rng(seed)
    shape = (o.N0, o.N[1], o.N[2])
    y = np.linspace(0, o.L1, o.N[1], endpoint=False)[None, :, None]
    z = np.linspace(0, o.L2, o.N[2], endpoint=False)[None, None, :]
    x = o.xb
    f = np.zeros(shape)
    for _ in range(6):
        a, b = rng.integers(0, K, 2)
        ph1, ph2 = rng.uniform(0, 2 * np.pi, 2)
        f = f + rng.normal() * (1 - x**2) * np.cos(2 * np.pi * a * y / o.L1 + ph1) * np.cos(
            2 * np.pi * b * z / o.L2 + ph2) * (1 + 0.3 * x)
    c = o.mask(o.tp_forward(f.astype(complex), o.SD, o.MD1))
    assert c.shape[0] == rows
    return c


def test_curl_matches_nodal_derivatives():
    o = MHDOracle(N=(17, 16, 16), Re=400.0, dt=1e-3)
    A = tuple(_rand_field(o, o.ND, s) for s in (1, 2, 3))
    B = o.B_from_A(A)
    bp = o.backward_B(B, padded=False)
    d = lambda c, **k: o.tp_backward(c, o.SD, padded=False, **k)  # noqa: E731
    ref = (d(A[2], k1=1) - d(A[1], k2=1), d(A[0], k2=1) - d(A[2], kx=1), d(A[1], kx=1) - d(A[0], k1=1))
    for got, want in zip(bp, ref):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    # div B = 0 to rounding
    assert o.diagnostics(MHDOracleState(flow=o.initial_state().flow, A=A))["divB_L2"] < 1e-12


def test_current_projection_is_quadrature_forward():
    """J_y, J_z in TD equal TD.forward of the nodal curl of B (exact quadrature:
    the products have degree <= 2 n0 - 3)."""
    o = MHDOracle(N=(17, 16, 16), Re=400.0, dt=1e-3)
    A = tuple(_rand_field(o, o.ND, s) for s in (4, 5, 6))
    B = o.B_from_A(A)
    J = o.J_from_B(B)
    bx = o.tp_backward(B[0], o.SD, padded=False)
    bT = lambda c, **k: o.tp_backward(c, None, padded=False, **k)  # noqa: E731
    dbx = lambda **k: o.tp_backward(B[0], o.SD, padded=False, **k)  # noqa: E731
    jy = dbx(k2=1) - bT(B[2], kx=1)
    jz = bT(B[1], kx=1) - dbx(k1=1)
    jx = bT(B[2], k1=1) - bT(B[1], k2=1)
    del bx
    for got, nodal in ((J[1], jy), (J[2], jz)):
        want = o.mask(o.tp_forward(nodal, o.SD, o.MD1))
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    jxn = o.tp_backward(J[0], None, padded=False)
    assert np.max(np.abs(jxn - jx)) <= 1e-12 * np.max(np.abs(jx))


def test_zero_field_reproduces_hydro_oracle():
    kw = dict(N=(17, 16, 16), Re=400.0, dt=2e-3, perturbation_amplitude=0.05)
    m = MHDOracle(magnetic_amplitude=0.0, **kw)
    h = PCFOracle(**kw)
    sm, sh = m.initial_state(), h.initial_state()
    sm, sh = m.solve(sm, 5), h.solve(sh, 5)
    for a, b in zip(sm.flow.leaves(), sh.leaves()):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    assert all(np.max(np.abs(a)) == 0 for a in sm.A)


def test_rotating_shear_reduces_to_conducting_mhd():
    """With Omega = 0, S = -1 (U_b = x) and the same B0 the rotating-shear
    oracle is the conducting-wall oracle bit for bit: the base-flow slope and
    the forcing hook are the only differences."""
    kw = dict(N=(17, 16, 16), Re=400.0, Rm=300.0, dt=2e-3, perturbation_amplitude=0.05, magnetic_amplitude=0.2,
              background_b=(0.0, 0.02, 0.1))
    r = RotatingShearMHDOracle(omega=0.0, shear_rate=-1.0, **kw)
    m = MHDOracle(U_wall=1.0, **kw)
    s0 = m.initial_state()
    sr, sm = r.solve(s0, 5), m.solve(s0, 5)
    for a, b in zip(sr.flow.leaves() + list(sr.A), sm.flow.leaves() + list(sm.A)):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    # the two classes' ax_yz seeds are the same field (written independently)
    for a, b in zip(r.initial_state().A, m.initial_state().A):
        assert np.max(np.abs(a - b)) <= 1e-15 * max(1.0, np.max(np.abs(b)))


def test_coriolis_does_no_work():
    """u . F_Coriolis = 0 pointwise for any velocity, and the forcing is linear
    in Omega with the sign pattern (-u1, +u0, 0)."""
    rng = np.random.default_rng(4)
    o = RotatingShearMHDOracle(N=(9, 8, 8), omega=0.7, shear_rate=1.5)
    up = tuple(rng.standard_normal((9, 12, 12)) for _ in range(3))
    zero = tuple(np.zeros_like(u) for u in up)
    f = o.flow_forcing(zero, up)
    assert np.max(np.abs(sum(a * b for a, b in zip(f, up)))) < 1e-14
    assert np.max(np.abs(f[2])) == 0.0
    o2 = RotatingShearMHDOracle(N=(9, 8, 8), omega=1.4, shear_rate=1.5)
    f2 = o2.flow_forcing(zero, up)
    assert all(np.allclose(b, 2 * a, rtol=0, atol=1e-14) for a, b in zip(f, f2))


def test_A_decay_follows_scalar_recurrence():
    """Ax = eps f(x) cos(z), f the least-damped eigenvector of eta M^{-1}(C2 - kz^2 C0):
    b = (0, dAx/dz, 0) and u_tot x b = 0 exactly, so the A update is linear and
    every step multiplies the mode by the integrator's scalar amplification."""
    dt, Rm = 0.01, 50.0
    o = MHDOracle(N=(17, 8, 8), Re=400.0, Rm=Rm, dt=dt, perturbation_amplitude=0.0, magnetic_amplitude=0.0)
    kz = o.K2[0, 0, 1]
    lam, vec = sla.eig(o.eta * (o.D_C2 - kz**2 * o.D_C0), o.D_C0)
    i = int(np.argmax(lam.real))
    lam, f = float(lam[i].real), np.real(vec[:, i])
    eps = 1e-7
    Ax = np.zeros((o.ND, o.N[1], o.N[2]), complex)
    Ax[:, 0, 1] = Ax[:, 0, -1] = 0.5 * eps * f
    s0 = o.initial_state()  # zero flow perturbation and zero A at these amplitudes
    s = o.reset_mhd_history(MHDOracleState(flow=s0.flow, A=(Ax, 0 * Ax, 0 * Ax)))
    # scalar reference: IMEXRK3 stages (1 - g l) y' = (1 + g l) y, then SBDF3
    ys = [1.0]
    y = 1.0
    for _ in range(2):
        for rk in range(3):
            g = (IMEXRK3_A[rk] + IMEXRK3_B[rk]) * dt / 2.0
            y = (1 + g * lam) * y / (1 - g * lam)
        ys.append(y)
    for _ in range(6):
        y = (18 * ys[-1] - 9 * ys[-2] + 2 * ys[-3]) / 11 / (1 - 6 * dt / 11 * lam)
        ys.append(y)
    for k in range(1, len(ys)):
        s = o.step(s)
        want = ys[k] * 0.5 * eps * f
        err = np.max(np.abs(s.A[0][:, 0, 1] - want)) / np.max(np.abs(0.5 * eps * f))
        assert err < 1e-11, (k, err, ys[k])
        assert np.max(np.abs(s.A[1])) < 1e-15 * eps and np.max(np.abs(s.A[2])) < 1e-15 * eps


def test_A_decay_follows_ars222_recurrence():
    """The same resistive eigenmode under IMEXRK222: each step is
    y1 = y / (1 - g h l), y' = (y + (1 - g) h l y1) / (1 - g h l) with
    g = (2 - sqrt 2) / 2 (pins the A stages, the explicit eta-Laplacian of the
    second stage and its coefficient)."""
    from oracle.kmm import ARS222_G as G

    dt, Rm = 0.01, 50.0
    o = MHDOracle(N=(17, 8, 8), Re=400.0, Rm=Rm, dt=dt, perturbation_amplitude=0.0, magnetic_amplitude=0.0,
                  time_integrator="IMEXRK222")
    kz = o.K2[0, 0, 1]
    lam, vec = sla.eig(o.eta * (o.D_C2 - kz**2 * o.D_C0), o.D_C0)
    i = int(np.argmax(lam.real))
    lam, f = float(lam[i].real), np.real(vec[:, i])
    eps = 1e-7
    Ax = np.zeros((o.ND, o.N[1], o.N[2]), complex)
    Ax[:, 0, 1] = Ax[:, 0, -1] = 0.5 * eps * f
    s0 = o.initial_state()
    s = o.reset_mhd_history(MHDOracleState(flow=s0.flow, A=(Ax, 0 * Ax, 0 * Ax)))
    y, h = 1.0, dt
    for k in range(6):
        y1 = y / (1 - G * h * lam)
        y = (y + (1 - G) * h * lam * y1) / (1 - G * h * lam)
        s = o.step(s)
        err = np.max(np.abs(s.A[0][:, 0, 1] - y * 0.5 * eps * f)) / np.max(np.abs(0.5 * eps * f))
        assert err < 1e-11, (k, err, y)


def test_A_decay_follows_crank_nicolson_recurrence_with_variable_dt():
    """The same resistive eigenmode under CNAB2 with dt changed between steps:
    the A nonlinear rows vanish, so each step is y' = (1 + h l / 2) y / (1 - h l / 2)
    at that step's h (pins the dt/2 weight of the explicit eta-Laplacian and of
    the implicit systems)."""
    dt, Rm = 0.01, 50.0
    o = MHDOracle(N=(17, 8, 8), Re=400.0, Rm=Rm, dt=dt, perturbation_amplitude=0.0, magnetic_amplitude=0.0,
                  time_integrator="CNAB2")
    kz = o.K2[0, 0, 1]
    lam, vec = sla.eig(o.eta * (o.D_C2 - kz**2 * o.D_C0), o.D_C0)
    i = int(np.argmax(lam.real))
    lam, f = float(lam[i].real), np.real(vec[:, i])
    eps = 1e-7
    Ax = np.zeros((o.ND, o.N[1], o.N[2]), complex)
    Ax[:, 0, 1] = Ax[:, 0, -1] = 0.5 * eps * f
    s0 = o.initial_state()
    s = MHDOracleState(flow=s0.flow, A=(Ax, 0 * Ax, 0 * Ax))
    y = 1.0
    for k, h in enumerate((dt, dt, 0.5 * dt, 2.0 * dt, 0.7 * dt)):
        o.set_dt(h)
        y = (1 + 0.5 * h * lam) * y / (1 - 0.5 * h * lam)
        s = o.step(s)
        err = np.max(np.abs(s.A[0][:, 0, 1] - y * 0.5 * eps * f)) / np.max(np.abs(0.5 * eps * f))
        assert err < 1e-11, (k, err, y)
        assert s.flow.previous_dt == h and s.flow.have_old == 1.0


def test_zero_field_cnab2_reproduces_hydro_oracle():
    kw = dict(N=(17, 16, 16), Re=400.0, dt=2e-3, perturbation_amplitude=0.05, time_integrator="CNAB2")
    m = MHDOracle(magnetic_amplitude=0.0, **kw)
    h = PCFOracle(**kw)
    sm, sh = m.initial_state(), h.initial_state()
    sm = MHDOracleState(flow=sh, A=sm.A)  # no history: the CNAB2 bootstrap
    sm, sh = m.solve(sm, 4), h.solve(sh, 4)
    for a, b in zip(list(sm.flow.u) + [sm.flow.g] + list(sm.flow.nonlinear_old),
                    list(sh.u) + [sh.g] + list(sh.nonlinear_old)):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    assert all(np.max(np.abs(a)) == 0 for a in sm.A + sm.nonlinear_A_old)
    # the CNAB2 MHD state keeps the step's nonlinear rows only (pcf_mhd_jax.py:383-387)
    assert sm.flow.solution_old is None and sm.A_old is None and sm.nonlinear_A_older is None
    assert sm.flow.have_old == 1.0 and sm.flow.previous_dt == 2e-3


def test_zero_field_ars222_reproduces_hydro_oracle():
    kw = dict(N=(17, 16, 16), Re=400.0, dt=2e-3, perturbation_amplitude=0.05, time_integrator="IMEXRK222")
    m = MHDOracle(magnetic_amplitude=0.0, **kw)
    h = PCFOracle(**kw)
    sm, sh = m.initial_state(), h.initial_state()
    sm, sh = m.solve(sm, 4), h.solve(sh, 4)
    for a, b in zip(list(sm.flow.u) + [sm.flow.g], list(sh.u) + [sh.g]):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    assert all(np.max(np.abs(a)) == 0 for a in sm.A)
    # the MHD ARS step keeps every history field (pcf_mhd_jax.py:482-486)
    assert sm.flow.solution_old is not None and sm.nonlinear_A_old is not None


@pytest.mark.slow
@pytest.mark.parametrize("case", GOLDEN["cases"], ids=lambda c: "x".join(map(str, c["N"])))
def test_published_mhd_golden(case):
    cfg = GOLDEN["config"]
    o = MHDOracle(N=tuple(case["N"]), Re=cfg["Re"], Rm=cfg["Rm"], dt=cfg["dt"],
                  perturbation_amplitude=cfg["velocity_amplitude"], magnetic_amplitude=cfg["magnetic_amplitude"])
    s = o.solve(o.initial_state(), case["steps"])
    d = o.diagnostics(s)
    for k in ("Epert", "Etot", "Emag"):
        assert math.isclose(d[k], case["diagnostics"][k], rel_tol=1e-10), (k, d[k], case["diagnostics"][k])
    for k in ("wall_shear_top", "wall_shear_bot", "mean_shear"):
        assert abs(d[k] - case["diagnostics"][k]) < 1e-12, k
    assert d["divB_L2"] < 1e-15 and d["divL2"] < 1e-14
    c = o.checksum(s)
    ref = case["checksum"]
    assert c["leaf_count"] == ref["leaf_count"] and c["element_count"] == ref["element_count"]
    assert math.isclose(c["l2_norm"], ref["l2_norm"], rel_tol=1e-11)
    assert math.isclose(c["real_sum"], ref["real_sum"], rel_tol=1e-9)
