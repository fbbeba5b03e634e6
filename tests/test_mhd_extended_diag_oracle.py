"""Pins of the oracle's rotating-shear diagnostic set (shearpy :178-513).

A manufactured state whose fields are known in closed form:
    u = (a0 (1-x^2)^2 cos ky y,  a1 (1-x^2) cos ky y,  a2 (1-x^2) cos kz z),
    A = c ((1-x^2) cos kz z,  (1-x^2) cos kz z,  0),
so b = curl A = c (kz f sin kz z, -kz f sin kz z, f' cos kz z) with f = 1 - x^2,
and J, the vorticity of the total velocity (base flow -S x) and every
electric-field part follow by hand.  The expected values below are evaluated
pointwise from those formulas on the unpadded grid (Fejer-I weights in x,
uniform in y and z), with no spectral transform, so the oracle's
coefficient-space path (B from A, backward_primitive, the TC forward and its
derivatives, the wall traces) is checked against mathematics it does not
share.  Every product has Chebyshev degree below N0 and Fourier content
|k| <= 2 < N/2, so the oracle's projections are exact and the agreement is to
rounding.  A sign slip in a curl, a missing base flow or background field, the
wrong wall, or a dropped factor of the stresses changes one of the keys.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.kmm import fejer_weights
from oracle.mhd import MHDOracleState, RotatingShearMHDOracle
from oracle.mhd_insulating import InsulatingMHDOracle

N = (17, 8, 8)
S, OMEGA, B0 = 0.7, 0.4, (0.05, -0.03, 0.11)
A0, A1, A2, CA = 0.3, -0.2, 0.25, 0.4
ETA_RM, NU_RE = 250.0, 400.0


def _oracle(cls):
    return cls(N=N, Re=NU_RE, Rm=ETA_RM, omega=OMEGA, shear_rate=S, background_b=B0, dt=0.01,
               perturbation_amplitude=0.0, magnetic_amplitude=0.0)


def _state(o):
    x, y, z = o.mesh()
    shp = (o.N0, N[1], N[2])
    ky, kz = 2 * np.pi / o.L1, 2 * np.pi / o.L2
    f = 1.0 - x**2
    u0 = np.broadcast_to(A0 * f**2 * np.cos(ky * y), shp).astype(complex)
    u1 = np.broadcast_to(A1 * f * np.cos(ky * y), shp).astype(complex)
    u2 = np.broadcast_to(A2 * f * np.cos(kz * z), shp).astype(complex)
    ax = np.broadcast_to(CA * f * np.cos(kz * z), shp).astype(complex)
    uc = (o.mask(o.tp_forward(u0, o.SB, o.MB1)), o.mask(o.tp_forward(u1, o.SD, o.MD1)),
          o.mask(o.tp_forward(u2, o.SD, o.MD1)))
    g = o.mask(1j * o.K1 * uc[2] - 1j * o.K2 * uc[1])
    from oracle.kmm import OracleState

    flow = o.reset_history(OracleState(u=uc, g=g))
    a_td = o.mask(o.tp_forward(ax, o.SD, o.MD1))
    if isinstance(o, InsulatingMHDOracle):
        a_tc = o.mask(o.tc_forward(ax))
        A = (a_td, a_tc, np.zeros_like(a_tc))
    else:
        A = (a_td, a_td.copy(), np.zeros_like(a_td))
    return MHDOracleState(flow=flow, A=A)


def _expected(o):
    """Every key from the closed-form fields on the unpadded grid."""
    xs = o.x
    ny, nz = N[1], N[2]
    y = np.linspace(0, o.L1, ny, endpoint=False)
    z = np.linspace(0, o.L2, nz, endpoint=False)
    ky, kz = 2 * np.pi / o.L1, 2 * np.pi / o.L2
    V = 2.0 * o.L1 * o.L2

    def fields(x):
        X, Y, Z = np.meshgrid(np.atleast_1d(x), y, z, indexing="ij")
        f, fp, fpp = 1 - X**2, -2 * X, -2.0 + 0 * X
        cy, sy, cz, sz = np.cos(ky * Y), np.sin(ky * Y), np.cos(kz * Z), np.sin(kz * Z)
        u = (A0 * f**2 * cy, A1 * f * cy, A2 * f * cz)
        grad_u = ((-4 * A0 * X * f * cy, -A0 * ky * f**2 * sy, 0 * X),
                  (A1 * fp * cy, -A1 * ky * f * sy, 0 * X),
                  (A2 * fp * cz, 0 * X, -A2 * kz * f * sz))
        b = (CA * kz * f * sz, -CA * kz * f * sz, CA * fp * cz)
        grad_b = ((CA * kz * fp * sz, 0 * X, CA * kz**2 * f * cz),
                  (-CA * kz * fp * sz, 0 * X, -CA * kz**2 * f * cz),
                  (CA * fpp * cz, 0 * X, -CA * kz * fp * sz))
        J = (grad_b[2][1] - grad_b[1][2], grad_b[0][2] - grad_b[2][0], grad_b[1][0] - grad_b[0][1])
        ut = (u[0], u[1] - S * X, u[2])
        grad_ut = (grad_u[0], (grad_u[1][0] - S, grad_u[1][1], grad_u[1][2]), grad_u[2])
        w = (grad_ut[2][1] - grad_ut[1][2], grad_ut[0][2] - grad_ut[2][0], grad_ut[1][0] - grad_ut[0][1])
        bt = tuple(bi + b0 for bi, b0 in zip(b, B0))
        cross = (ut[1] * bt[2] - ut[2] * bt[1], ut[2] * bt[0] - ut[0] * bt[2], ut[0] * bt[1] - ut[1] * bt[0])
        ideal = tuple(-c for c in cross)
        res = tuple(o.eta * j for j in J)
        # div(u x b) = b . curl u - u . curl b
        div_ideal = -(sum(bi * wi for bi, wi in zip(bt, w)) - sum(ui * ji for ui, ji in zip(ut, J)))
        # divergence of curl A
        div_a = CA * fp * cz
        return dict(u=u, grad_u=grad_u, b=b, bt=bt, grad_b=grad_b, J=J, ideal=ideal, res=res,
                    div_ideal=div_ideal, div_a=div_a)

    F = fields(xs)
    wq = fejer_weights(o.N0)[:, None, None] * (o.L1 / ny) * (o.L2 / nz)
    integ = lambda v: float(np.sum(v * wq))  # noqa: E731
    W = fields(np.array([-1.0, 1.0]))
    tot = tuple(a + b for a, b in zip(F["ideal"], F["res"]))
    tot_w = tuple(a + b for a, b in zip(W["ideal"], W["res"]))
    reynolds = integ(F["u"][0] * F["u"][1]) / V
    maxwell = -integ(F["bt"][0] * F["bt"][1]) / V
    emag_total = sum(integ(b * b) for b in F["bt"])
    mean_b = tuple(integ(b) / V for b in F["bt"])
    kz1_u = 2 * abs(A2) / math.sqrt(15.0)
    kz1_b = abs(CA) * math.sqrt(8 * kz**2 / 15 + 2.0 / 3.0)
    out = {
        "reynolds_stress": reynolds, "maxwell_stress": maxwell, "total_stress": reynolds + maxwell,
        "Emag_total": emag_total, "mean_bx": mean_b[0], "mean_by": mean_b[1], "mean_bz": mean_b[2],
        "mag_energy_fluct_total": sum(integ(b * b) for b in F["b"]),
        "mag_energy_mean": V * sum(m * m for m in mean_b),
        "dissipation_kinetic": o.nu * sum(integ(g * g) for gu in F["grad_u"] for g in gu),
        "dissipation_magnetic": o.eta * sum(integ(g * g) for gb in F["grad_b"] for g in gb),
        "channel_kz1_velocity_rms": kz1_u, "channel_kz1_magnetic_rms": kz1_b,
        "bmax": max(float(np.max(np.abs(b))) for b in F["b"]),
        "bmax_total": max(float(np.max(np.abs(b))) for b in F["bt"]),
        "electric_ideal_l2": math.sqrt(sum(integ(e * e) for e in F["ideal"])),
        "electric_resistive_l2": math.sqrt(sum(integ(e * e) for e in F["res"])),
        "electric_total_l2": math.sqrt(sum(integ(e * e) for e in tot)),
        "divergence_e_ideal_l2": math.sqrt(integ(F["div_ideal"] ** 2)),
        "divergence_e_l2": math.sqrt(integ(F["div_ideal"] ** 2)),
        "divergence_a_l2": math.sqrt(integ(F["div_a"] ** 2)),
        "divergence_a_wall_linf": float(np.max(np.abs(W["div_a"]))),
        "electric_ideal_wall_tangential_linf": max(float(np.max(np.abs(W["ideal"][i]))) for i in (1, 2)),
        "electric_resistive_wall_tangential_linf": max(float(np.max(np.abs(W["res"][i]))) for i in (1, 2)),
        "electric_wall_tangential_linf": max(float(np.max(np.abs(tot_w[i]))) for i in (1, 2)),
        "faraday_mean_by_tendency": float(np.mean(tot_w[2][1]) - np.mean(tot_w[2][0])) / 2.0,
        "faraday_mean_bz_tendency": -float(np.mean(tot_w[1][1]) - np.mean(tot_w[1][0])) / 2.0,
        "mean_bx_trace": B0[0], "mean_by_trace": B0[1], "mean_bz_trace": B0[2],
    }
    for label, fs in (("ideal", F["ideal"]), ("resistive", F["res"]), ("total", tot)):
        for axis, f in zip("xyz", fs):
            out[f"mean_e{axis}_{label}"] = integ(f) / V
    out["mag_energy_fluct"] = out["Emag_total"] - out["mag_energy_mean"]
    out["alpha_Sh"] = out["total_stress"] / S**2
    out["alpha"] = out["total_stress"] / B0[2] ** 2
    return out


@pytest.mark.parametrize("cls", [RotatingShearMHDOracle, InsulatingMHDOracle])
def test_extended_diagnostics_match_closed_forms(cls):
    o = _oracle(cls)
    d = o.extended_diagnostics(_state(o))
    exp = _expected(o)
    scale = max(abs(v) for v in exp.values())
    for key, ref in exp.items():
        assert abs(d[key] - ref) <= 1e-12 * max(abs(ref), 1e-3 * scale), (key, d[key], ref)
    # closed forms that the pointwise evaluation above reproduces as numbers
    kz = 2 * np.pi / o.L2
    assert abs(d["divergence_a_wall_linf"] - 2 * abs(CA)) < 1e-13
    # u = 0 and b_x = 0 at both walls, so the ideal field there is u_tot,y B0x = -S x B0x
    assert abs(d["electric_ideal_wall_tangential_linf"] - abs(S * B0[0])) < 1e-13
    assert abs(d["maxwell_stress"] - CA**2 * kz**2 * 4 / 15 + B0[0] * B0[1]) < 1e-13
    assert d["divergence_e_resistive_l2"] < 1e-13  # div curl = 0
    assert d["mean_b_trace_mismatch_linf"] < 1e-14


def test_planted_mistakes_fail_the_closed_forms():
    """Planted-mistake copies of the oracle (Maxwell stress of the fluctuation
    field only; the walls swapped in the trace; the base flow left out of the
    electric field) each miss the closed forms."""
    o = _oracle(RotatingShearMHDOracle)
    exp = _expected(o)
    s = _state(o)
    # insulating walls admit a mean A_y = d x (T_1 in TC): b_z gains d, and the
    # wall trace of A_y gives mean_bz = B0z + d only with the walls in order
    d_mean = 0.02
    oi = _oracle(InsulatingMHDOracle)
    si = _state(oi)
    ay = si.A[1].copy()
    ay[1, 0, 0] += d_mean
    si = MHDOracleState(flow=si.flow, A=(si.A[0], ay, si.A[2]))
    assert abs(oi.extended_diagnostics(si)["mean_bz_trace"] - (B0[2] + d_mean)) < 1e-14
    assert abs(oi.extended_diagnostics(si)["mean_bz"] - (B0[2] + d_mean)) < 1e-14

    class NoB0(RotatingShearMHDOracle):
        def extended_diagnostics(self, st):
            saved = self.background_b
            d = super().extended_diagnostics(st)
            self.background_b = (0.0, 0.0, 0.0)
            d2 = super().extended_diagnostics(st)
            self.background_b = saved
            d["maxwell_stress"] = d2["maxwell_stress"]
            return d

    class SwappedWalls(InsulatingMHDOracle):
        def wall_trace(self, c, S_, derivative=0):
            return super().wall_trace(c, S_, derivative)[::-1]

    class NoBaseFlow(RotatingShearMHDOracle):
        def extended_diagnostics(self, st):
            saved = self.dUb
            self.dUb = 0.0
            d = super().extended_diagnostics(st)
            self.dUb = saved
            return d

    for cls, key in ((NoB0, "maxwell_stress"), (NoBaseFlow, "electric_ideal_l2")):
        bad = cls(N=N, Re=NU_RE, Rm=ETA_RM, omega=OMEGA, shear_rate=S, background_b=B0, dt=0.01,
                  perturbation_amplitude=0.0, magnetic_amplitude=0.0)
        d = bad.extended_diagnostics(s)
        assert abs(d[key] - exp[key]) > 1e-6 * max(abs(exp[key]), 1e-12), (cls.__name__, key)
    bad = SwappedWalls(N=N, Re=NU_RE, Rm=ETA_RM, omega=OMEGA, shear_rate=S, background_b=B0, dt=0.01,
                       perturbation_amplitude=0.0, magnetic_amplitude=0.0)
    assert abs(bad.extended_diagnostics(si)["mean_bz_trace"] - (B0[2] + d_mean)) > 1e-3


def test_energy_split_identity_on_a_stepped_state():
    """Emag_total = V |mean_b|^2 + (int |b|^2 - V |<b>|^2) + ... : the split of
    the total field's energy into the mean part and the fluctuation part holds
    for a seeded, stepped state (both wall types)."""
    for cls in (RotatingShearMHDOracle, InsulatingMHDOracle):
        o = cls(N=(17, 16, 16), Re=400, Rm=300, omega=0.3, shear_rate=1.0, background_b=(0.05, 0.02, 0.1),
                dt=0.01, perturbation_amplitude=0.05, magnetic_amplitude=0.05)
        s = o.solve(o.initial_state(), 3)
        d = o.extended_diagnostics(s)
        V = 2.0 * o.L1 * o.L2
        mb = np.array([d["mean_bx"], d["mean_by"], d["mean_bz"]]) - np.array(o.background_b)
        lhs = d["Emag_total"] - d["mag_energy_mean"]
        rhs = d["mag_energy_fluct_total"] - V * float(mb @ mb)
        assert abs(lhs - rhs) <= 1e-12 * d["Emag_total"]
        assert d["mean_b_trace_mismatch_linf"] < 1e-13
