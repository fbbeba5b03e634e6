"""CPU FP64 reverse mode of the conducting-wall MHD oracle step — TEST INFRASTRUCTURE ONLY.

The reference differentiates ``PlaneCouetteMRIShearpyJax`` with SBDF3 through
``jax.value_and_grad`` over ``lax.scan(jax.checkpoint(step))``
(``production/autograd_cost_benchmark.py:56-61,228-330,395``).  jax is not
available here, so this module states that derivative by hand: the
vector-Jacobian product of every stage of ``oracle.mhd.MHDOracle.step`` (and of
its rotating-shear subclass, whose Coriolis forcing and background field
enter through the same hooks), for A in TD^3 (conducting walls), composed in
reverse.  Conventions as ``oracle.adjoint``: cotangents for the real inner
product Re sum conj(a) b; a linear stage propagates by its conjugate
transpose.

Pinned by central finite differences of the oracle forward along seeded
directions in every leaf (flow, A and their histories) and by the adjoint
identity of each linear stage (``tests/test_adjoint_mhd_oracle.py``).
"""

from __future__ import annotations

from dataclasses import replace

import numpy as np

from .adjoint import OracleAdjoint, _add, _add_tuple, _cross
from .kmm import IMEXRK3_A, IMEXRK3_B, OracleState, cheb_backward, cheb_derivative
from .mhd import MHDOracle, MHDOracleState


class MHDOracleAdjoint(OracleAdjoint):
    """Vector-Jacobian products of ``MHDOracle.step`` (conducting walls, SBDF3
    with the IMEXRK3 start-up) and of the energy objective Epert + Emag."""

    def __init__(self, o: MHDOracle):
        super().__init__(o)
        n0 = o.N0
        eye = np.eye(n0)
        for kx in (0, 1):
            a = eye if not kx else cheb_derivative(eye, kx)
            self.Xb[("C", kx)] = cheb_backward(a, n0)
        # P_D = M_D^{-1} S_D diag(h): the TD projection of orthogonal coefficients
        self.PD = np.linalg.solve(o.MD1, o.SD * o.h[None, :])

    def _name(self, S):
        if S is None:
            return "C"
        return "B" if S is self.o.SB else "D"

    # ------------------------------------------------------------ linear pieces
    def _dcoef_adj(self, c, kx=0, k1=0, k2=0):
        """Adjoint of ``_dcoef`` (wavenumber factors, then the x-derivative)."""
        o = self.o
        z = self._x(self.D.T, c) if kx else c
        if k2:
            K = o.K2e if k2 % 2 else o.K2
            z = np.conj((1j * K) ** k2) * z
        if k1:
            K = o.K1e if k1 % 2 else o.K1
            z = np.conj((1j * K) ** k1) * z
        return z

    def _proj_adj(self, c):
        """Adjoint of ``project_D`` (M_D^{-1} S_D diag(h) along x)."""
        return self._x(self.PD.T, c)

    def B_from_A_adj(self, mB):
        """Adjoint of ``B_from_A`` (``examples/pcf_mhd_jax.py:269-289``)."""
        o = self.o
        mbx = self._proj_adj(o.mask(mB[0]))
        mby, mbz = o.mask(mB[1]), o.mask(mB[2])
        d = self._dcoef_adj
        mAO0 = d(mby, k2=1) - d(mbz, k1=1)
        mAO1 = -d(mbx, k2=1) + d(mbz, kx=1)
        mAO2 = d(mbx, k1=1) - d(mby, kx=1)
        return tuple(self._x(o.SD, m) for m in (mAO0, mAO1, mAO2))

    def J_from_B_adj(self, mJ):
        """Adjoint of ``J_from_B`` (``examples/pcf_mhd_jax.py:291-311``)."""
        o = self.o
        mjx = o.mask(mJ[0])
        mjy, mjz = self._proj_adj(o.mask(mJ[1])), self._proj_adj(o.mask(mJ[2]))
        d = self._dcoef_adj
        mBO0 = d(mjy, k2=1) - d(mjz, k1=1)
        mBO1 = -d(mjx, k2=1) + d(mjz, kx=1)
        mBO2 = d(mjx, k1=1) - d(mjy, kx=1)
        return (self._x(o.SD, mBO0), mBO1, mBO2)

    def backward_B_adj(self, m, padded=True):
        A = self.tp_backward_adj
        return (A(m[0], self.o.SD, padded), A(m[1], None, padded), A(m[2], None, padded))

    def backward_J_adj(self, m, padded=True):
        A = self.tp_backward_adj
        return (A(m[0], None, padded), A(m[1], self.o.SD, padded), A(m[2], self.o.SD, padded))

    def A_mass_adj(self, r):
        """Adjoint of ``A_mass`` (M_A = W C0_D per component)."""
        o = self.o
        return tuple(self._terms_adj(c, [(o.D_C0, o.W)]) for c in r)

    def A_lap_adj(self, r):
        """Adjoint of ``A_lap`` (eta W (C2_D - k^2 C0_D) per component)."""
        o = self.o
        k2 = o.K1**2 + o.K2**2
        e = o.eta * o.W
        return tuple(self._terms_adj(c, [(o.D_C2, e), (o.D_C0, -e * k2)]) for c in r)

    def A_solve_adj(self, gamma, mu):
        """Adjoint of ``A_solve_gamma``: (M_A - gamma L_A)^{-T} per mode, masked."""
        o = self.o
        k2 = o.k2flat
        S = o.MA - gamma * o.LA
        St = np.swapaxes(S, 1, 2)
        out = []
        for c in mu:
            m = o._modes(o.mask(c))
            re = np.linalg.solve(St, m.real[:, :, None])[:, :, 0]
            im = np.linalg.solve(St, m.imag[:, :, None])[:, :, 0]
            out.append(o.mask(o._unmodes(re + 1j * im)))
        del k2
        return tuple(out)

    # ------------------------------------------------------------ nonlinear term
    def mhd_nonlinear_adj(self, s: MHDOracleState, mu_n, mu_nA):
        """Adjoint of (nonlinear_rhs(H), nonlinear_A(HA)) with (H, HA) =
        ``mhd_convection(s)`` (``examples/pcf_mhd_jax.py:339-357``;
        ``examples/pcf_mhd_mri_shearpy_jax.py:153-176``) at the state s:
        cotangents of the rows -> cotangents of (u, A)."""
        o = self.o
        SB, SD = o.SB, o.SD
        u0, u1, u2 = s.flow.u
        up = (o.tp_backward(u0, SB), o.tp_backward(u1, SD), o.tp_backward(u2, SD))
        w0 = o.tp_backward(u2, SD, k1=1) - o.tp_backward(u1, SD, k2=1)
        w1 = o.tp_backward(u0, SB, k2=1) - o.tp_backward(u2, SD, kx=1)
        w2 = o.tp_backward(u1, SD, kx=1) - o.tp_backward(u0, SB, k1=1)
        a = (up[0], up[1] + o.dUb * o.xb, up[2])
        b = (w0, w1, w2 + o.dUb)
        B = o.B_from_A(s.A)
        J = o.J_from_B(B)
        bp = tuple(x + b0 for x, b0 in zip(o.backward_B(B), o.background_b))
        jp = o.backward_J(J)
        mH = self.nonlinear_rhs_adj(mu_n)
        mHA = self.A_mass_adj(mu_nA)
        P = tuple(self.tp_forward_D_adj(o.mask(m)) for m in mH)  # cotangent of the physical forcing
        PE = tuple(self.tp_forward_D_adj(o.mask(m)) for m in mHA)  # cotangent of the physical EMF
        mP = tuple(-x for x in P)
        cj = lambda v: tuple(np.conj(x) for x in v)  # noqa: E731
        # -(a x w_tot) and the EMF a x b_tot: cotangents of the total velocity a
        ma = _add(_cross(cj(b), mP), _cross(cj(bp), PE))
        mb = _cross(mP, cj(a))
        # Coriolis (-2 Omega u1, 2 Omega u0, 0) on the fluctuation velocity
        omega = float(getattr(o, "omega", 0.0))
        if omega != 0.0:
            ma = (ma[0] + 2.0 * omega * P[1], ma[1] - 2.0 * omega * P[0], ma[2])
        # -(j x b_tot) and the EMF's b_tot
        mj = _cross(cj(bp), mP)
        mbp = _add(_cross(mP, cj(jp)), _cross(PE, cj(a)))
        A = self.tp_backward_adj
        mu0 = A(ma[0], SB) + A(mb[1], SB, k2=1) - A(mb[2], SB, k1=1)
        mu1 = A(ma[1], SD) - A(mb[0], SD, k2=1) + A(mb[2], SD, kx=1)
        mu2 = A(ma[2], SD) + A(mb[0], SD, k1=1) - A(mb[1], SD, kx=1)
        mB = _add(self.backward_B_adj(mbp), self.J_from_B_adj(self.backward_J_adj(mj)))
        mA = self.B_from_A_adj(mB)
        return (mu0, mu1, mu2), mA

    # ------------------------------------------------------------ step VJPs
    def mhd_step_vjp(self, s: MHDOracleState, mu: MHDOracleState) -> MHDOracleState:
        """Cotangent of the input of ``MHDOracle.step`` (SBDF3 / IMEXRK3 start-up)."""
        o = self.o
        dt = o.dt
        if s.flow.history_steps < 2:
            return self._imexrk3_mhd_vjp(s, mu)
        f, mf = s.flow, mu.flow
        gamma = dt * 6.0 / 11.0
        r = self.solve_adj(gamma, _add_tuple(self.reconstruct_adj(mf.u), (0, mf.g, 0, 0)))
        mr = self.mass_adj(r)
        rA = self.A_solve_adj(gamma, mu.A)
        mrA = self.A_mass_adj(rA)
        mcur_n = _add(tuple(dt * 18.0 / 11.0 * x for x in r), mf.nonlinear_old)
        mcur_nA = _add(tuple(dt * 18.0 / 11.0 * x for x in rA), mu.nonlinear_A_old)
        mu_u, mu_A = self.mhd_nonlinear_adj(s, mcur_n, mcur_nA)
        mu_u, mu_g = self._primary_adj(mu_u, np.zeros_like(f.g),
                                       _add(tuple(18.0 / 11.0 * x for x in mr), mf.solution_old))
        flow = replace(f, u=mu_u, g=mu_g,
                       nonlinear_old=_add(tuple(-dt * 18.0 / 11.0 * x for x in r), mf.nonlinear_older),
                       nonlinear_older=tuple(dt * 6.0 / 11.0 * x for x in r),
                       solution_old=_add(tuple(-9.0 / 11.0 * x for x in mr), mf.solution_older),
                       solution_older=tuple(2.0 / 11.0 * x for x in mr))
        return MHDOracleState(
            flow=flow,
            A=_add(_add(tuple(18.0 / 11.0 * x for x in mrA), mu.A_old), mu_A),
            nonlinear_A_old=_add(tuple(-dt * 18.0 / 11.0 * x for x in rA), mu.nonlinear_A_older),
            nonlinear_A_older=tuple(dt * 6.0 / 11.0 * x for x in rA),
            A_old=_add(tuple(-9.0 / 11.0 * x for x in mrA), mu.A_older),
            A_older=tuple(2.0 / 11.0 * x for x in mrA))

    def _mhd_stage_inputs(self, s: MHDOracleState):
        """(flow, A) entering each IMEXRK3 stage (``MHDOracle.step_imexrk3_mhd``)."""
        o = self.o
        dt = o.dt
        out = []
        u_stage, g_stage, A_stage = s.flow.u, s.flow.g, s.A
        prev = (np.zeros_like(s.flow.u[0]), np.zeros_like(s.flow.g), np.zeros(o.ND), np.zeros(o.ND))
        prevA = o.zero_A()
        for rk in range(3):
            st = MHDOracleState(flow=OracleState(u=u_stage, g=g_stage), A=A_stage)
            out.append(st)
            H, HA = o.mhd_convection(st)
            cur, curA = o.nonlinear_rhs(H), o.nonlinear_A(HA)
            a, b = IMEXRK3_A[rk], IMEXRK3_B[rk]
            gamma = (a + b) * dt / 2.0
            v00 = np.real(u_stage[1][:, 0, 0])
            w00 = np.real(u_stage[2][:, 0, 0])
            rhs_u = o.Mu_apply(u_stage[0]) + gamma * o.Lu_apply(u_stage[0]) + dt * (a * cur[0] + b * prev[0])
            rhs_g = o.Mg_apply(g_stage) + gamma * o.Lg_apply(g_stage) + dt * (a * cur[1] + b * prev[1])
            rhs_v = o.M00 @ v00 + gamma * (o.L00 @ v00) + dt * (a * cur[2] + b * prev[2])
            rhs_w = o.M00 @ w00 + gamma * (o.L00 @ w00) + dt * (a * cur[3] + b * prev[3])
            mA, lA = o.A_mass(A_stage), o.A_lap(A_stage)
            rhs_A = [mA[i] + gamma * lA[i] + dt * (a * curA[i] + b * prevA[i]) for i in range(3)]
            u0n, gn, vn, wn = o._solve(o.startup[rk], rhs_u, rhs_g, rhs_v, rhs_w)
            u_stage = o.reconstruct(u0n, gn, vn, wn)
            g_stage = gn
            A_stage = o.A_solve_gamma(gamma, rhs_A)
            prev, prevA = cur, curA
        return out

    def _imexrk3_mhd_vjp(self, s: MHDOracleState, mu: MHDOracleState) -> MHDOracleState:
        o = self.o
        dt = o.dt
        stages = self._mhd_stage_inputs(s)
        mu_u, mu_g, mu_Ast = mu.flow.u, mu.flow.g, mu.A
        rs, rsA = [None] * 3, [None] * 3
        for rk in (2, 1, 0):
            a, b = IMEXRK3_A[rk], IMEXRK3_B[rk]
            gamma = (a + b) * dt / 2.0
            r = self.solve_adj(gamma, _add_tuple(self.reconstruct_adj(mu_u), (0, mu_g, 0, 0)))
            rA = self.A_solve_adj(gamma, mu_Ast)
            rs[rk], rsA[rk] = r, rA
            lin = tuple(x + gamma * y for x, y in zip(self.mass_adj(r), self.lin_adj(r)))
            linA = tuple(x + gamma * y for x, y in zip(self.A_mass_adj(rA), self.A_lap_adj(rA)))
            mcur = tuple(dt * a * x for x in r)
            mcurA = tuple(dt * a * x for x in rA)
            if rk < 2:
                mcur = _add(mcur, tuple(dt * IMEXRK3_B[rk + 1] * x for x in rs[rk + 1]))
                mcurA = _add(mcurA, tuple(dt * IMEXRK3_B[rk + 1] * x for x in rsA[rk + 1]))
            if rk == 0:
                mcur = _add(mcur, mu.flow.nonlinear_old)
                mcurA = _add(mcurA, mu.nonlinear_A_old)
            mu_in, mA_in = self.mhd_nonlinear_adj(stages[rk], mcur, mcurA)
            mu_u, mu_g = self._primary_adj(mu_in, np.zeros_like(s.flow.g), lin)
            mu_Ast = _add(mA_in, linA)
        mu_u, mu_g = self._primary_adj(mu_u, mu_g, mu.flow.solution_old)
        zf = lambda grp: tuple(np.zeros_like(np.asarray(x)) for x in grp)  # noqa: E731
        flow = replace(s.flow, u=mu_u, g=mu_g, nonlinear_old=mu.flow.nonlinear_older,
                       nonlinear_older=zf(s.flow.nonlinear_old), solution_old=mu.flow.solution_older,
                       solution_older=zf(s.flow.solution_old))
        return MHDOracleState(flow=flow, A=_add(mu_Ast, mu.A_old), nonlinear_A_old=mu.nonlinear_A_older,
                              nonlinear_A_older=zf(s.A), A_old=mu.A_older, A_older=zf(s.A))

    # ------------------------------------------------------------ objective
    def mhd_energy(self, s: MHDOracleState) -> float:
        """Epert + Emag of the state (``diagnostics``, ``examples/pcf_mhd_jax.py:760-789``)."""
        d = self.o.diagnostics(s)
        return d["Epert"] + d["Emag"]

    def mhd_energy_grad(self, s: MHDOracleState):
        """Cotangents of Epert + Emag with respect to u and A."""
        o = self.o
        w = o._weights()
        mu = self.energy_grad(s.flow)
        bp = o.backward_B(o.B_from_A(s.A), padded=False)
        mB = self.backward_B_adj(tuple(2.0 * w * x for x in bp), padded=False)
        return mu, self.B_from_A_adj(mB)

    def zeros(self, s: MHDOracleState) -> MHDOracleState:
        from .adjoint import zeros_like_state

        z = lambda grp: tuple(np.zeros_like(np.asarray(x)) for x in grp)  # noqa: E731
        return MHDOracleState(flow=zeros_like_state(s.flow), A=z(s.A), nonlinear_A_old=z(s.A),
                              nonlinear_A_older=z(s.A), A_old=z(s.A), A_older=z(s.A))

    def mhd_value_and_grad(self, s0: MHDOracleState, steps: int):
        """Epert + Emag after ``steps`` steps and its cotangent w.r.t. every leaf of s0."""
        o = self.o
        traj = [s0]
        for _ in range(int(steps)):
            traj.append(o.step(traj[-1]))
        value = self.mhd_energy(traj[-1])
        mu_u, mu_A = self.mhd_energy_grad(traj[-1])
        mu = self.zeros(traj[-1])
        mu = replace(mu, flow=replace(mu.flow, u=mu_u), A=mu_A)
        for k in range(int(steps) - 1, -1, -1):
            mu = self.mhd_step_vjp(traj[k], mu)
        return value, mu

    def mhd_vjp(self, s0: MHDOracleState, steps: int, mu_final: MHDOracleState) -> MHDOracleState:
        o = self.o
        traj = [s0]
        for _ in range(int(steps) - 1):
            traj.append(o.step(traj[-1]))
        mu = mu_final
        for k in range(int(steps) - 1, -1, -1):
            mu = self.mhd_step_vjp(traj[k], mu)
        return mu


def mhd_leaves(s: MHDOracleState) -> list:
    out = list(s.flow.u) + [s.flow.g]
    for grp in (s.flow.nonlinear_old, s.flow.nonlinear_older, s.flow.solution_old, s.flow.solution_older):
        out.extend(np.asarray(x) for x in grp)
    for grp in (s.A, s.nonlinear_A_old, s.nonlinear_A_older, s.A_old, s.A_older):
        out.extend(np.asarray(x) for x in grp)
    return out


def mhd_pairing(mu: MHDOracleState, d: MHDOracleState) -> float:
    """<mu, d> = Re sum conj(mu) d over every leaf."""
    return float(sum(np.real(np.sum(np.conj(a) * b)) for a, b in zip(mhd_leaves(mu), mhd_leaves(d))))
