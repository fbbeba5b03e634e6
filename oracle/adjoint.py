"""CPU FP64 reverse-mode (adjoint) of the oracle step — TEST INFRASTRUCTURE ONLY.

The reference differentiates its step with ``jax.value_and_grad`` over
``lax.scan(jax.checkpoint(step))`` (``production/autograd_cost_benchmark.py:228-330``,
``src/jaxfun/integrators/cnab2.py:55``).  jax is not available here, so this
module states the same derivative by hand: the vector-Jacobian product of every
stage of ``oracle.kmm.PCFOracle.step``, composed in reverse.  It is the checker
for the HIP adjoint, and it is itself pinned by central finite differences of
the oracle forward along seeded directions (the reference's own check,
``tests/couette/test_differentiability_jax.py:81-136``) and by the adjoint
identity <J v, w> = <v, J^T w> stage by stage (``tests/test_adjoint_oracle.py``).

Convention.  Cotangents ``mu`` are gradients for the real inner product
<a, b> = Re sum conj(a) b over every complex leaf (real leaves: sum a b).  A
complex-linear operator A then propagates cotangents by its conjugate
transpose A^H.  jax's gradient of a real function of complex leaves is
``conj(mu)`` (directional derivative Re sum(g d),
``examples/pcf_minimal_seed_jax.py:78-94``); ``jax_convention`` converts.
"""

from __future__ import annotations

from dataclasses import replace

import numpy as np

from .kmm import (
    IMEXRK3_A,
    IMEXRK3_B,
    OracleState,
    PCFOracle,
    cheb_backward,
    cheb_derivative,
    cheb_scalar_product,
    fourier_backward,
    fourier_forward,
)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _zeros_like_state(s: OracleState) -> OracleState:
    z = lambda grp: tuple(np.zeros_like(np.asarray(a)) for a in grp)  # noqa: E731
    return replace(s, u=z(s.u), g=np.zeros_like(s.g), nonlinear_old=z(s.nonlinear_old),
                   nonlinear_older=z(s.nonlinear_older), solution_old=z(s.solution_old),
                   solution_older=z(s.solution_older))


def _add(a, b):
    return tuple(x + y for x, y in zip(a, b))


class OracleAdjoint:
    """Vector-Jacobian products of ``PCFOracle.step`` and of the energy objective."""

    def __init__(self, o: PCFOracle):
        self.o = o
        n0 = o.N0
        eye = lambda n: np.eye(n)  # noqa: E731
        # x-axis matrices of the separable transforms, applied to identity columns
        self.D = cheb_derivative(eye(n0), 1)  # orthogonal coefficients -> derivative coefficients
        self.Xb = {}
        for name, S in (("B", o.SB), ("D", o.SD)):
            for kx in (0, 1):
                a = S.T @ eye(S.shape[0])
                if kx:
                    a = cheb_derivative(a, kx)
                self.Xb[(name, kx)] = cheb_backward(a, n0)  # (n0, nS)
        P = cheb_scalar_product(eye(n0), n0)  # (n0, n0): scalar products of nodal values
        self.FxD = np.linalg.solve(o.MD1, o.SD @ P)  # nodal -> Dirichlet coefficients
        self.R = np.linalg.solve(o.MD1, o.SD @ (o.h[:, None] * (self.D @ o.SB.T)))  # (nd, nb)
        self._S = {}

    # ------------------------------------------------------------ 1-D adjoints
    @staticmethod
    def _fb_adj(mu, axis, N):
        """Adjoint of ``fourier_backward`` (zero-insert + ifft norm='forward')."""
        n = mu.shape[axis]
        return n * fourier_forward(mu, axis, N)

    @staticmethod
    def _ff_adj(mu, axis, n):
        """Adjoint of ``fourier_forward`` (fft norm='forward' + truncation)."""
        return fourier_backward(mu, axis, n) / n

    def _x(self, A, c):
        return np.tensordot(A, c, axes=(1, 0))

    def _name(self, S):
        return "B" if S is self.o.SB else "D"

    def tp_backward_adj(self, mu, S, padded=True, kx=0, k1=0, k2=0):
        """Adjoint of ``PCFOracle.tp_backward``."""
        o = self.o
        z = self._x(self.Xb[(self._name(S), kx)].T, mu)
        z = self._fb_adj(z, 2, o.N[2])
        if k2:
            K = o.K2e if k2 % 2 else o.K2
            z = np.conj((1j * K) ** k2) * z
        z = self._fb_adj(z, 1, o.N[1])
        if k1:
            K = o.K1e if k1 % 2 else o.K1
            z = np.conj((1j * K) ** k1) * z
        return z

    def tp_forward_D_adj(self, mu):
        """Adjoint of ``PCFOracle.tp_forward_D``."""
        o = self.o
        z = self._ff_adj(mu, 1, o.pad[1])
        z = self._ff_adj(z, 2, o.pad[2])
        return self._x(self.FxD.T, z)

    # ------------------------------------------------------------ stage adjoints
    def energy(self, s: OracleState) -> float:
        """``perturbation_energy`` (``channelflow_kmm.py:1302-1308``) = diagnostics Epert."""
        return self.o.diagnostics(s)["Epert"]

    def energy_grad(self, s: OracleState):
        """Cotangent of Epert with respect to the velocity leaves u."""
        o = self.o
        w = o._weights()
        out = []
        for ui, S in zip(s.u, (o.SB, o.SD, o.SD)):
            up = o.tp_backward(ui, S, padded=False)
            out.append(self.tp_backward_adj(2.0 * w * up, S, padded=False))
        return tuple(out)

    def nonlinear_rhs_adj(self, mu_n):
        """Adjoint of ``PCFOracle.nonlinear_rhs`` (cotangents of Nu, Ng, Nv, Nw -> H)."""
        o = self.o
        mNu, mNg, mNv, mNw = mu_n
        h = o.h[:, None, None]
        mHu = o.W * h * self._x(o.SB.T, mNu)
        mHg = o.W * h * self._x(o.SD.T, mNg)
        DT = self.D.T
        mHo0 = (o.K1**2 + o.K2**2) * mHu
        mHo1 = np.conj(1j * o.K1e) * self._x(DT, mHu) + np.conj(1j * o.K2e) * mHg
        mHo2 = np.conj(1j * o.K2e) * self._x(DT, mHu) - np.conj(1j * o.K1e) * mHg
        mH = [self._x(o.SD, m) for m in (mHo0, mHo1, mHo2)]
        mH[1] = mH[1].astype(complex)
        mH[2] = mH[2].astype(complex)
        mH[1][:, 0, 0] += -(o.M00.T @ np.real(mNv))
        mH[2][:, 0, 0] += -(o.M00.T @ np.real(mNw))
        return tuple(mH)

    def convection_adj(self, u, mu_h):
        """Adjoint of ``PCFOracle.convection`` at velocity u (cotangent of H -> u)."""
        o = self.o
        SB, SD = o.SB, o.SD
        u0, u1, u2 = u
        up = (o.tp_backward(u0, SB), o.tp_backward(u1, SD), o.tp_backward(u2, SD))
        w0 = o.tp_backward(u2, SD, k1=1) - o.tp_backward(u1, SD, k2=1)
        w1 = o.tp_backward(u0, SB, k2=1) - o.tp_backward(u2, SD, kx=1)
        w2 = o.tp_backward(u1, SD, kx=1) - o.tp_backward(u0, SB, k1=1)
        a = (up[0], up[1] + o.dUb * o.xb, up[2])
        b = (w0, w1, w2 + o.dUb)
        mc = tuple(-self.tp_forward_D_adj(o.mask(m)) for m in mu_h)
        ma = _cross(tuple(np.conj(v) for v in b), mc)
        mb = _cross(mc, tuple(np.conj(v) for v in a))
        A = self.tp_backward_adj
        mu0 = A(ma[0], SB) + A(mb[1], SB, k2=1) - A(mb[2], SB, k1=1)
        mu1 = A(ma[1], SD) - A(mb[0], SD, k2=1) + A(mb[2], SD, kx=1)
        mu2 = A(ma[2], SD) + A(mb[0], SD, k1=1) - A(mb[1], SD, kx=1)
        return (mu0, mu1, mu2)

    def nonlinear_adj(self, u, mu_n):
        return self.convection_adj(u, self.nonlinear_rhs_adj(mu_n))

    def reconstruct_adj(self, mu_u):
        """Adjoint of ``PCFOracle.reconstruct`` -> cotangents of (u0, g, v00, w00)."""
        o = self.o
        m0, m1, m2 = (o.mask(m) for m in mu_u)
        mv = np.real(m1[:, 0, 0]).copy()
        mw = np.real(m2[:, 0, 0]).copy()
        m1 = m1.copy()
        m2 = m2.copy()
        m1[:, 0, 0] = 0
        m2[:, 0, 0] = 0
        a, b = o.KoK2
        mf = np.conj(1j * a) * m1 + np.conj(1j * b) * m2
        mg = np.conj(1j * b) * m1 - np.conj(1j * a) * m2
        mu0 = m0 + self._x(self.R.T, mf)
        return mu0, mg, mv, mw

    def _dense(self, gamma):
        key = round(gamma, 17)
        if key not in self._S:
            o = self.o
            self._S[key] = (o.Mu - gamma * o.Lu, o.Mg - gamma * o.Lg, o.M00 - gamma * o.L00)
        return self._S[key]

    def solve_adj(self, gamma, mu):
        """Adjoint of ``PCFOracle._solve`` with the factors of (M - gamma L)."""
        o = self.o
        Su, Sg, S00 = self._dense(gamma)
        mu0, mg, mv, mw = mu

        def tsolve(S, c):
            m = o._modes(c)
            St = np.swapaxes(S, 1, 2)
            re = np.linalg.solve(St, m.real[:, :, None])[:, :, 0]
            im = np.linalg.solve(St, m.imag[:, :, None])[:, :, 0]
            return o._unmodes(re + 1j * im)

        ru = o.mask(tsolve(Su, mu0))
        rg = o.mask(tsolve(Sg, o.mask(mg)))
        rv = np.linalg.solve(S00.T, np.real(mv))
        rw = np.linalg.solve(S00.T, np.real(mw))
        return ru, rg, rv, rw

    def _terms_adj(self, c, mats):
        out = 0
        for A, sk in mats:
            out = out + np.conj(sk) * self._x(A.T, c)
        return out

    def mass_adj(self, r):
        o = self.o
        k2 = o.K1**2 + o.K2**2
        mu = self._terms_adj(r[0], [(o.B_C2, o.W), (o.B_C0, -o.W * k2)])
        mg = self._terms_adj(r[1], [(o.D_C0, o.W)])
        return (mu, mg, o.M00.T @ r[2], o.M00.T @ r[3])

    def lin_adj(self, r):
        o = self.o
        k2 = o.K1**2 + o.K2**2
        nu = o.nu * o.W
        lu = self._terms_adj(r[0], [(o.B_C4, nu), (o.B_C2, -2 * nu * k2), (o.B_C0, nu * k2**2)])
        lg = self._terms_adj(r[1], [(o.D_C2, nu), (o.D_C0, -nu * k2)])
        return (lu, lg, o.L00.T @ r[2], o.L00.T @ r[3])

    @staticmethod
    def _primary_adj(mu_u, mu_g, mp):
        """Add the cotangent of primary(s) = (u0, g, Re u1[:,0,0], Re u2[:,0,0])."""
        u0, u1, u2 = (np.array(a, dtype=complex) for a in mu_u)
        u0 = u0 + mp[0]
        g = mu_g + mp[1]
        u1[:, 0, 0] += np.real(mp[2])
        u2[:, 0, 0] += np.real(mp[3])
        return (u0, u1, u2), g

    # ------------------------------------------------------------ step VJPs
    def step_vjp(self, s: OracleState, mu: OracleState) -> OracleState:
        """Cotangent of the input state of ``PCFOracle.step`` from that of its output."""
        o = self.o
        dt = o.dt
        if s.history_steps < 2:
            return self._imexrk3_vjp(s, mu)
        r = self.solve_adj(dt * 6.0 / 11.0, _add_tuple(self.reconstruct_adj(mu.u), (0, mu.g, 0, 0)))
        mr = self.mass_adj(r)
        mcur_p = _add(tuple(18.0 / 11.0 * x for x in mr), mu.solution_old)
        msol_old = _add(tuple(-9.0 / 11.0 * x for x in mr), mu.solution_older)
        msol_older = tuple(2.0 / 11.0 * x for x in mr)
        mcur_n = _add(tuple(dt * 18.0 / 11.0 * x for x in r), mu.nonlinear_old)
        mnl_old = _add(tuple(-dt * 18.0 / 11.0 * x for x in r), mu.nonlinear_older)
        mnl_older = tuple(dt * 6.0 / 11.0 * x for x in r)
        mu_u = self.nonlinear_adj(s.u, mcur_n)
        mu_u, mu_g = self._primary_adj(mu_u, np.zeros_like(s.g), mcur_p)
        return replace(s, u=mu_u, g=mu_g, nonlinear_old=mnl_old, nonlinear_older=mnl_older,
                       solution_old=msol_old, solution_older=msol_older)

    def _imexrk3_vjp(self, s: OracleState, mu: OracleState) -> OracleState:
        o = self.o
        dt = o.dt
        stages = self._imexrk3_stage_inputs(s)
        mu_u, mu_g = mu.u, mu.g
        rs = [None, None, None]
        for rk in (2, 1, 0):
            a, b = IMEXRK3_A[rk], IMEXRK3_B[rk]
            gamma = (a + b) * dt / 2.0
            r = self.solve_adj(gamma, _add_tuple(self.reconstruct_adj(mu_u), (0, mu_g, 0, 0)))
            rs[rk] = r
            mr = self.mass_adj(r)
            lr = self.lin_adj(r)
            lin = tuple(x + gamma * y for x, y in zip(mr, lr))
            mcur = tuple(dt * a * x for x in r)
            if rk < 2:
                mcur = _add(mcur, tuple(dt * IMEXRK3_B[rk + 1] * x for x in rs[rk + 1]))
            if rk == 0:
                mcur = _add(mcur, mu.nonlinear_old)
            u_in, _ = stages[rk]
            mu_in = self.nonlinear_adj(u_in, mcur)
            mu_in, mg_in = self._primary_adj(mu_in, np.zeros_like(s.g), lin)
            mu_u, mu_g = mu_in, mg_in
        mu_u, mu_g = self._primary_adj(mu_u, mu_g, mu.solution_old)
        zeros = tuple(np.zeros_like(np.asarray(x)) for x in s.nonlinear_old)
        return replace(s, u=mu_u, g=mu_g, nonlinear_old=mu.nonlinear_older,
                       nonlinear_older=zeros, solution_old=mu.solution_older,
                       solution_older=tuple(np.zeros_like(np.asarray(x)) for x in s.solution_old))

    def _imexrk3_stage_inputs(self, s: OracleState):
        """(u, g) entering each IMEXRK3 stage (``PCFOracle.step_imexrk3``)."""
        o = self.o
        dt = o.dt
        out = []
        u_stage, g_stage = s.u, s.g
        prev = (np.zeros_like(s.u[0]), np.zeros_like(s.g), np.zeros(o.ND), np.zeros(o.ND))
        for rk in range(3):
            out.append((u_stage, g_stage))
            cur = o.nonlinear_rhs(o.convection(OracleState(u=u_stage, g=g_stage)))
            a, b = IMEXRK3_A[rk], IMEXRK3_B[rk]
            gamma = (a + b) * dt / 2.0
            v00 = np.real(u_stage[1][:, 0, 0])
            w00 = np.real(u_stage[2][:, 0, 0])
            rhs_u = o.Mu_apply(u_stage[0]) + gamma * o.Lu_apply(u_stage[0]) + dt * (a * cur[0] + b * prev[0])
            rhs_g = o.Mg_apply(g_stage) + gamma * o.Lg_apply(g_stage) + dt * (a * cur[1] + b * prev[1])
            rhs_v = o.M00 @ v00 + gamma * (o.L00 @ v00) + dt * (a * cur[2] + b * prev[2])
            rhs_w = o.M00 @ w00 + gamma * (o.L00 @ w00) + dt * (a * cur[3] + b * prev[3])
            u0n, gn, vn, wn = o._solve(o.startup[rk], rhs_u, rhs_g, rhs_v, rhs_w)
            u_stage = o.reconstruct(u0n, gn, vn, wn)
            g_stage = gn
            prev = cur
        return out

    # ------------------------------------------------------------ objective
    def vjp(self, s0: OracleState, steps: int, mu_final: OracleState) -> OracleState:
        """Pull the cotangent of the state after ``steps`` steps back to ``s0``."""
        o = self.o
        traj = [s0]
        for _ in range(int(steps) - 1):
            traj.append(o.step(traj[-1]))
        mu = mu_final
        for k in range(int(steps) - 1, -1, -1):
            mu = self.step_vjp(traj[k], mu)
        return mu

    def value_and_grad(self, s0: OracleState, steps: int):
        """Epert after ``steps`` steps and its cotangent with respect to every leaf of s0."""
        o = self.o
        traj = [s0]
        for _ in range(int(steps)):
            traj.append(o.step(traj[-1]))
        value = self.energy(traj[-1])
        mu = _zeros_like_state(traj[-1])
        mu = replace(mu, u=self.energy_grad(traj[-1]))
        for k in range(int(steps) - 1, -1, -1):
            mu = self.step_vjp(traj[k], mu)
        return value, mu


def from_jax_convention(g: OracleState) -> OracleState:
    """Inverse of ``jax_convention`` (conj is an involution)."""
    return jax_convention(g)


def zeros_like_state(s: OracleState) -> OracleState:
    return _zeros_like_state(s)


def _add_tuple(a, b):
    return tuple(x + y for x, y in zip(a, b))


def jax_convention(mu: OracleState) -> OracleState:
    """conj of every complex cotangent: the gradient ``jax.grad`` would return."""
    c = lambda grp: tuple(np.conj(x) if np.iscomplexobj(x) else x for x in grp)  # noqa: E731
    return replace(mu, u=c(mu.u), g=np.conj(mu.g), nonlinear_old=c(mu.nonlinear_old),
                   nonlinear_older=c(mu.nonlinear_older), solution_old=c(mu.solution_old),
                   solution_older=c(mu.solution_older))


def pairing(mu: OracleState, d: OracleState) -> float:
    """<mu, d> = Re sum conj(mu) d over every leaf: the directional derivative along d."""
    tot = 0.0
    for a, b in zip(_leaves(mu), _leaves(d)):
        tot += float(np.real(np.sum(np.conj(a) * b)))
    return tot


def _leaves(s: OracleState):
    out = list(s.u) + [s.g]
    for grp in (s.nonlinear_old, s.nonlinear_older, s.solution_old, s.solution_older):
        out.extend(np.asarray(x) for x in grp)
    return out
