"""CPU FP64 oracle of the conducting-wall plane-Couette MHD step.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  A numpy restatement,
in the reference's full-complex layout, of ``PlaneCouetteMHDJax``
(``examples/pcf_mhd_jax.py:69-789``) for SBDF3 with the IMEXRK3 start-up, the
rotational flow nonlinearity and the optimized coefficient path:

* the magnetic vector potential A = (Ax, Ay, Az), each component in the
  Dirichlet Shen space TD (``:143-155``);
* B = curl A in (TD, TC, TC) and J = curl B in (TC, TD, TD), formed on
  Chebyshev coefficients with Galerkin projections onto TD
  (``:269-311``; ``tensorproductspace.py:706-727``, ``composite.py:293-299``);
* the flow forcing H = forward_TD(-(u_tot x w_tot) - J x B) and the induction
  forcing HA = forward_TD(u_tot x B) on the 3/2-padded grid (``:339-357``);
* per-mode Helmholtz solves S_A = M_A - gamma L_A with L_A the eta-weighted
  Laplacian, eta = U_wall / Rm (``:116-139``);
* the coupled SBDF3 / IMEXRK3 timeline of flow and A (``:488-715``);
* diagnostics Emag and divB_L2 (``:760-789``).

Pins (``tests/test_mhd_oracle.py``): the reference's published FP64
diagnostics and 38-leaf state checksums after 325 steps at 65x64x64 and 50
steps at 97x96x96 (``benchmark-results/mhd-sbdf3-rotational-*.json``);
curl identities against physical derivatives of the backward transforms; the
TD projection against quadrature forward(backward(c)); the hydro oracle
recovered exactly at zero magnetic amplitude; and the A subsystem's linear
decay equal to the hydro g subsystem's at Rm = Re.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np

from .kmm import IMEXRK3_A, IMEXRK3_B, SBDF3_GAMMA, OracleState, PCFOracle, _lu_solve, _nopivot_lu, fourier_backward


@dataclass
class MHDOracleState:
    """Mirror of ``MHDState`` (``examples/pcf_mhd_jax.py:34-66``)."""

    flow: OracleState
    A: tuple
    nonlinear_A_old: tuple | None = None
    nonlinear_A_older: tuple | None = None
    A_old: tuple | None = None
    A_older: tuple | None = None

    def leaves(self) -> list:
        out = self.flow.leaves()
        for group in (self.A, self.nonlinear_A_old, self.nonlinear_A_older, self.A_old, self.A_older):
            out.extend(group)
        return out


class MHDOracle(PCFOracle):
    """``PlaneCouetteMHDJax`` (``examples/pcf_mhd_jax.py:69-114``) for SBDF3/rotational."""

    def __init__(self, N=(17, 16, 16), domain=((-1.0, 1.0), (0.0, 4.0 * math.pi), (0.0, 2.0 * math.pi)),
                 Re: float = 400.0, Rm: float | None = None, U_wall: float = 1.0, dt: float = 0.01,
                 padding_factor=(1.0, 1.5, 1.5), perturbation_amplitude: float = 0.1,
                 magnetic_amplitude: float = 0.05, background_b=(0.0, 0.0, 0.0),
                 time_integrator: str = "SBDF3") -> None:
        if time_integrator not in ("SBDF3", "IMEXRK222", "CNAB2"):
            raise ValueError("the MHD oracle steps SBDF3, IMEXRK222 and CNAB2")
        self.Rm = float(Re if Rm is None else Rm)
        # uniform background field added to B in physical space
        # (``_total_B_physical``, examples/pcf_mhd_mri_shearpy_jax.py:149-151)
        self.background_b = tuple(float(v) for v in background_b)
        self.eta = float(U_wall) / self.Rm
        self.magnetic_amplitude = float(magnetic_amplitude)
        super().__init__(N=N, domain=domain, Re=Re, U_wall=U_wall, dt=dt, padding_factor=padding_factor,
                         perturbation_amplitude=perturbation_amplitude, time_integrator=time_integrator)
        self._build_A_operators()

    # ------------------------------------------------------------ A operators
    def _build_A_operators(self):
        """M_A = (h, a), L_A = (h, eta lap a) on TD^3 (``pcf_mhd_jax.py:116-139``)."""
        k2 = self.k2flat
        self.MA = self._mass_g(k2)
        self.LA = self.W * self.eta * (self.D_C2[None] - k2[:, None, None] * self.D_C0[None])
        self._A_factors = {}

        def fac(gamma):
            key = round(gamma, 17)
            if key not in self._A_factors:
                self._A_factors[key] = _nopivot_lu(self.MA - gamma * self.LA)
            return self._A_factors[key]

        self._A_fac = fac
        self.SA = fac(self.dt * SBDF3_GAMMA)
        self.SA_startup = tuple(fac((IMEXRK3_A[rk] + IMEXRK3_B[rk]) * self.dt / 2.0) for rk in range(3))

    def MA_apply(self, c):
        return self.Mg_apply(c)

    def LA_apply(self, c):
        k2 = self.K1**2 + self.K2**2
        e = self.eta * self.W
        return self._terms(c, [(self.D_C2, e), (self.D_C0, -e * k2)])

    def A_solve(self, fac, rhs):
        """``_A_solve`` (``pcf_mhd_jax.py:172-178``): per-mode solve, Nyquist mask."""
        return tuple(self.mask(self._unmodes(_lu_solve(fac, self._modes(r)))) for r in rhs)

    # hooks of the A subsystem (``pcf_mhd_jax.py:143-178``), overridden for other walls
    def A_mass(self, A):
        """``_A_mass_rhs``: mass rows of each component."""
        return tuple(self.MA_apply(a) for a in A)

    def A_lap(self, A):
        """``_A_eta_lap``: eta-Laplacian rows of each component."""
        return tuple(self.LA_apply(a) for a in A)

    def A_solve_gamma(self, gamma, rhs):
        """Solve (M_A - gamma L_A) A = rhs."""
        return self.A_solve(self._A_fac(gamma), rhs)

    def forward_emf(self, emf):
        """``_A_forward_emf``: TD forward transform of each EMF component."""
        return tuple(self.mask(self.tp_forward_D(e)) for e in emf)

    def flow_forcing(self, n, up):
        """Extra physical flow forcing before the Lorentz force (none here)."""
        return n

    # ------------------------------------------------------------ projections
    def project_D(self, c):
        """TD ``project_orthogonal_coeffs`` along x (``composite.py:293-299``):
        M_D^{-1} S_D diag(h) c."""
        w = self._apply0(self.SD, c * self.h[:, None, None])
        shp = w.shape
        return np.linalg.solve(self.MD1, w.reshape(shp[0], -1)).reshape(shp)

    def B_from_A(self, A):
        """``update_B_from_A`` (``pcf_mhd_jax.py:269-289``): B in (TD, TC, TC)."""
        AO = tuple(self.to_orth(a, self.SD) for a in A)
        bx = self._dcoef(AO[2], 0, 1, 0) - self._dcoef(AO[1], 0, 0, 1)
        by = self._dcoef(AO[0], 0, 0, 1) - self._dcoef(AO[2], 1, 0, 0)
        bz = self._dcoef(AO[1], 1, 0, 0) - self._dcoef(AO[0], 0, 1, 0)
        return (self.mask(self.project_D(bx)), self.mask(by), self.mask(bz))

    def J_from_B(self, B):
        """``update_J_from_B`` (``pcf_mhd_jax.py:291-311``): J in (TC, TD, TD)."""
        BO = (self.to_orth(B[0], self.SD), B[1], B[2])
        jx = self._dcoef(BO[2], 0, 1, 0) - self._dcoef(BO[1], 0, 0, 1)
        jy = self._dcoef(BO[0], 0, 0, 1) - self._dcoef(BO[2], 1, 0, 0)
        jz = self._dcoef(BO[1], 1, 0, 0) - self._dcoef(BO[0], 0, 1, 0)
        return (self.mask(jx), self.mask(self.project_D(jy)), self.mask(self.project_D(jz)))

    def backward_B(self, B, padded=True):
        """``_backward_B`` (``pcf_mhd_jax.py:314-333``): spaces (TD, TC, TC)."""
        return (self.tp_backward(B[0], self.SD, padded), self.tp_backward(B[1], None, padded),
                self.tp_backward(B[2], None, padded))

    def backward_J(self, J, padded=True):
        """``_backward_J`` (``pcf_mhd_jax.py:335-337``): spaces (TC, TD, TD)."""
        return (self.tp_backward(J[0], None, padded), self.tp_backward(J[1], self.SD, padded),
                self.tp_backward(J[2], self.SD, padded))

    # ------------------------------------------------------------ state
    def zero_A(self):
        return tuple(np.zeros((self.ND, self.N[1], self.N[2]), dtype=complex) for _ in range(3))

    def reset_mhd_history(self, s: MHDOracleState) -> MHDOracleState:
        """``_ensure_mhd_history`` restart (``pcf_mhd_jax.py:217-248``)."""
        z = self.zero_A()
        return MHDOracleState(flow=self.reset_history(s.flow), A=s.A, nonlinear_A_old=z, nonlinear_A_older=z,
                              A_old=z, A_older=z)

    def initial_state(self) -> MHDOracleState:
        """``initial_state`` (``pcf_mhd_jax.py:253-267``)."""
        flow = super().initial_state()
        x, y, z = self.mesh()
        shape = (self.N0, self.N[1], self.N[2])
        ax = np.zeros(shape)
        amp = self.magnetic_amplitude
        if amp != 0.0:
            ky, kz = 2.0 * np.pi / self.L1, 2.0 * np.pi / self.L2
            ax = np.broadcast_to(amp * (1.0 - x**2) * (1.0 / kz) * np.sin(ky * y) * np.sin(kz * z), shape)
        A = tuple(self.mask(self.tp_forward(v.astype(complex), self.SD, self.MD1)) for v in (ax, np.zeros(shape),
                                                                                              np.zeros(shape)))
        return self.reset_mhd_history(MHDOracleState(flow=flow, A=A))

    # ------------------------------------------------------------ physics
    def mhd_convection(self, s: MHDOracleState):
        """``_mhd_convection`` (``pcf_mhd_jax.py:339-357``): (H, HA)."""
        n, up = self.convection_physical(s.flow)
        n = self.flow_forcing(n, up)
        B = self.B_from_A(s.A)
        J = self.J_from_B(B)
        bp = tuple(b + b0 for b, b0 in zip(self.backward_B(B), self.background_b))
        jp = self.backward_J(J)
        lor = (jp[1] * bp[2] - jp[2] * bp[1], jp[2] * bp[0] - jp[0] * bp[2], jp[0] * bp[1] - jp[1] * bp[0])
        H = tuple(self.mask(self.tp_forward_D(a - b)) for a, b in zip(n, lor))
        ut = (up[0], up[1] + self.dUb * self.xb, up[2])
        emf = (ut[1] * bp[2] - ut[2] * bp[1], ut[2] * bp[0] - ut[0] * bp[2], ut[0] * bp[1] - ut[1] * bp[0])
        return H, self.forward_emf(emf)

    def field_maxima(self, s):
        """Total velocity of the flow and total magnetic field b + B0
        (``_total_B_physical(B, padded=False)``, ``pcf_mhd_mri_shearpy_jax.py:149-151``)
        on the unpadded grid, max |.| per component."""
        u_max, _ = super().field_maxima(s.flow)
        bp = self.backward_B(self.B_from_A(s.A), padded=False)
        return u_max, tuple(float(np.max(np.abs(b + b0))) for b, b0 in zip(bp, self.background_b))

    def cfl(self, s, eta=None, rotation_rate=None, shear_rate=None) -> dict:
        from .kmm import cfl_scalars

        u_max, b_max = self.field_maxima(s)
        return cfl_scalars(u_max, b_max, self.grid_spacings(), self.dt, self.nu, self.eta,
                           getattr(self, "omega", 0.0), getattr(self, "shear_rate", 0.0))

    def nonlinear_A(self, HA):
        """``_A_mass_rhs`` of HA (``pcf_mhd_jax.py:157-159``)."""
        return self.A_mass(HA)

    def step_imexrk3_mhd(self, s: MHDOracleState) -> MHDOracleState:
        """``_step_imexrk3_mhd`` (``pcf_mhd_jax.py:488-561``) with the start-up factors."""
        dt = self.dt
        u_stage, g_stage, A_stage = s.flow.u, s.flow.g, s.A
        prev = (np.zeros_like(s.flow.u[0]), np.zeros_like(s.flow.g), np.zeros(self.ND), np.zeros(self.ND))
        prevA = self.zero_A()
        for rk in range(3):
            H, HA = self.mhd_convection(MHDOracleState(flow=OracleState(u=u_stage, g=g_stage), A=A_stage))
            cur = self.nonlinear_rhs(H)
            curA = self.nonlinear_A(HA)
            a, b = IMEXRK3_A[rk], IMEXRK3_B[rk]
            gamma = (a + b) * dt / 2.0
            v00 = np.real(u_stage[1][:, 0, 0])
            w00 = np.real(u_stage[2][:, 0, 0])
            rhs_u = self.Mu_apply(u_stage[0]) + gamma * self.Lu_apply(u_stage[0]) + dt * (a * cur[0] + b * prev[0])
            rhs_g = self.Mg_apply(g_stage) + gamma * self.Lg_apply(g_stage) + dt * (a * cur[1] + b * prev[1])
            rhs_v = self.M00 @ v00 + gamma * (self.L00 @ v00) + dt * (a * cur[2] + b * prev[2])
            rhs_w = self.M00 @ w00 + gamma * (self.L00 @ w00) + dt * (a * cur[3] + b * prev[3])
            mA, lA = self.A_mass(A_stage), self.A_lap(A_stage)
            rhs_A = [mA[i] + gamma * lA[i] + dt * (a * curA[i] + b * prevA[i]) for i in range(3)]
            u0n, gn, vn, wn = self._solve(self.startup[rk], rhs_u, rhs_g, rhs_v, rhs_w)
            u_stage = self.reconstruct(u0n, gn, vn, wn)
            g_stage = gn
            A_stage = self.A_solve_gamma(gamma, rhs_A)
            prev, prevA = cur, curA
        return replace(s, flow=replace(s.flow, u=u_stage, g=g_stage), A=A_stage)

    def step_ars222_mhd(self, s: MHDOracleState) -> MHDOracleState:
        """``_step_imex_mhd`` (``pcf_mhd_jax.py:389-486``) with the IMEXRK222 tableau:
        flow and A advance by the same two ARS stages, explicit linear terms
        L x_1 and eta-Laplacian(A_1) in the second; every history field passes
        through (``replace`` at :482-486)."""
        from .kmm import ARS222_D, ARS222_G

        dt = self.dt
        a = ((0.0, 0.0, 0.0), (0.0, ARS222_G, 0.0), (0.0, 1.0 - ARS222_G, ARS222_G))
        b = ((0.0, 0.0, 0.0), (ARS222_G, 0.0, 0.0), (ARS222_D, 1.0 - ARS222_D, 0.0))
        base = self.mass_rows(self.primary(s.flow))
        baseA = self.A_mass(s.A)
        stage = MHDOracleState(flow=OracleState(u=s.flow.u, g=s.flow.g), A=s.A)
        nl, nlA, lin, linA = [], [], [], []
        for rk in range(2):
            H, HA = self.mhd_convection(stage)
            nl.append(self.nonlinear_rhs(H))
            nlA.append(self.nonlinear_A(HA))
            if rk > 0:
                lin.append(self.linear_rows(self.primary(stage.flow)))
                linA.append(self.A_lap(stage.A))
            rhs, rhsA = list(base), list(baseA)
            for j in range(rk + 1):
                rhs = [r + dt * b[rk + 1][j] * n for r, n in zip(rhs, nl[j])]
                rhsA = [r + dt * b[rk + 1][j] * n for r, n in zip(rhsA, nlA[j])]
            for j in range(rk):
                rhs = [r + dt * a[rk + 1][j + 1] * ell for r, ell in zip(rhs, lin[j])]
                rhsA = [r + dt * a[rk + 1][j + 1] * ell for r, ell in zip(rhsA, linA[j])]
            u0n, gn, vn, wn = self._solve(self.ars, *rhs)
            stage = MHDOracleState(flow=OracleState(u=self.reconstruct(u0n, gn, vn, wn), g=gn),
                                   A=self.A_solve_gamma(dt * ARS222_G, rhsA))
        return replace(s, flow=replace(s.flow, u=stage.flow.u, g=stage.flow.g), A=stage.A)

    def step_cnab2_mhd(self, s: MHDOracleState) -> MHDOracleState:
        """``_step_cnab2_mhd`` (``pcf_mhd_jax.py:359-387``): the flow CNAB2 update
        and the same Crank-Nicolson / variable AB2 update of A (its nonlinear
        level extrapolated with the flow's ``have_old`` and ``previous_dt``);
        the new state keeps the step's nonlinear rows of flow and A only."""
        from .kmm import variable_ab2_extrapolate

        dt = self.dt
        f = s.flow
        H, HA = self.mhd_convection(s)
        cur, curA = self.nonlinear_rhs(H), self.nonlinear_A(HA)
        oldA = s.nonlinear_A_old if s.nonlinear_A_old is not None else self.zero_A()
        ext = variable_ab2_extrapolate(cur, f.nonlinear_old, f.have_old, dt, f.previous_dt)
        extA = variable_ab2_extrapolate(curA, oldA if f.nonlinear_old is not None else None, f.have_old, dt,
                                        f.previous_dt)
        p = self.primary(f)
        rhs = [a + 0.5 * dt * b + dt * e for a, b, e in zip(self.mass_rows(p), self.linear_rows(p), ext)]
        u0n, gn, vn, wn = self._solve(self._fac(0.5 * dt), *rhs)
        mA, lA = self.A_mass(s.A), self.A_lap(s.A)
        rhsA = [mA[i] + 0.5 * dt * lA[i] + dt * extA[i] for i in range(3)]
        flow = OracleState(u=self.reconstruct(u0n, gn, vn, wn), g=gn, nonlinear_old=cur, have_old=1.0,
                           previous_dt=dt)
        return MHDOracleState(flow=flow, A=self.A_solve_gamma(0.5 * dt, rhsA), nonlinear_A_old=curA)

    def step(self, s: MHDOracleState) -> MHDOracleState:
        """``_step_sbdf3_mhd`` / ``_step_sbdf3_steady`` (``pcf_mhd_jax.py:563-715``),
        or the IMEXRK222 / CNAB2 step."""
        if self.time_integrator == "IMEXRK222":
            return self.step_ars222_mhd(s)
        if self.time_integrator == "CNAB2":
            return self.step_cnab2_mhd(s)
        f = s.flow
        cur_p = self.primary(f)
        H, HA = self.mhd_convection(s)
        cur_n = self.nonlinear_rhs(H)
        cur_nA = self.nonlinear_A(HA)
        if f.history_steps < 2:
            adv = self.step_imexrk3_mhd(s)
            hist = min(2.0, f.history_steps + 1.0)
        else:
            m0, m1, m2 = self.mass_rows(cur_p), self.mass_rows(f.solution_old), self.mass_rows(f.solution_older)
            rhs = tuple(
                (18.0 * a - 9.0 * b + 2.0 * c) / 11.0 + self.dt * ((18.0 * n0 - 18.0 * n1 + 6.0 * n2) / 11.0)
                for a, b, c, n0, n1, n2 in zip(m0, m1, m2, cur_n, f.nonlinear_old, f.nonlinear_older))
            a0, a1, a2 = self.A_mass(s.A), self.A_mass(s.A_old), self.A_mass(s.A_older)
            rhs_A = [
                (18.0 * a0[i] - 9.0 * a1[i] + 2.0 * a2[i]) / 11.0
                + self.dt * ((18.0 * cur_nA[i] - 18.0 * s.nonlinear_A_old[i] + 6.0 * s.nonlinear_A_older[i]) / 11.0)
                for i in range(3)]
            u0n, gn, vn, wn = self._solve(self.steady, *rhs)
            adv = replace(s, flow=replace(f, u=self.reconstruct(u0n, gn, vn, wn), g=gn),
                          A=self.A_solve_gamma(self.dt * SBDF3_GAMMA, rhs_A))
            hist = 2.0
        flow = replace(adv.flow, nonlinear_old=cur_n, nonlinear_older=f.nonlinear_old, solution_old=cur_p,
                       solution_older=f.solution_old, history_steps=hist, have_old=1.0, previous_dt=self.dt)
        return MHDOracleState(flow=flow, A=adv.A, nonlinear_A_old=cur_nA, nonlinear_A_older=s.nonlinear_A_old,
                              A_old=s.A, A_older=s.A_old)

    def solve(self, s: MHDOracleState, steps: int) -> MHDOracleState:
        for _ in range(int(steps)):
            s = self.step(s)
        return s

    # ------------------------------------------------------------ diagnostics
    def diagnostics(self, s: MHDOracleState) -> dict:
        """``diagnostics`` + ``magnetic_divergence_l2`` (``pcf_mhd_jax.py:760-789``)."""
        out = super().diagnostics(s.flow)
        B = self.B_from_A(s.A)
        bp = self.backward_B(B, padded=False)
        out["Emag"] = float(sum(np.real(self.integrate(np.conj(b) * b)) for b in bp))
        divb = self.divB(B)
        out["divB_L2"] = float(np.sqrt(np.real(self.integrate(np.conj(divb) * divb))))
        return out

    def divB(self, B):
        """div B on the unpadded grid, B in (TD, TC, TC) (``pcf_mhd_jax.py:760-768``)."""
        return (self.tp_backward(B[0], self.SD, padded=False, kx=1) + self.tp_backward(B[1], None, padded=False, k1=1)
                + self.tp_backward(B[2], None, padded=False, k2=1))


__all__ = ["MHDOracle", "MHDOracleState"]


class RotatingShearMHDOracle(MHDOracle):
    """``PlaneCouetteMRIShearpyJax`` (``examples/pcf_mhd_mri_shearpy_jax.py:30-514``)
    with conducting walls, SBDF3/rotational: base flow U_b = -S x (``:86-92``),
    Coriolis forcing (``:153-160``) and a uniform background field B0 added to B
    (``:149-151``); A in TD^3 as in ``MHDOracle``."""

    def __init__(self, N=(17, 16, 16), domain=((-1.0, 1.0), (0.0, 4.0 * math.pi), (0.0, 2.0 * math.pi)),
                 Re: float = 400.0, Rm: float | None = None, omega: float = 1.0, shear_rate: float = 1.0,
                 background_b=(0.0, 0.0, 0.025), dt: float = 0.01, padding_factor=(1.0, 1.5, 1.5),
                 perturbation_amplitude: float = 0.05, magnetic_amplitude: float = 0.0,
                 time_integrator: str = "SBDF3") -> None:
        self.omega = float(omega)
        self.shear_rate = float(shear_rate)
        super().__init__(N=N, domain=domain, Re=Re, Rm=Rm, U_wall=1.0, dt=dt, padding_factor=padding_factor,
                         perturbation_amplitude=perturbation_amplitude, magnetic_amplitude=magnetic_amplitude,
                         background_b=background_b, time_integrator=time_integrator)
        # U_b = -S x, dU_b/dx = -S (shearpy :89-92); nu, eta keep U_wall = 1
        self.dUb = -self.shear_rate

    def flow_forcing(self, n, up):
        """Coriolis forcing (shearpy :153-160)."""
        return (n[0] - 2.0 * self.omega * up[1], n[1] + 2.0 * self.omega * up[0], n[2])

    def shearpy_seed(self):
        """Physical fields of ``initial_state`` (shearpy :99-147) with the default
        channel-mode velocity seed and the ``ax_yz`` magnetic seed: the flow
        state and the nodal A_x."""
        from .kmm import OracleState

        x, y, z = self.mesh()
        shape = (self.N0, self.N[1], self.N[2])
        wall = 1.0 - x**2
        amp = self.perturbation_amplitude
        ky, kz = 2.0 * np.pi / self.L1, 2.0 * np.pi / self.L2
        u0 = np.zeros(shape)
        u1 = np.zeros(shape)
        u2 = np.zeros(shape)
        if amp != 0.0:
            for hm in (1, 2, 3):
                u0 = u0 + amp * wall * np.cos(hm * kz * z)
                u1 = u1 + amp * wall * np.sin(hm * kz * z)
            u0 = u0 + 0.1 * amp * wall * np.sin(ky * y) * np.cos(kz * z)
            u1 = u1 + 0.1 * amp * wall * np.cos(ky * y) * np.sin(kz * z)
            u2 = u2 + 0.1 * amp * wall * np.sin(2.0 * ky * y) * np.cos(2.0 * kz * z)
        flow = self.state_from_physical(tuple(np.broadcast_to(v, shape).astype(complex) for v in (u0, u1, u2)))
        ax = np.zeros(shape)
        if self.magnetic_amplitude != 0.0:
            ax = np.broadcast_to(self.magnetic_amplitude * wall * (1.0 / kz) * np.sin(ky * y) * np.sin(kz * z), shape)
        return OracleState(u=flow.u, g=flow.g), ax

    def diagnostics(self, s) -> dict:
        """Adds Emag_total, the energy of the total field b + B0 (shearpy :431-513)."""
        out = super().diagnostics(s)
        bp = self.backward_B(self.B_from_A(s.A), padded=False)
        out["Emag_total"] = float(sum(np.real(self.integrate(np.conj(b + b0) * (b + b0)))
                                      for b, b0 in zip(bp, self.background_b)))
        return out

    def initial_state(self):
        """``initial_state`` (shearpy :99-147), A projected onto TD^3."""
        flow, ax = self.shearpy_seed()
        z = self.zero_A()
        A = (self.mask(self.tp_forward(ax.astype(complex), self.SD, self.MD1)), z[1], z[2])
        return self.reset_mhd_history(MHDOracleState(flow=flow, A=A))

    # ------------------------------------------------ extended diagnostics
    # The rotating-shear family's diagnostics (shearpy :178-513): stresses,
    # alpha, mean-field split, kz = 1 channel amplitudes, dissipation and the
    # electromagnetic health set, all on the unpadded grid.
    def b_stencils(self):
        """Stencils of ``b_coeff_spaces``: (TD, TC, TC) for conducting walls."""
        return (self.SD, None, None)

    def a_stencils(self):
        """Stencils of ``a_coeff_spaces``: TD^3 for conducting walls."""
        return (self.SD, self.SD, self.SD)

    def _bt(self, c, S, kx=0, k1=0, k2=0):
        """``backward_primitive`` on the unpadded grid."""
        return self.tp_backward(c, S, padded=False, kx=kx, k1=k1, k2=k2)

    def wall_trace(self, c, S, derivative=0):
        """``_wall_trace`` (``examples/pcf_fluctuations_jax.py:118-129``): basis
        values (or x-derivatives) at x = -1, +1 applied along axis 0, then the
        2-D Fourier backward on the unpadded (y, z) grid; shape (2, N1, N2)."""
        k = np.arange(self.N0, dtype=float)
        if derivative == 0:
            rows = np.stack([(-1.0) ** k, np.ones_like(k)])
        else:  # T_n'(-1) = (-1)^(n+1) n^2, T_n'(1) = n^2
            rows = np.stack([(-1.0) ** (k + 1) * k**2, k**2])
        if S is not None:
            rows = rows @ S.T
        w = np.tensordot(rows, c, axes=(1, 0))
        w = fourier_backward(w, 1, self.N[1])
        return fourier_backward(w, 2, self.N[2])

    def tc_forward(self, u):
        """``TC.forward`` of an unpadded physical field: Chebyshev coefficients
        (``Chebyshev.py:282-301``) and the two Fourier forwards."""
        from .kmm import cheb_forward, fourier_forward

        c = cheb_forward(u, self.N0)
        c = fourier_forward(c, 2, self.N[2])
        return fourier_forward(c, 1, self.N[1])

    def _kz1_rms(self, fields):
        """``_kz1_rms`` (shearpy :203-215)."""
        from .kmm import fejer_weights

        nz = self.N[2]
        w = fejer_weights(self.N0)
        w = w / np.sum(w)
        power = 0.0
        for f in fields:
            profile = np.mean(np.real(f), axis=1)
            coef = np.fft.fft(profile, axis=1)[:, 1] / nz
            power += 2.0 * np.sum(w * np.abs(coef) ** 2)
        return float(np.sqrt(power))

    @staticmethod
    def _cross(a, b):
        """``physical_cross`` (``src/jaxfun/integrators/nonlinear.py:17-29``)."""
        return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])

    def _l2(self, fields):
        return float(np.sqrt(sum(np.real(self.integrate(np.conj(f) * f)) for f in fields)))

    def extended_diagnostics(self, s) -> dict:
        """``PlaneCouetteMRIShearpyJax.diagnostics`` (shearpy :431-513) with
        ``_dissipation_parts`` (:178-201), ``_kz1_rms`` (:203-215),
        ``_a_divergence`` (:243-249), ``magnetic_flux_from_a`` (:251-263),
        ``electric_field_parts`` (:265-276), ``_current_physical`` (:278-294),
        ``_current_wall_values`` / ``_electric_wall_parts`` (:296-353) and
        ``_electromagnetic_diagnostics`` (:355-429)."""
        out = dict(self.diagnostics(s))
        Sb, Sa = self.b_stencils(), self.a_stencils()
        Su = (self.SB, self.SD, self.SD)
        V = 2.0 * self.L1 * self.L2
        lx = 2.0
        bt = self._bt
        up = self.backward_velocity(s.flow)
        B = self.B_from_A(s.A)
        bf = tuple(bt(b, S) for b, S in zip(B, Sb))
        bp = tuple(b + b0 for b, b0 in zip(bf, self.background_b))
        reynolds = float(np.real(self.integrate(np.real(up[0] * np.conj(up[1]))))) / V
        maxwell = -float(np.real(self.integrate(np.real(bp[0] * np.conj(bp[1]))))) / V
        transport = reynolds + maxwell
        emag_total = float(sum(np.real(self.integrate(np.conj(b) * b)) for b in bp))
        dk = sum(np.real(self.integrate(np.conj(g) * g)) for c, S in zip(s.flow.u, Su)
                 for g in (bt(c, S, kx=1), bt(c, S, k1=1), bt(c, S, k2=1)))
        dm = sum(np.real(self.integrate(np.conj(g) * g)) for c, S in zip(B, Sb)
                 for g in (bt(c, S, kx=1), bt(c, S, k1=1), bt(c, S, k2=1)))
        mean_b = tuple(float(np.real(self.integrate(np.real(b)))) / V for b in bp)
        emag_fluct = float(sum(np.real(self.integrate(np.conj(b) * b)) for b in bf))
        mag_energy_mean = V * sum(m * m for m in mean_b)
        kz1_u, kz1_b = self._kz1_rms(up), self._kz1_rms(bf)
        # electric field parts (physical, unpadded)
        ut = (up[0], up[1] + self.dUb * self.xb, up[2])
        J = (bt(B[2], Sb[2], k1=1) - bt(B[1], Sb[1], k2=1), bt(B[0], Sb[0], k2=1) - bt(B[2], Sb[2], kx=1),
             bt(B[1], Sb[1], kx=1) - bt(B[0], Sb[0], k1=1))
        ideal = tuple(-e for e in self._cross(ut, bp))
        resistive = tuple(self.eta * j for j in J)
        electric = tuple(a + b for a, b in zip(ideal, resistive))

        def div_coeffs(fields):
            c = [self.mask(self.tc_forward(f)) for f in fields]
            return bt(c[0], None, kx=1) + bt(c[1], None, k1=1) + bt(c[2], None, k2=1)

        div_i, div_r = div_coeffs(ideal), div_coeffs(resistive)
        div_e = div_i + div_r
        div_a = bt(s.A[0], Sa[0], kx=1) + bt(s.A[1], Sa[1], k1=1) + bt(s.A[2], Sa[2], k2=1)
        wt = self.wall_trace
        iK1, iK2 = 1j * self.K1, 1j * self.K2
        div_a_wall = wt(s.A[0], Sa[0], 1) + wt(iK1 * s.A[1], Sa[1]) + wt(iK2 * s.A[2], Sa[2])
        # wall parts: velocity with the base flow at x = -1, +1, total field, curl B
        xw = np.array([-1.0, 1.0])[:, None, None]
        vw = tuple(wt(c, S) for c, S in zip(s.flow.u, Su))
        vw = (vw[0], vw[1] + self.dUb * xw, vw[2])
        mw = tuple(wt(c, S) + b0 for c, S, b0 in zip(B, Sb, self.background_b))
        cw = (wt(iK1 * B[2], Sb[2]) - wt(iK2 * B[1], Sb[1]), wt(iK2 * B[0], Sb[0]) - wt(B[2], Sb[2], 1),
              wt(B[1], Sb[1], 1) - wt(iK1 * B[0], Sb[0]))
        ideal_w = tuple(-e for e in self._cross(vw, mw))
        res_w = tuple(self.eta * c for c in cw)
        tot_w = tuple(a + b for a, b in zip(ideal_w, res_w))

        def tangential_linf(w):
            return float(max(np.max(np.abs(w[i])) for i in (1, 2)))

        faraday_by = float(np.real(np.mean(tot_w[2][1]) - np.mean(tot_w[2][0]))) / lx
        faraday_bz = -float(np.real(np.mean(tot_w[1][1]) - np.mean(tot_w[1][0]))) / lx
        ay_w, az_w = wt(s.A[1], Sa[1]), wt(s.A[2], Sa[2])
        b0 = self.background_b
        trace = (b0[0], float(np.real(-(np.mean(az_w[1]) - np.mean(az_w[0])) / lx)) + b0[1],
                 float(np.real((np.mean(ay_w[1]) - np.mean(ay_w[0])) / lx)) + b0[2])
        sq = lambda f: float(np.sqrt(np.real(self.integrate(np.conj(f) * f))))  # noqa: E731
        out.update({
            "Emag_total": emag_total,
            "bmax": float(max(np.max(np.abs(b)) for b in bf)),
            "bmax_total": float(max(np.max(np.abs(b)) for b in bp)),
            "reynolds_xy": reynolds, "maxwell_xy": maxwell, "reynolds_stress": reynolds,
            "maxwell_stress": maxwell, "transport_xy": transport, "total_stress": transport,
            "mean_bx": mean_b[0], "mean_by": mean_b[1], "mean_bz": mean_b[2],
            "mag_energy_fluct_total": emag_fluct, "mag_energy_mean": mag_energy_mean,
            "mag_energy_fluct": emag_total - mag_energy_mean,
            "dissipation_kinetic": self.nu * float(dk), "dissipation_magnetic": self.eta * float(dm),
            "channel_kz1_velocity_rms": kz1_u, "channel_kz1_magnetic_rms": kz1_b,
            "channel_kz1_total_rms": math.sqrt(kz1_u**2 + kz1_b**2),
            "electric_ideal_l2": self._l2(ideal), "electric_resistive_l2": self._l2(resistive),
            "electric_total_l2": self._l2(electric),
            "divergence_e_l2": sq(div_e), "divergence_e_ideal_l2": sq(div_i),
            "divergence_e_resistive_l2": sq(div_r), "divergence_a_l2": sq(div_a),
            "divergence_a_wall_linf": float(np.max(np.abs(div_a_wall))),
            "electric_ideal_wall_tangential_linf": tangential_linf(ideal_w),
            "electric_resistive_wall_tangential_linf": tangential_linf(res_w),
            "electric_wall_tangential_linf": tangential_linf(tot_w),
            "faraday_mean_bx_tendency": 0.0, "faraday_mean_by_tendency": faraday_by,
            "faraday_mean_bz_tendency": faraday_bz,
            "mean_bx_trace": trace[0], "mean_by_trace": trace[1], "mean_bz_trace": trace[2],
            "mean_b_trace_mismatch_linf": float(max(abs(mean_b[i] - trace[i]) for i in range(3))),
        })
        for label, fields in (("ideal", ideal), ("resistive", resistive), ("total", electric)):
            for axis, f in zip("xyz", fields):
                out[f"mean_e{axis}_{label}"] = float(np.real(self.integrate(f))) / V
        out["q_shear"] = self.shear_rate / self.omega if self.omega != 0 else math.inf
        out["kappa2"] = 2.0 * self.omega * (2.0 * self.omega - self.shear_rate)
        sh2 = (self.shear_rate * 1.0) ** 2
        out["alpha_Sh"] = transport / sh2 if sh2 > 0.0 else float("nan")
        if b0[2] ** 2 > 0.0:
            out["alpha"] = out["alpha_B0"] = transport / b0[2] ** 2
        return out
