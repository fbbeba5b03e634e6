"""CPU FP64 restatement of the axisymmetric Taylor-Couette DNS step — TEST INFRASTRUCTURE ONLY.

Nothing in the product imports this module: it is the checker of the HIP
Taylor-Couette path (tests, ``__graft_entry__.smoke`` and the CPU baseline of
``bench.py`` only).

It restates ``AxisymmetricTCDNSJax`` (``examples/taylor_couette_dns_jax.py:217-1078``):
the perturbation equations about the circular-Couette flow V(r) = a r + b / r
in primitive variables (u_r, u_theta, u_z, p) on Fourier(z) x Chebyshev(r),
Dirichlet velocity modes phi_k = T_k - T_{k+2} and the truncated orthogonal
pressure space T_0..T_{N-3} (the P_N / P_{N-2} pair), with per-axial-mode
coupled saddle-point blocks solved by pivoted LU after pinning the k = 0
pressure constant; the SBDF3 step with the ARS(4,4,3) start-up
(``:885-946``), IMEXRK3 alone (``:797-850``) and CNAB2 (``:758-795``).  The
linear eigenproblem that seeds the golden run restates
``TaylorCouetteLinearJax`` (``examples/taylor_couette_linear_jax.py:92-240``).

Weak forms.  Every bilinear form (v, c(r) d^m u / dr^m) is the N-point
Gauss-Chebyshev quadrature of the mapped integral, as jaxfun assembles a form
with a non-polynomial coefficient at the test space's quadrature points
(``src/jaxfun/galerkin/inner.py:306-349``; polynomial products of these
degrees are exact at N points, so the same rule covers them):
    (1/df) sum_q (pi/N) v(x_q) c(r_q) df^m u^(m)(x_q),  df = 2 / (R2 - R1),
times the axial factor Lz (i kz)^m of the Fourier part.

Pinned by the published golden run ``taylor_couette_hydro_dns_v1``
(``production/goldens/taylor_couette_hydro_dns_v1``: linear growth rate, the
seeded and final kinetic energy after 400 CNAB2 steps, the divergence) and
by the checks of ``tests/test_tc_oracle.py``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np
import scipy.linalg as sla
from numpy.polynomial import chebyshev as npcheb

from .kmm import (
    cheb_backward,
    cheb_derivative,
    cheb_forward,
    cheb_nodes,
    cheb_scalar_product,
    fejer_weights,
    fourier_backward,
    fourier_forward,
    fourier_wavenumbers,
    stencil_dirichlet,
)

# ARS(4,4,3) tableau (src/jaxfun/integrators/imex_rk.py:183-198)
ARS443_A = np.array([
    [0.0, 0.0, 0.0, 0.0, 0.0],
    [0.0, 1.0 / 2.0, 0.0, 0.0, 0.0],
    [0.0, 1.0 / 6.0, 1.0 / 2.0, 0.0, 0.0],
    [0.0, -1.0 / 2.0, 1.0 / 2.0, 1.0 / 2.0, 0.0],
    [0.0, 3.0 / 2.0, -3.0 / 2.0, 1.0 / 2.0, 1.0 / 2.0],
])
ARS443_B = np.array([
    [0.0, 0.0, 0.0, 0.0, 0.0],
    [1.0 / 2.0, 0.0, 0.0, 0.0, 0.0],
    [11.0 / 18.0, 1.0 / 18.0, 0.0, 0.0, 0.0],
    [5.0 / 6.0, -5.0 / 6.0, 1.0 / 2.0, 0.0, 0.0],
    [1.0 / 4.0, 7.0 / 4.0, 3.0 / 4.0, -7.0 / 4.0, 0.0],
])
SBDF3_SCALE = 6.0 / 11.0  # src/jaxfun/integrators/sbdf3.py:11-12


class CircularCouette:
    """Base flow V = a r + b / r between cylinders R1 < R2 rotating at Omega1,
    Omega2 (``examples/taylor_couette_linear_jax.py:31-76``)."""

    def __init__(self, R1=1.0, R2=2.0, Omega1=1.0, Omega2=0.0):
        if not (R2 > R1 > 0):
            raise ValueError("require 0 < R1 < R2")
        self.R1, self.R2 = float(R1), float(R2)
        self.Omega1, self.Omega2 = float(Omega1), float(Omega2)
        d2 = self.R2**2 - self.R1**2
        self.a = (self.Omega2 * self.R2**2 - self.Omega1 * self.R1**2) / d2
        self.b = (self.Omega1 - self.Omega2) * self.R1**2 * self.R2**2 / d2
        self.gap = self.R2 - self.R1


class RadialForms:
    """Galerkin radial blocks on N Gauss-Chebyshev points of [R1, R2]."""

    def __init__(self, N: int, R1: float, R2: float):
        self.N = N
        self.df = 2.0 / (R2 - R1)
        self.x = cheb_nodes(N)
        self.r = R1 + (self.x + 1.0) / self.df
        self.V = npcheb.chebvander(self.x, N - 1)  # T_n(x_q)
        self.SD = stencil_dirichlet(N)  # (N-2, N)
        self.SP = np.eye(N)[: N - 2]  # T_0..T_{N-3}

    def values(self, S, deriv=0):
        """(q, j) values of the basis S (rows of orthogonal coefficients) and its
        r-derivatives at the nodes."""
        C = S.T.copy()
        if deriv:
            C = cheb_derivative(C, deriv) * self.df**deriv
        return self.V @ C

    def form(self, test, trial, deriv=0, coef=None):
        c = np.ones(self.N) if coef is None else np.asarray(coef(self.r), dtype=float)
        vt = self.values(test)
        vu = self.values(trial, deriv)
        return (math.pi / self.N) / self.df * (vt * c[:, None]).T @ vu


@dataclass
class TCState:
    """``AxisymmetricTCState`` (``examples/taylor_couette_dns_jax.py:114-165``): full
    complex coefficients, u components (Nz, Nr-2), p (Nz, Nr-2)."""

    u: tuple
    p: np.ndarray
    nonlinear_old: tuple
    have_old: float = 0.0
    nonlinear_older: tuple | None = None
    solution_old: tuple | None = None
    solution_older: tuple | None = None
    history_steps: float = 0.0

    def leaves(self) -> list:
        out = list(self.u) + [self.p] + list(self.nonlinear_old)
        for grp in (self.nonlinear_older, self.solution_old, self.solution_older):
            if grp is not None:
                out += list(grp)
        return out


class TCOracle:
    """``AxisymmetricTCDNSJax`` restated (``examples/taylor_couette_dns_jax.py:217-1078``)."""

    def __init__(self, base: CircularCouette, nu=1.0e-2, Nr=32, Nz=16, Lz=None, dt=2.0e-3, dealias=1.5,
                 time_integrator="CNAB2"):
        self.base = base
        self.nu = float(nu)
        self.Nr, self.Nz = int(Nr), int(Nz)
        self.dt = float(dt)
        ti = str(time_integrator).upper()
        if ti not in ("CNAB2", "SBDF3", "IMEXRK3"):
            raise ValueError("time_integrator must be one of {'CNAB2', 'SBDF3', 'IMEXRK3'}")
        self.time_integrator = ti
        self.dealias = float(dealias)
        self.Lz = float(Lz) if Lz is not None else 2.0 * math.pi / 3.13 * base.gap
        self.rf = RadialForms(self.Nr, base.R1, base.R2)
        self.Nd = self.Nr - 2
        self.Np = self.Nr - 2
        self.kint = fourier_wavenumbers(self.Nz)
        self.kz = self.kint * (2.0 * math.pi / self.Lz)
        self.kz_odd = fourier_wavenumbers(self.Nz, eliminate_highest=True) * (2.0 * math.pi / self.Lz)
        if self.dealias > 1.0:
            self.Mz = int(self.Nz * self.dealias)
            self.Mr = int(self.Nr * self.dealias)
        else:
            self.Mz, self.Mr = self.Nz, self.Nr
        self.r_pad = base.R1 + (cheb_nodes(self.Mr) + 1.0) / self.rf.df
        self._radial_blocks()
        self._build_time_operators()

    # ------------------------------------------------------------ operators
    def _radial_blocks(self):
        rf, b = self.rf, self.base
        SD, SP = rf.SD, rf.SP
        self.MD = rf.form(SD, SD)
        self.D2 = rf.form(SD, SD, 2)
        self.D1r = rf.form(SD, SD, 1, lambda r: 1.0 / r)
        self.R2D = rf.form(SD, SD, 0, lambda r: 1.0 / r**2)
        self.TwoOm = rf.form(SD, SD, 0, lambda r: 2.0 * (b.a + b.b / r**2))
        self.Gr = rf.form(SD, SP, 1)
        self.Gz0 = rf.form(SD, SP, 0)
        self.Dr = rf.form(SP, SD, 1) + rf.form(SP, SD, 0, lambda r: 1.0 / r)
        self.Dz0 = rf.form(SP, SD, 0)

    def _lap(self, kz):
        return self.D2 + self.D1r - kz**2 * self.MD

    def _build_operators(self, ls):
        """Per-mode Limp (4n x 4n), Lexp and mass (3n x 3n), each times Lz
        (``examples/taylor_couette_dns_jax.py:523-606``)."""
        n, nu, dt, a = self.Nd, self.nu, self.dt, self.base.a
        Lz = self.Lz
        nm = self.Nz
        limp = np.zeros((nm, 4 * n, 4 * n), dtype=complex)
        lexp = np.zeros((nm, 3 * n, 3 * n), dtype=complex)
        mass = np.zeros((nm, 3 * n, 3 * n), dtype=complex)
        S = lambda i: slice(i * n, (i + 1) * n)  # noqa: E731
        for m in range(nm):
            kz = self.kz[m]
            lap = self._lap(kz)
            A = limp[m]
            for i in range(3):
                A[S(i), S(i)] += self.MD / dt
            A[S(0), S(0)] += -ls * nu * lap + ls * nu * self.R2D
            A[S(0), S(1)] += -ls * self.TwoOm
            A[S(1), S(1)] += -ls * nu * lap + ls * nu * self.R2D
            A[S(1), S(0)] += -ls * (-2.0 * a) * self.MD
            A[S(2), S(2)] += -ls * nu * lap
            A[S(0), S(3)] += self.Gr
            A[S(2), S(3)] += 1j * kz * self.Gz0
            A[S(3), S(0)] += self.Dr
            A[S(3), S(2)] += 1j * kz * self.Dz0
            E = lexp[m]
            for i in range(3):
                E[S(i), S(i)] += self.MD / dt
                mass[m][S(i), S(i)] = self.MD / dt
            E[S(0), S(0)] += ls * nu * lap - ls * nu * self.R2D
            E[S(0), S(1)] += ls * self.TwoOm
            E[S(1), S(1)] += ls * nu * lap - ls * nu * self.R2D
            E[S(1), S(0)] += ls * (-2.0 * a) * self.MD
            E[S(2), S(2)] += ls * nu * lap
        return Lz * limp, Lz * lexp, Lz * mass

    def _pin(self, limp):
        """Pin the k = 0 pressure constant (``:608-612``)."""
        limp = limp.copy()
        row = 3 * self.Nd
        limp[0, row, :] = 0.0
        limp[0, row, row] = 1.0
        return limp

    def _factor(self, limp):
        return [sla.lu_factor(A) for A in self._pin(limp)]

    def _build_time_operators(self):
        """``:303-338``."""
        if self.time_integrator == "CNAB2":
            self.Limp, self.Lexp, self.mass = self._build_operators(0.5)
            self.Limp_lu = self._factor(self.Limp)
            return
        s_limp, s_lexp, mass = self._build_operators(0.5)
        self.mass = mass
        self.linear = (s_lexp - mass) / 0.5
        self.startup_lu = self._factor(s_limp)
        if self.time_integrator == "IMEXRK3":
            self.Limp, self.Lexp, self.Limp_lu = s_limp, s_lexp, self.startup_lu
            return
        self.Limp, self.Lexp, _ = self._build_operators(SBDF3_SCALE)
        self.Limp_lu = self._factor(self.Limp)

    # --------------------------------------------------------------- transforms
    def _to_orth(self, c):
        return c @ self.rf.SD  # (Nz, Nr)

    def _cheb_r(self, c, n):
        """values along r (axis 1) of orthogonal coefficients at n Gauss points."""
        return cheb_backward(c.T, n).T

    def _phys(self, c):
        """u, du/dr, du/dz on the (Mz, Mr) grid (``:709-714``)."""
        orth = self._to_orth(c)
        val = self._cheb_r(fourier_backward(orth, 0, self.Mz), self.Mr)
        dr = self.rf.df * self._cheb_r(fourier_backward(cheb_derivative(orth.T, 1).T, 0, self.Mz), self.Mr)
        dz = self._cheb_r(fourier_backward(1j * self.kz_odd[:, None] * orth, 0, self.Mz), self.Mr)
        return val, dr, dz

    def _dealias_to_standard(self, v):
        """``:703-707``: padded forward to the N coefficients, back on the N grid."""
        if self.dealias <= 1.0:
            return v
        c = fourier_forward(v, 0, self.Nz)
        c = cheb_forward(c.T, self.Nr).T
        return self._cheb_r(fourier_backward(c, 0, self.Nz), self.Nr)

    def scalar_product(self, v):
        """TD.scalar_product of values on the (Nz, Nr) grid, Nyquist masked."""
        c = self.Lz * fourier_forward(v, 0, self.Nz)
        s = cheb_scalar_product(c.T, self.Nr).T / self.rf.df
        out = s @ self.rf.SD.T
        return self.mask(out)

    def mask(self, c):
        c = np.array(c)
        if self.Nz % 2 == 0:
            c[self.Nz // 2] = 0.0
        return c

    def forward_D(self, v):
        """TD.forward of (Nz, Nr) values: Fourier forward, then the Dirichlet
        mass solve of the weighted scalar products."""
        c = fourier_forward(v, 0, self.Nz)
        s = cheb_scalar_product(c.T, self.Nr).T @ self.rf.SD.T
        M = self.rf.SD @ np.diag(np.where(np.arange(self.Nr) == 0, math.pi, math.pi / 2)) @ self.rf.SD.T
        return np.linalg.solve(M, s.T).T

    def backward(self, c):
        return self._cheb_r(fourier_backward(self._to_orth(c), 0, self.Nz), self.Nr)

    # ---------------------------------------------------------------- physics
    def nonlinear(self, u):
        """inner(v_i, N_i) of the cylindrical advection (``:716-739``)."""
        (ur, urr, urz), (ut, utr, utz), (uz, uzr, uzz) = (self._phys(c) for c in u)
        invr = (1.0 / self.r_pad)[None, :]
        n_r = ur * urr + uz * urz - ut * ut * invr
        n_t = ur * utr + uz * utz + ur * ut * invr
        n_z = ur * uzr + uz * uzz
        return tuple(self.scalar_product(self._dealias_to_standard(f)) for f in (n_r, n_t, n_z))

    def _modes(self, u):
        return np.concatenate(list(u), axis=1)  # (Nz, 3n)

    def _unmodes(self, x):
        n = self.Nd
        return tuple(x[:, i * n:(i + 1) * n] for i in range(3))

    def apply(self, mats, u):
        return self._unmodes(np.einsum("kij,kj->ki", mats, self._modes(u)))

    def solve_rhs(self, rhs_u, lu):
        """``:750-756``, ``:614-637``."""
        rhs = np.concatenate([self._modes(rhs_u), np.zeros((self.Nz, self.Np), dtype=complex)], axis=1)
        rhs[0, 3 * self.Nd] = 0.0
        sol = np.stack([sla.lu_solve(lu[m], rhs[m]) for m in range(self.Nz)])
        return self._unmodes(sol[:, :3 * self.Nd]), sol[:, 3 * self.Nd:]

    # ------------------------------------------------------------------ states
    def zero_state(self) -> TCState:
        z = tuple(np.zeros((self.Nz, self.Nd), dtype=complex) for _ in range(3))
        return TCState(u=z, p=np.zeros((self.Nz, self.Np), dtype=complex),
                       nonlinear_old=tuple(np.zeros_like(a) for a in z))

    def mesh(self):
        z = np.arange(self.Nz) * self.Lz / self.Nz
        return z[:, None], self.rf.r[None, :]

    def state_from_physical(self, values) -> TCState:
        u = tuple(self.forward_D(np.asarray(v, dtype=complex)) for v in values)
        return TCState(u=u, p=np.zeros((self.Nz, self.Np), dtype=complex),
                       nonlinear_old=tuple(np.zeros_like(a) for a in u))

    def initial_state(self, amp=1.0e-3, kz_mode=1) -> TCState:
        """Divergence-free streamfunction seed (``:651-667``)."""
        R1, d = self.base.R1, self.base.gap
        Z, R = self.mesh()
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        arg = math.pi * (R - R1) / d
        g = np.sin(arg) ** 2
        gp = (2.0 * math.pi / d) * np.sin(arg) * np.cos(arg)
        ur = amp * (kz / R) * g * np.sin(kz * Z)
        uz = amp * (1.0 / R) * gp * np.cos(kz * Z)
        ut = amp * np.sin(arg) * np.cos(kz * Z) + 0.0 * Z
        return self.state_from_physical((ur, ut, uz))

    def seed_linear_eigenmode(self, kz_mode=1, amp=1.0e-6, which=0):
        """``:669-701``: the real part of the leading linear eigenmode."""
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        lin = TCLinearOracle(self.base, nu=self.nu, N=self.Nr)
        w, vecs = lin.eigs(0, kz, n_return=which + 1)
        vec = positive_pivot_phase(vecs[:, which])
        n = lin.n
        st = self.zero_state()
        comps = []
        kpos, kneg = int(kz_mode) % self.Nz, (-int(kz_mode)) % self.Nz
        for c in range(3):
            block = vec[c * n:(c + 1) * n] * amp
            arr = np.zeros((self.Nz, self.Nd), dtype=complex)
            if kpos == kneg:
                arr[kpos, :n] = np.real(block)
            else:
                arr[kpos, :n] = 0.5 * block
                arr[kneg, :n] = 0.5 * np.conj(block)
            comps.append(arr)
        return replace(st, u=tuple(comps), nonlinear_old=tuple(np.zeros_like(a) for a in comps),
                       have_old=0.0), complex(w[which])

    # ------------------------------------------------------------- integrators
    def step_cnab2(self, s: TCState) -> TCState:
        """``:758-795``."""
        N = self.nonlinear(s.u)
        ex = self.apply(self.Lexp, s.u)
        if s.have_old:
            nab = tuple(1.5 * a - 0.5 * b for a, b in zip(N, s.nonlinear_old))
        else:
            nab = N
        u, p = self.solve_rhs(tuple(e - b for e, b in zip(ex, nab)), self.Limp_lu)
        return replace(s, u=u, p=p, nonlinear_old=N, have_old=1.0)

    def step_imexrk3(self, s: TCState, lu) -> TCState:
        """ARS(4,4,3) stages (``:797-850``, ``src/jaxfun/integrators/coupled.py:13-44``)."""
        base_rhs = self.apply(self.mass, s.u)
        fields, p = s.u, s.p
        nh, lh = [], []
        for rk in range(4):
            nonl = self.nonlinear(fields)
            nh.append(tuple(-v for v in nonl))
            if rk > 0:
                lh.append(self.apply(self.linear, fields))
            rhs = list(base_rhs)
            for j in range(rk + 1):
                rhs = [r + ARS443_B[rk + 1, j] * v for r, v in zip(rhs, nh[j])]
            for j in range(rk):
                rhs = [r + ARS443_A[rk + 1, j + 1] * v for r, v in zip(rhs, lh[j])]
            fields, p = self.solve_rhs(tuple(rhs), lu)
        return replace(s, u=fields, p=p)

    @staticmethod
    def ensure_history(s: TCState) -> TCState:
        """``:851-863``."""
        z = tuple(np.zeros_like(v) for v in s.u)
        return replace(s, nonlinear_older=z if s.nonlinear_older is None else s.nonlinear_older,
                       solution_old=s.u if s.solution_old is None else s.solution_old,
                       solution_older=s.u if s.solution_older is None else s.solution_older)

    @staticmethod
    def _shift(s: TCState, adv: TCState, cur) -> TCState:
        """``:865-883``."""
        return replace(adv, nonlinear_old=cur, nonlinear_older=s.nonlinear_old, solution_old=s.u,
                       solution_older=s.solution_old, history_steps=min(2.0, s.history_steps + 1.0), have_old=1.0)

    def step_sbdf3(self, s: TCState) -> TCState:
        """``:885-946``."""
        s = self.ensure_history(s)
        cur = self.nonlinear(s.u)
        if s.history_steps < 2:
            return self._shift(s, self.step_imexrk3(s, self.startup_lu), cur)
        m0, m1, m2 = (self.apply(self.mass, v) for v in (s.u, s.solution_old, s.solution_older))
        rhs = tuple((18.0 * a - 9.0 * b + 2.0 * c) / 11.0 - (18.0 * x - 18.0 * y + 6.0 * z) / 11.0
                    for a, b, c, x, y, z in zip(m0, m1, m2, cur, s.nonlinear_old, s.nonlinear_older))
        u, p = self.solve_rhs(rhs, self.Limp_lu)
        return self._shift(s, replace(s, u=u, p=p), cur)

    def step(self, s: TCState) -> TCState:
        if self.time_integrator == "CNAB2":
            return self.step_cnab2(s)
        if self.time_integrator == "IMEXRK3":
            return self.step_imexrk3(s, self.Limp_lu)
        return self.step_sbdf3(s)

    def solve(self, s: TCState, steps: int) -> TCState:
        if self.time_integrator == "SBDF3":
            s = self.ensure_history(s)
        for _ in range(int(steps)):
            s = self.step(s)
        return s

    # ------------------------------------------------------------- diagnostics
    def velocity_physical(self, s):
        return tuple(self.backward(c) for c in s.u)

    def _integrate(self, f):
        w = (fejer_weights(self.Nr) / self.rf.df)[None, :] * (self.Lz / self.Nz)
        return float(np.real(np.sum(f * w)))

    def energy(self, s):
        r = self.rf.r[None, :]
        return self._integrate(0.5 * sum(np.real(np.conj(v) * v) for v in self.velocity_physical(s)) * r)

    def divergence(self, s):
        orth_r = self._to_orth(s.u[0])
        dur = self.rf.df * self._cheb_r(fourier_backward(cheb_derivative(orth_r.T, 1).T, 0, self.Nz), self.Nr)
        ur = self.backward(s.u[0])
        duz = self._cheb_r(fourier_backward(1j * self.kz_odd[:, None] * self._to_orth(s.u[2]), 0, self.Nz), self.Nr)
        return dur + ur / self.rf.r[None, :] + duz

    def continuity_l2(self, s):
        """``:1025-1034,1053-1059``."""
        res = (self.Lz * (s.u[0] @ self.Dr.T)
               + self.Lz * (1j * self.kz[:, None]) * (s.u[2] @ self.Dz0.T))
        return float(np.linalg.norm(res))

    def wall(self, s):
        """Coefficient wall trace (``src/jaxfun/diagnostics.py:79-120``)."""
        ends = npcheb.chebvander(np.array([-1.0, 1.0]), self.Nr - 1) @ self.rf.SD.T  # (2, n)
        vals = [np.abs(fourier_backward(c @ ends.T, 0, self.Nz)) for c in s.u]
        return float(max(np.max(v) for v in vals))

    def diagnostics(self, s) -> dict:
        """``:1061-1069``."""
        vel = self.velocity_physical(s)
        r = self.rf.r[None, :]
        return {"E": self.energy(s), "div_linf": float(np.max(np.abs(self.divergence(s)))),
                "continuity_l2": self.continuity_l2(s), "wall": self.wall(s),
                "Eth": self._integrate(np.real(np.conj(vel[1]) * vel[1]) * r)}


def positive_pivot_phase(vec):
    """``examples/taylor_couette_dns_jax.py:99-111``."""
    i = int(np.argmax(np.abs(vec)))
    piv = vec[i]
    pa = float(abs(piv))
    if pa == 0.0:
        return vec
    if piv.real < 0.0 or (abs(piv.real) <= 1.0e-14 * pa and piv.imag < 0.0):
        return -vec
    return vec


class TCLinearOracle:
    """Axisymmetric-or-not linear eigenproblem L q = lambda M q
    (``examples/taylor_couette_linear_jax.py:92-240``; ``src/jaxfun/la/eig.py:145-195``)."""

    def __init__(self, base: CircularCouette, nu=1.0e-3, N=48):
        self.base, self.nu, self.N = base, float(nu), int(N)
        self.rf = RadialForms(self.N, base.R1, base.R2)
        self.n = self.N - 2

    def assemble(self, m, kz):
        rf, b, n = self.rf, self.base, self.n
        SD, SP = rf.SD, rf.SP
        A = lambda deriv, coef=None: rf.form(SD, SD, deriv, coef)  # noqa: E731
        lap = A(2) + A(1, lambda r: 1.0 / r) - A(0, lambda r: m**2 / r**2 + kz**2)
        Mvv = A(0)
        lv_diag = lap - A(0, lambda r: 1.0 / r**2)
        adv = -1j * A(0, lambda r: m * (b.a + b.b / r**2)) if m else np.zeros((n, n), dtype=complex)
        two_om = A(0, lambda r: 2 * b.a + 2 * b.b / r**2)
        shear = A(0, lambda r: 2 * b.a + 0 * r)
        couple = 1j * A(0, lambda r: 2 * m / r**2) if m else np.zeros((n, n), dtype=complex)
        Gr = rf.form(SD, SP, 1)
        Gt = 1j * rf.form(SD, SP, 0, lambda r: m / r) if m else np.zeros((n, n), dtype=complex)
        Gz = 1j * kz * rf.form(SD, SP, 0)
        Dr = rf.form(SP, SD, 1) + rf.form(SP, SD, 0, lambda r: 1.0 / r)
        Dt = 1j * rf.form(SP, SD, 0, lambda r: m / r) if m else np.zeros((n, n), dtype=complex)
        Dz = 1j * kz * rf.form(SP, SD, 0)
        L0 = np.zeros((4 * n, 4 * n), dtype=complex)
        Lv = np.zeros_like(L0)
        M = np.zeros_like(L0)

        def put(X, i, j, v):
            X[i * n:(i + 1) * n, j * n:(j + 1) * n] = v

        put(L0, 0, 0, adv), put(Lv, 0, 0, lv_diag), put(L0, 0, 1, two_om), put(Lv, 0, 1, -couple)
        put(L0, 0, 3, -Gr), put(M, 0, 0, Mvv)
        put(L0, 1, 0, -shear), put(Lv, 1, 0, couple), put(L0, 1, 1, adv), put(Lv, 1, 1, lv_diag)
        put(L0, 1, 3, -Gt), put(M, 1, 1, Mvv)
        put(L0, 2, 2, adv), put(Lv, 2, 2, lap), put(L0, 2, 3, -Gz), put(M, 2, 2, Mvv)
        put(L0, 3, 0, Dr), put(L0, 3, 1, Dt), put(L0, 3, 2, Dz)
        return L0 + self.nu * Lv, M

    def eigs(self, m, kz, n_return=8, finite_cap=1.0e6):
        L, M = self.assemble(m, kz)
        w, V = sla.eig(L, M)
        keep = np.isfinite(w) & (np.abs(w) < finite_cap)
        w, V = w[keep], V[:, keep]
        order = np.lexsort((-w.imag, -w.real))
        return w[order][:n_return], V[:, order][:, :n_return]


class TC3DOracle(TCOracle):
    """``TaylorCouetteDNSJax`` restated (``examples/taylor_couette_dns_jax.py:1081-1356``):
    Fourier(theta) x Fourier(z) x Chebyshev(r), one saddle-point block per
    (m, kz) with the base-flow advection -Omega d/dtheta, the 2/r^2 azimuthal
    couplings and the 3D Laplacian (:1154-1232); every form carries the
    Fourier factor 2 pi Lz (i m)^a (i kz)^b.  Arrays are (Ntheta, Nz, Nr - 2)."""

    def __init__(self, base: CircularCouette, nu=1.0e-2, Nr=24, Ntheta=8, Nz=16, Lz=None, dt=2.0e-3, dealias=1.5,
                 time_integrator="CNAB2"):
        self.Ntheta = int(Ntheta)
        self.mint = fourier_wavenumbers(self.Ntheta)
        self.m_odd = fourier_wavenumbers(self.Ntheta, eliminate_highest=True)
        super().__init__(base, nu=nu, Nr=Nr, Nz=Nz, Lz=Lz, dt=dt, dealias=dealias, time_integrator=time_integrator)
        self.Mth = int(self.Ntheta * self.dealias) if self.dealias > 1.0 else self.Ntheta

    def _radial_blocks(self):
        super()._radial_blocks()
        rf, b = self.rf, self.base
        self.Om = rf.form(rf.SD, rf.SD, 0, lambda r: b.a + b.b / r**2)
        self.Gt0 = rf.form(rf.SD, rf.SP, 0, lambda r: 1.0 / r)
        self.Dt0 = rf.form(rf.SP, rf.SD, 0, lambda r: 1.0 / r)

    def _operator_part(self, m, kz, sign):
        """The _add_avv_terms blocks of one mode times sign (3n x 3n)."""
        n, nu, a = self.Nd, self.nu, self.base.a
        S = lambda i: slice(i * n, (i + 1) * n)  # noqa: E731
        lap = self.D2 + self.D1r - kz**2 * self.MD - m**2 * self.R2D
        X = np.zeros((3 * n, 3 * n), dtype=complex)
        X[S(0), S(0)] = sign * (nu * lap - nu * self.R2D - 1j * m * self.Om)
        X[S(0), S(1)] = sign * (-nu * 2j * m * self.R2D + self.TwoOm)
        X[S(1), S(1)] = sign * (nu * lap - nu * self.R2D - 1j * m * self.Om)
        X[S(1), S(0)] = sign * (nu * 2j * m * self.R2D - 2.0 * a * self.MD)
        X[S(2), S(2)] = sign * (nu * lap - 1j * m * self.Om)
        return X

    def _build_operators(self, ls):
        """``:1192-1232``, per (m, kz) mode, times 2 pi Lz."""
        n, dt = self.Nd, self.dt
        fz = 2.0 * math.pi * self.Lz
        nm = self.Ntheta * self.Nz
        limp = np.zeros((nm, 4 * n, 4 * n), dtype=complex)
        lexp = np.zeros((nm, 3 * n, 3 * n), dtype=complex)
        mass = np.zeros((nm, 3 * n, 3 * n), dtype=complex)
        S = lambda i: slice(i * n, (i + 1) * n)  # noqa: E731
        for it in range(self.Ntheta):
            for iz in range(self.Nz):
                q = it * self.Nz + iz
                m, kz = self.mint[it], self.kz[iz]
                M3 = np.kron(np.eye(3), self.MD / dt)
                limp[q][:3 * n, :3 * n] = M3 + self._operator_part(m, kz, -ls)
                lexp[q] = M3 + self._operator_part(m, kz, ls)
                mass[q] = M3
                A = limp[q]
                A[S(0), S(3)] = self.Gr
                A[S(1), S(3)] = 1j * m * self.Gt0
                A[S(2), S(3)] = 1j * kz * self.Gz0
                A[S(3), S(0)] = self.Dr
                A[S(3), S(1)] = 1j * m * self.Dt0
                A[S(3), S(2)] = 1j * kz * self.Dz0
        return fz * limp, fz * lexp, fz * mass

    # transforms over (theta, z)
    def _orth(self, c):
        return c @ self.rf.SD  # (Nth, Nz, Nr)

    def _cheb_last(self, c, n):
        return np.moveaxis(cheb_backward(np.moveaxis(c, -1, 0), n), 0, -1)

    def _synth(self, c, nth, nz, nr):
        v = fourier_backward(fourier_backward(c, 0, nth), 1, nz)
        return self._cheb_last(v, nr)

    def _phys(self, c):
        orth = self._orth(c)
        d = lambda x: self._synth(x, self.Mth, self.Mz, self.Mr)  # noqa: E731
        dr = np.moveaxis(cheb_derivative(np.moveaxis(orth, -1, 0), 1), 0, -1)
        return (d(orth), self.rf.df * d(dr), d(1j * self.m_odd[:, None, None] * orth),
                d(1j * self.kz_odd[None, :, None] * orth))

    def _dealias_to_standard(self, v):
        if self.dealias <= 1.0:
            return v
        c = fourier_forward(fourier_forward(v, 0, self.Ntheta), 1, self.Nz)
        c = np.moveaxis(cheb_forward(np.moveaxis(c, -1, 0), self.Nr), 0, -1)
        return self._synth(c, self.Ntheta, self.Nz, self.Nr)

    def scalar_product(self, v):
        c = 2.0 * math.pi * self.Lz * fourier_forward(fourier_forward(v, 0, self.Ntheta), 1, self.Nz)
        s = np.moveaxis(cheb_scalar_product(np.moveaxis(c, -1, 0), self.Nr), 0, -1) / self.rf.df
        return self.mask(s @ self.rf.SD.T)

    def mask(self, c):
        c = np.array(c)
        if self.Ntheta % 2 == 0:
            c[self.Ntheta // 2] = 0.0
        if self.Nz % 2 == 0:
            c[:, self.Nz // 2] = 0.0
        return c

    def nonlinear(self, u):
        """``:1242-1262``."""
        (ur, urr, urt, urz), (ut, utr, utt, utz), (uz, uzr, uzt, uzz) = (self._phys(c) for c in u)
        invr = (1.0 / self.r_pad)[None, None, :]
        n_r = ur * urr + (ut * invr) * urt + uz * urz - ut * ut * invr
        n_t = ur * utr + (ut * invr) * utt + uz * utz + ur * ut * invr
        n_z = ur * uzr + (ut * invr) * uzt + uz * uzz
        return tuple(self.scalar_product(self._dealias_to_standard(f)) for f in (n_r, n_t, n_z))

    def _modes(self, u):
        return np.concatenate([c.reshape(self.Ntheta * self.Nz, self.Nd) for c in u], axis=1)

    def _unmodes(self, x):
        n = self.Nd
        return tuple(x[:, i * n:(i + 1) * n].reshape(self.Ntheta, self.Nz, n) for i in range(3))

    def solve_rhs(self, rhs_u, lu):
        nm = self.Ntheta * self.Nz
        rhs = np.concatenate([self._modes(rhs_u), np.zeros((nm, self.Np), dtype=complex)], axis=1)
        rhs[0, 3 * self.Nd] = 0.0
        sol = np.stack([sla.lu_solve(lu[q], rhs[q]) for q in range(nm)])
        return self._unmodes(sol[:, :3 * self.Nd]), sol[:, 3 * self.Nd:].reshape(self.Ntheta, self.Nz, self.Np)

    def zero_state(self) -> TCState:
        sh = (self.Ntheta, self.Nz, self.Nd)
        z = tuple(np.zeros(sh, dtype=complex) for _ in range(3))
        return TCState(u=z, p=np.zeros((self.Ntheta, self.Nz, self.Np), dtype=complex),
                       nonlinear_old=tuple(np.zeros_like(a) for a in z))

    def mesh(self):
        th = np.arange(self.Ntheta) * 2.0 * math.pi / self.Ntheta
        z = np.arange(self.Nz) * self.Lz / self.Nz
        return th[:, None, None], z[None, :, None], self.rf.r[None, None, :]

    def forward_D(self, v):
        c = fourier_forward(fourier_forward(np.asarray(v, dtype=complex), 0, self.Ntheta), 1, self.Nz)
        s = np.moveaxis(cheb_scalar_product(np.moveaxis(c, -1, 0), self.Nr), 0, -1) @ self.rf.SD.T
        M = self.rf.SD @ np.diag(np.where(np.arange(self.Nr) == 0, math.pi, math.pi / 2)) @ self.rf.SD.T
        return np.linalg.solve(M, s.reshape(-1, self.Nd).T).T.reshape(s.shape)

    def state_from_physical(self, values) -> TCState:
        u = tuple(self.forward_D(v) for v in values)
        st = self.zero_state()
        return replace(st, u=u)

    def initial_state(self, amp=1.0e-3, m=1, kz_mode=1) -> TCState:
        """``:1264-1275``."""
        TH, Z, R = self.mesh()
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        shape = np.sin(math.pi * (R - self.base.R1) / self.base.gap)
        f = amp * shape * np.cos(int(m) * TH) * np.cos(kz * Z)
        return self.state_from_physical((f, f, f))

    def seed_linear_eigenmode(self, m=1, kz_mode=1, amp=1.0e-6, which=0):
        """``:1277-1308``: no pivot phase for m != 0."""
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        lin = TCLinearOracle(self.base, nu=self.nu, N=self.Nr)
        w, vecs = lin.eigs(int(m), kz, n_return=which + 1)
        vec = vecs[:, which]
        if int(m) == 0:
            vec = positive_pivot_phase(vec)
        n = lin.n
        mpos, mneg = int(m) % self.Ntheta, (-int(m)) % self.Ntheta
        kpos, kneg = int(kz_mode) % self.Nz, (-int(kz_mode)) % self.Nz
        comps = []
        for c in range(3):
            block = vec[c * n:(c + 1) * n] * amp
            arr = np.zeros((self.Ntheta, self.Nz, self.Nd), dtype=complex)
            if mpos == mneg and kpos == kneg:
                arr[mpos, kpos, :n] = np.real(block)
            else:
                arr[mpos, kpos, :n] = 0.5 * block
                arr[mneg, kneg, :n] = 0.5 * np.conj(block)
            comps.append(arr)
        st = self.zero_state()
        return replace(st, u=tuple(comps), nonlinear_old=tuple(np.zeros_like(a) for a in comps)), complex(w[which])

    def backward(self, c):
        return self._synth(self._orth(c), self.Ntheta, self.Nz, self.Nr)

    def _integrate(self, f):
        w = (fejer_weights(self.Nr) / self.rf.df)[None, None, :] * (self.Lz / self.Nz) * (2.0 * math.pi / self.Ntheta)
        return float(np.real(np.sum(f * w)))

    def energy(self, s):
        r = self.rf.r[None, None, :]
        return self._integrate(0.5 * sum(np.real(np.conj(v) * v) for v in self.velocity_physical(s)) * r)

    def divergence(self, s):
        orth_r = self._orth(s.u[0])
        dr = np.moveaxis(cheb_derivative(np.moveaxis(orth_r, -1, 0), 1), 0, -1)
        syn = lambda x: self._synth(x, self.Ntheta, self.Nz, self.Nr)  # noqa: E731
        r = self.rf.r[None, None, :]
        return (self.rf.df * syn(dr) + syn(orth_r) / r
                + syn(1j * self.m_odd[:, None, None] * self._orth(s.u[1])) / r
                + syn(1j * self.kz_odd[None, :, None] * self._orth(s.u[2])))

    def continuity_l2(self, s):
        """``:1310-1321,1338-1346``."""
        fz = 2.0 * math.pi * self.Lz
        res = (fz * (s.u[0] @ self.Dr.T) + fz * (1j * self.mint[:, None, None]) * (s.u[1] @ self.Dt0.T)
               + fz * (1j * self.kz[None, :, None]) * (s.u[2] @ self.Dz0.T))
        return float(np.linalg.norm(res))

    def wall(self, s):
        ends = npcheb.chebvander(np.array([-1.0, 1.0]), self.Nr - 1) @ self.rf.SD.T
        vals = [np.abs(fourier_backward(fourier_backward(c @ ends.T, 0, self.Ntheta), 1, self.Nz)) for c in s.u]
        return float(max(np.max(v) for v in vals))

    def diagnostics(self, s) -> dict:
        vel = self.velocity_physical(s)
        r = self.rf.r[None, None, :]
        return {"E": self.energy(s), "div_linf": float(np.max(np.abs(self.divergence(s)))),
                "continuity_l2": self.continuity_l2(s), "wall": self.wall(s),
                "Eth": self._integrate(np.real(np.conj(vel[1]) * vel[1]) * r)}


# ---------------------------------------------------------------------------
# Axisymmetric conducting-wall Taylor-Couette MHD (AxisymmetricMRIDNSJax)
# ---------------------------------------------------------------------------


def stencil_two_point(N: int, left, right) -> np.ndarray:
    """psi_i = T_i + d1 T_{i+1} + d2 T_{i+2} meeting one homogeneous condition
    per wall (``src/jaxfun/galerkin/composite.py:797-871``): ``left(j)`` and
    ``right(j)`` give the condition applied to T_j.  Shape (N - 2, N)."""
    S = np.zeros((N - 2, N))
    for i in range(N - 2):
        A = np.array([[left(i + 1), left(i + 2)], [right(i + 1), right(i + 2)]], dtype=float)
        d = np.linalg.solve(A, -np.array([left(i), right(i)], dtype=float))
        S[i, i], S[i, i + 1], S[i, i + 2] = 1.0, d[0], d[1]
    return S


def stencil_robin(N: int, alpha_left: float, alpha_right: float) -> np.ndarray:
    """psi + alpha psi_x = 0 at x = -1 and x = +1 (the "R" condition; T_j(+-1) =
    (+-1)^j, T_j'(+-1) = (+-1)^(j+1) j^2)."""
    return stencil_two_point(N, lambda j: (-1.0) ** j * (1.0 - alpha_left * j * j),
                             lambda j: 1.0 + alpha_right * j * j)


def stencil_neumann(N: int) -> np.ndarray:
    """psi_x = 0 at both walls: T_i - i^2 / (i + 2)^2 T_{i+2}, the closed form
    the reference's symbolic solve simplifies to (``composite.py:811-814``)."""
    S = np.zeros((N - 2, N))
    i = np.arange(N - 2)
    S[i, i] = 1.0
    S[i, i + 2] = -(i.astype(float) ** 2) / (i + 2.0) ** 2
    return S


class TCMRILinearOracle:
    """Conducting-wall MRI eigenproblem L q = lambda M q on (u_r, u_theta, u_z,
    p, b_r, b_theta, b_z) (``examples/taylor_couette_mri_jax.py:65-401``)."""

    def __init__(self, base: CircularCouette, B0=0.1, nu=1e-3, eta_mag=1e-3, N=48):
        self.base, self.B0, self.nu, self.eta = base, float(B0), float(nu), float(eta_mag)
        self.N = int(N)
        self.rf = RadialForms(self.N, base.R1, base.R2)
        self.n = self.N - 2
        df = self.rf.df
        self.SBT = stencil_robin(self.N, base.R1 * df, base.R2 * df)
        self.SBZ = stencil_neumann(self.N)

    def assemble(self, m, kz):
        """``assemble_parts`` (``:214-321``): L = L0 + nu Lnu + eta Leta."""
        rf, b, n = self.rf, self.base, self.n
        SD, SP, SBT, SBZ = rf.SD, rf.SP, self.SBT, self.SBZ
        sp_ = {"ur": SD, "ut": SD, "uz": SD, "p": SP, "br": SD, "bt": SBT, "bz": SBZ}
        idx = {"ur": 0, "ut": 1, "uz": 2, "p": 3, "br": 4, "bt": 5, "bz": 6}
        ikB = 1j * kz * self.B0
        om = lambda r: b.a + b.b / r**2  # noqa: E731

        def blk(row, col, terms):
            out = np.zeros((n, n), dtype=complex)
            for coef, deriv in terms:
                out = out + (rf.form(sp_[row], sp_[col], deriv) if coef is None else
                             rf.form(sp_[row], sp_[col], deriv, lambda r, c=coef: np.real(c(r)))
                             + 1j * rf.form(sp_[row], sp_[col], deriv, lambda r, c=coef: np.imag(c(r))))
            return out

        const = lambda v: (lambda r: v + 0.0 * r)  # noqa: E731
        lap = [(None, 2), (lambda r: 1.0 / r, 1), (lambda r: -(m * m / r**2 + kz * kz), 0)]
        lv = lap + [(lambda r: -1.0 / r**2, 0)]
        couple = [(lambda r: 2j * m / r**2, 0)]
        imOm = lambda r: 1j * m * om(r)  # noqa: E731
        L0 = np.zeros((7 * n, 7 * n), dtype=complex)
        Lnu, Leta, M = np.zeros_like(L0), np.zeros_like(L0), np.zeros_like(L0)

        def put(X, row, col, v):
            i, j = idx[row], idx[col]
            X[i * n:(i + 1) * n, j * n:(j + 1) * n] += v

        neg_imOm = [(lambda r: -imOm(r), 0)]
        put(L0, "ur", "ur", blk("ur", "ur", neg_imOm))
        put(Lnu, "ur", "ur", blk("ur", "ur", lv))
        put(L0, "ur", "ut", blk("ur", "ut", [(lambda r: 2.0 * om(r), 0)]))
        put(Lnu, "ur", "ut", -blk("ur", "ut", couple))
        put(L0, "ur", "p", -blk("ur", "p", [(None, 1)]))
        put(L0, "ur", "br", blk("ur", "br", [(const(ikB), 0)]))
        put(M, "ur", "ur", blk("ur", "ur", [(None, 0)]))
        put(L0, "ut", "ur", blk("ut", "ur", [(const(-2.0 * b.a), 0)]))
        put(Lnu, "ut", "ur", blk("ut", "ur", couple))
        put(L0, "ut", "ut", blk("ut", "ut", neg_imOm))
        put(Lnu, "ut", "ut", blk("ut", "ut", lv))
        put(L0, "ut", "p", -blk("ut", "p", [(lambda r: 1j * m / r, 0)]))
        put(L0, "ut", "bt", blk("ut", "bt", [(const(ikB), 0)]))
        put(M, "ut", "ut", blk("ut", "ut", [(None, 0)]))
        put(L0, "uz", "uz", blk("uz", "uz", neg_imOm))
        put(Lnu, "uz", "uz", blk("uz", "uz", lap))
        put(L0, "uz", "p", -blk("uz", "p", [(const(1j * kz), 0)]))
        put(L0, "uz", "bz", blk("uz", "bz", [(const(ikB), 0)]))
        put(M, "uz", "uz", blk("uz", "uz", [(None, 0)]))
        put(L0, "p", "ur", blk("p", "ur", [(None, 1), (lambda r: 1.0 / r, 0)]))
        put(L0, "p", "ut", blk("p", "ut", [(lambda r: 1j * m / r, 0)]))
        put(L0, "p", "uz", blk("p", "uz", [(const(1j * kz), 0)]))
        put(L0, "br", "ur", blk("br", "ur", [(const(ikB), 0)]))
        put(L0, "br", "br", blk("br", "br", neg_imOm))
        put(Leta, "br", "br", blk("br", "br", lv))
        put(Leta, "br", "bt", -blk("br", "bt", couple))
        put(M, "br", "br", blk("br", "br", [(None, 0)]))
        put(L0, "bt", "ut", blk("bt", "ut", [(const(ikB), 0)]))
        put(L0, "bt", "br", blk("bt", "br", [(lambda r: -2.0 * b.b / r**2, 0)]))
        put(Leta, "bt", "br", blk("bt", "br", couple))
        put(L0, "bt", "bt", blk("bt", "bt", neg_imOm))
        put(Leta, "bt", "bt", blk("bt", "bt", lv))
        put(M, "bt", "bt", blk("bt", "bt", [(None, 0)]))
        put(L0, "bz", "uz", blk("bz", "uz", [(const(ikB), 0)]))
        put(L0, "bz", "bz", blk("bz", "bz", neg_imOm))
        put(Leta, "bz", "bz", blk("bz", "bz", lap))
        put(M, "bz", "bz", blk("bz", "bz", [(None, 0)]))
        return L0 + self.nu * Lnu + self.eta * Leta, M

    def eigs(self, m, kz, n_return=8, finite_cap=1.0e6):
        L, M = self.assemble(m, kz)
        w, V = sla.eig(L, M)
        keep = np.isfinite(w) & (np.abs(w) < finite_cap)
        w, V = w[keep], V[:, keep]
        order = np.lexsort((-w.imag, -w.real))
        return w[order][:n_return], V[:, order][:, :n_return]


class TCMRIOracle(TCOracle):
    """``AxisymmetricMRIDNSJax`` restated (``examples/taylor_couette_dns_jax.py:1359-1903``):
    perturbations (u_r, u_theta, u_z, b_r, b_theta, b_z) and the total pressure
    about V(r) e_theta and B0 e_z; b_r in the Dirichlet basis, b_theta in the
    Robin basis of d(r b_theta)/dr = 0, b_z in the Neumann basis; per axial
    mode one 7-field saddle-point block (``:1501-1571``).  State ``u`` holds the
    six fields, each (Nz, Nr - 2)."""

    NF = 6

    def __init__(self, base: CircularCouette, B0=0.1, nu=1.0e-3, eta_mag=1.0e-3, Nr=24, Nz=16, Lz=None, dt=2.0e-3,
                 dealias=1.5, time_integrator="CNAB2"):
        self.B0, self.eta = float(B0), float(eta_mag)
        Lz = float(Lz) if Lz is not None else 2.0 * math.pi / 3.0 * base.gap
        df = 2.0 / (base.R2 - base.R1)
        self.SBT = stencil_robin(int(Nr), base.R1 * df, base.R2 * df)
        self.SBZ = stencil_neumann(int(Nr))
        super().__init__(base, nu=nu, Nr=Nr, Nz=Nz, Lz=Lz, dt=dt, dealias=dealias, time_integrator=time_integrator)

    @property
    def bases(self):
        SD = self.rf.SD
        return (SD, SD, SD, SD, self.SBT, self.SBZ)

    def _radial_blocks(self):
        super()._radial_blocks()

    def _lap_xy(self, X, Y, kz, hoop):
        rf = self.rf
        lap = rf.form(X, Y, 2) + rf.form(X, Y, 1, lambda r: 1.0 / r) - kz**2 * rf.form(X, Y)
        return lap - rf.form(X, Y, 0, lambda r: 1.0 / r**2) if hoop else lap

    def _operator_part(self, kz, sign):
        """``_add_mhd_terms`` (``:1460-1499``) on the six fields, times sign."""
        rf, b, n = self.rf, self.base, self.Nd
        nu, eta, B0 = self.nu, self.eta, self.B0
        B = self.bases
        S = lambda i: slice(i * n, (i + 1) * n)  # noqa: E731
        X = np.zeros((6 * n, 6 * n), dtype=complex)
        ikB = 1j * kz * B0
        X[S(0), S(0)] = nu * self._lap_xy(B[0], B[0], kz, True)
        X[S(0), S(1)] = rf.form(B[0], B[1], 0, lambda r: 2.0 * (b.a + b.b / r**2))
        X[S(0), S(3)] = ikB * rf.form(B[0], B[3])
        X[S(1), S(1)] = nu * self._lap_xy(B[1], B[1], kz, True)
        X[S(1), S(0)] = -2.0 * b.a * rf.form(B[1], B[0])
        X[S(1), S(4)] = ikB * rf.form(B[1], B[4])
        X[S(2), S(2)] = nu * self._lap_xy(B[2], B[2], kz, False)
        X[S(2), S(5)] = ikB * rf.form(B[2], B[5])
        X[S(3), S(3)] = eta * self._lap_xy(B[3], B[3], kz, True)
        X[S(3), S(0)] = ikB * rf.form(B[3], B[0])
        X[S(4), S(4)] = eta * self._lap_xy(B[4], B[4], kz, True)
        X[S(4), S(1)] = ikB * rf.form(B[4], B[1])
        X[S(4), S(3)] = rf.form(B[4], B[3], 0, lambda r: -2.0 * b.b / r**2)
        X[S(5), S(5)] = eta * self._lap_xy(B[5], B[5], kz, False)
        X[S(5), S(2)] = ikB * rf.form(B[5], B[2])
        return sign * X

    def _mass6(self):
        n, B = self.Nd, self.bases
        M = np.zeros((6 * n, 6 * n))
        for i in range(6):
            M[i * n:(i + 1) * n, i * n:(i + 1) * n] = self.rf.form(B[i], B[i]) / self.dt
        return M

    def _build_operators(self, ls):
        """``:1501-1571``, times Lz; Limp in the (u_r, u_theta, u_z, Pi, b_r,
        b_theta, b_z) order, Lexp and mass on the six fields."""
        n, rf = self.Nd, self.rf
        SD, SP = rf.SD, rf.SP
        M6 = self._mass6()
        limp = np.zeros((self.Nz, 7 * n, 7 * n), dtype=complex)
        lexp = np.zeros((self.Nz, 6 * n, 6 * n), dtype=complex)
        mass = np.zeros((self.Nz, 6 * n, 6 * n), dtype=complex)
        perm = np.r_[0:3 * n, 4 * n:7 * n]  # field rows/columns inside the 7-field block
        for m in range(self.Nz):
            kz = self.kz[m]
            A6 = M6 + self._operator_part(kz, -ls)
            A = limp[m]
            A[np.ix_(perm, perm)] = A6
            P = slice(3 * n, 4 * n)
            A[0:n, P] += rf.form(SD, SP, 1)
            A[2 * n:3 * n, P] += 1j * kz * rf.form(SD, SP, 0)
            A[P, 0:n] += rf.form(SP, SD, 1) + rf.form(SP, SD, 0, lambda r: 1.0 / r)
            A[P, 2 * n:3 * n] += 1j * kz * rf.form(SP, SD, 0)
            lexp[m] = M6 + self._operator_part(kz, ls)
            mass[m] = M6
        return self.Lz * limp, self.Lz * lexp, self.Lz * mass

    # --------------------------------------------------------------- transforms
    def _orth(self, c, f):
        return c @ self.bases[f]

    def _phys_f(self, c, f):
        orth = self._orth(c, f)
        val = self._cheb_r(fourier_backward(orth, 0, self.Mz), self.Mr)
        dr = self.rf.df * self._cheb_r(fourier_backward(cheb_derivative(orth.T, 1).T, 0, self.Mz), self.Mr)
        dz = self._cheb_r(fourier_backward(1j * self.kz_odd[:, None] * orth, 0, self.Mz), self.Mr)
        return val, dr, dz

    def scalar_product_basis(self, v, S):
        c = self.Lz * fourier_forward(v, 0, self.Nz)
        s = cheb_scalar_product(c.T, self.Nr).T / self.rf.df
        return self.mask(s @ S.T)

    def _std_deriv(self, chat, axis):
        """T0.backward_primitive of (Nz, Nr) Chebyshev coefficients on the
        standard grid: axis 0 is d/dz, axis 1 d/dr."""
        if axis == 0:
            return self._cheb_r(fourier_backward(1j * self.kz_odd[:, None] * chat, 0, self.Nz), self.Nr)
        return self.rf.df * self._cheb_r(fourier_backward(cheb_derivative(chat.T, 1).T, 0, self.Nz), self.Nr)

    def nonlinear(self, x):
        """``:1603-1649``: advection minus Lorentz force on the velocity rows,
        minus the curl of the EMF u x b on the magnetic rows."""
        F = [self._phys_f(c, f) for f, c in enumerate(x)]
        (ur, urr, urz), (ut, utr, utz), (uz, uzr, uzz) = F[0], F[1], F[2]
        (br, brr, brz), (bt, btr, btz), (bz, bzr, bzz) = F[3], F[4], F[5]
        invr = (1.0 / self.r_pad)[None, :]
        au_r = ur * urr + uz * urz - ut * ut * invr
        au_t = ur * utr + uz * utz + ur * ut * invr
        au_z = ur * uzr + uz * uzz
        lb_r = br * brr + bz * brz - bt * bt * invr
        lb_t = br * btr + bz * btz + br * bt * invr
        lb_z = br * bzr + bz * bzz
        std = self._dealias_to_standard
        nu_r, nu_t, nu_z = std(au_r - lb_r), std(au_t - lb_t), std(au_z - lb_z)
        eps_r, eps_t, eps_z = std(ut * bz - uz * bt), std(uz * br - ur * bz), std(ur * bt - ut * br)
        coef = lambda v: cheb_forward(fourier_forward(v, 0, self.Nz).T, self.Nr).T  # noqa: E731
        er, et, ez = coef(eps_r), coef(eps_t), coef(eps_z)
        nb_r = self._std_deriv(et, 0)
        nb_t = -self._std_deriv(er, 0) + self._std_deriv(ez, 1)
        nb_z = -self._std_deriv(et, 1) - eps_t / self.rf.r[None, :]
        SD = self.rf.SD
        return (self.scalar_product_basis(nu_r, SD), self.scalar_product_basis(nu_t, SD),
                self.scalar_product_basis(nu_z, SD), self.scalar_product_basis(nb_r, SD),
                self.scalar_product_basis(nb_t, self.SBT), self.scalar_product_basis(nb_z, self.SBZ))

    def _modes(self, u):
        return np.concatenate(list(u), axis=1)

    def _unmodes(self, x):
        n = self.Nd
        return tuple(x[:, i * n:(i + 1) * n] for i in range(6))

    def solve_rhs(self, rhs_x, lu):
        """``_solve_mhd_rhs`` (``:1662-1669``): zero pressure rows, pinned mode 0."""
        n = self.Nd
        r6 = self._modes(rhs_x)
        rhs = np.concatenate([r6[:, :3 * n], np.zeros((self.Nz, n), dtype=complex), r6[:, 3 * n:]], axis=1)
        rhs[0, 3 * n] = 0.0
        sol = np.stack([sla.lu_solve(lu[m], rhs[m]) for m in range(self.Nz)])
        x = np.concatenate([sol[:, :3 * n], sol[:, 4 * n:]], axis=1)
        return self._unmodes(x), sol[:, 3 * n:4 * n]

    # ------------------------------------------------------------------ states
    def zero_state(self) -> TCState:
        z = tuple(np.zeros((self.Nz, self.Nd), dtype=complex) for _ in range(6))
        return TCState(u=z, p=np.zeros((self.Nz, self.Np), dtype=complex),
                       nonlinear_old=tuple(np.zeros_like(a) for a in z))

    def forward_basis(self, v, S):
        c = fourier_forward(np.asarray(v, dtype=complex), 0, self.Nz)
        s = cheb_scalar_product(c.T, self.Nr).T @ S.T
        M = S @ np.diag(np.where(np.arange(self.Nr) == 0, math.pi, math.pi / 2)) @ S.T
        return np.linalg.solve(M, s.T).T

    def state_from_physical(self, values) -> TCState:
        x = tuple(self.forward_basis(v, S) for v, S in zip(values, self.bases))
        return replace(self.zero_state(), u=x)

    def seed_linear_eigenmode(self, kz_mode=1, amp=1.0e-6, which=0):
        """``:1809-1842``: conducting MRI eigenmode, fields from blocks (0,1,2,4,5,6)."""
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        lin = TCMRILinearOracle(self.base, B0=self.B0, nu=self.nu, eta_mag=self.eta, N=self.Nr)
        w, vecs = lin.eigs(0, kz, n_return=which + 1)
        vec = positive_pivot_phase(vecs[:, which])
        n = lin.n
        kpos, kneg = int(kz_mode) % self.Nz, (-int(kz_mode)) % self.Nz
        comps = []
        for blk in (0, 1, 2, 4, 5, 6):
            block = vec[blk * n:(blk + 1) * n] * amp
            arr = np.zeros((self.Nz, self.Nd), dtype=complex)
            if kpos == kneg:
                arr[kpos, :n] = np.real(block)
            else:
                arr[kpos, :n] = 0.5 * block
                arr[kneg, :n] = 0.5 * np.conj(block)
            comps.append(arr)
        st = self.zero_state()
        return replace(st, u=tuple(comps), nonlinear_old=tuple(np.zeros_like(a) for a in comps),
                       have_old=0.0), complex(w[which])

    # ------------------------------------------------------------- diagnostics
    def backward_f(self, c, f):
        return self._cheb_r(fourier_backward(self._orth(c, f), 0, self.Nz), self.Nr)

    def fields_physical(self, s):
        return tuple(self.backward_f(c, f) for f, c in enumerate(s.u))

    def energy_parts(self, s):
        r = self.rf.r[None, :]
        F = self.fields_physical(s)
        ek = self._integrate(0.5 * sum(np.real(np.conj(v) * v) for v in F[:3]) * r)
        em = self._integrate(0.5 * sum(np.real(np.conj(v) * v) for v in F[3:]) * r)
        return ek, em

    def energy(self, s):
        return sum(self.energy_parts(s))

    def _div(self, s, fr, fz):
        orth_r = self._orth(s.u[fr], fr)
        d_r = self.rf.df * self._cheb_r(fourier_backward(cheb_derivative(orth_r.T, 1).T, 0, self.Nz), self.Nr)
        v_r = self.backward_f(s.u[fr], fr)
        d_z = self._cheb_r(fourier_backward(1j * self.kz_odd[:, None] * self._orth(s.u[fz], fz), 0, self.Nz), self.Nr)
        return d_r + v_r / self.rf.r[None, :] + d_z

    def diagnostics(self, s) -> dict:
        """``:1876-1886``."""
        ek, em = self.energy_parts(s)
        ends = npcheb.chebvander(np.array([-1.0, 1.0]), self.Nr - 1) @ self.rf.SD.T
        wall = max(float(np.max(np.abs(fourier_backward(c @ ends.T, 0, self.Nz)))) for c in s.u[:3])
        return {"Ekin": ek, "Emag": em, "E": ek + em,
                "divu": float(np.max(np.abs(self._div(s, 0, 2)))),
                "divb": float(np.max(np.abs(self._div(s, 3, 5)))), "wall_u": wall}


class TC3DMRIOracle(TCMRIOracle):
    """``TaylorCouetteMRIDNSJax`` restated (``examples/taylor_couette_dns_jax.py:1906-2245``):
    the conducting-wall MHD step on Fourier(theta) x Fourier(z) x Chebyshev(r),
    one 7-field block per (m, kz) with the base-flow advection -Omega d/dtheta
    of every field and the 2/r^2 azimuthal couplings of u and of b
    (``:2005-2059``); forms times 2 pi Lz.  Arrays are (Ntheta, Nz, Nr - 2)."""

    _synth = TC3DOracle._synth
    _cheb_last = TC3DOracle._cheb_last
    _dealias_to_standard = TC3DOracle._dealias_to_standard
    mask = TC3DOracle.mask
    mesh = TC3DOracle.mesh
    _integrate = TC3DOracle._integrate

    def __init__(self, base: CircularCouette, B0=0.1, nu=1.0e-3, eta_mag=1.0e-3, Nr=24, Ntheta=8, Nz=16, Lz=None,
                 dt=2.0e-3, dealias=1.5, time_integrator="CNAB2"):
        self.Ntheta = int(Ntheta)
        self.mint = fourier_wavenumbers(self.Ntheta)
        self.m_odd = fourier_wavenumbers(self.Ntheta, eliminate_highest=True)
        self.Mth = int(self.Ntheta * dealias) if dealias > 1.0 else self.Ntheta
        super().__init__(base, B0=B0, nu=nu, eta_mag=eta_mag, Nr=Nr, Nz=Nz, Lz=Lz, dt=dt, dealias=dealias,
                         time_integrator=time_integrator)

    def _operator_part3(self, m, kz, sign):
        """``_add_mhd_terms`` of the 3D class, times sign (6n x 6n)."""
        rf, b, n = self.rf, self.base, self.Nd
        B = self.bases
        S = lambda i: slice(i * n, (i + 1) * n)  # noqa: E731
        X = self._operator_part(kz, 1.0).astype(complex)
        for f in range(6):  # -Omega d/dtheta and the m^2 / r^2 part of the Laplacian
            nu_f = self.nu if f < 3 else self.eta
            X[S(f), S(f)] += (-1j * m * rf.form(B[f], B[f], 0, lambda r: b.a + b.b / r**2)
                              - nu_f * m * m * rf.form(B[f], B[f], 0, lambda r: 1.0 / r**2))
        X[S(0), S(1)] += -self.nu * 2j * m * rf.form(B[0], B[1], 0, lambda r: 1.0 / r**2)
        X[S(1), S(0)] += self.nu * 2j * m * rf.form(B[1], B[0], 0, lambda r: 1.0 / r**2)
        X[S(3), S(4)] += -self.eta * 2j * m * rf.form(B[3], B[4], 0, lambda r: 1.0 / r**2)
        X[S(4), S(3)] += self.eta * 2j * m * rf.form(B[4], B[3], 0, lambda r: 1.0 / r**2)
        return sign * X

    def _build_operators(self, ls):
        n, rf = self.Nd, self.rf
        SD, SP = rf.SD, rf.SP
        M6 = self._mass6()
        nm = self.Ntheta * self.Nz
        limp = np.zeros((nm, 7 * n, 7 * n), dtype=complex)
        lexp = np.zeros((nm, 6 * n, 6 * n), dtype=complex)
        mass = np.zeros((nm, 6 * n, 6 * n), dtype=complex)
        perm = np.r_[0:3 * n, 4 * n:7 * n]
        P = slice(3 * n, 4 * n)
        for it in range(self.Ntheta):
            for iz in range(self.Nz):
                q = it * self.Nz + iz
                m, kz = self.mint[it], self.kz[iz]
                A = limp[q]
                A[np.ix_(perm, perm)] = M6 + self._operator_part3(m, kz, -ls)
                A[0:n, P] += rf.form(SD, SP, 1)
                A[n:2 * n, P] += 1j * m * rf.form(SD, SP, 0, lambda r: 1.0 / r)
                A[2 * n:3 * n, P] += 1j * kz * rf.form(SD, SP, 0)
                A[P, 0:n] += rf.form(SP, SD, 1) + rf.form(SP, SD, 0, lambda r: 1.0 / r)
                A[P, n:2 * n] += 1j * m * rf.form(SP, SD, 0, lambda r: 1.0 / r)
                A[P, 2 * n:3 * n] += 1j * kz * rf.form(SP, SD, 0)
                lexp[q] = M6 + self._operator_part3(m, kz, ls)
                mass[q] = M6
        fz = 2.0 * math.pi * self.Lz
        return fz * limp, fz * lexp, fz * mass

    # --------------------------------------------------------------- transforms
    def _phys_f(self, c, f):
        orth = self._orth(c, f)
        d = lambda x: self._synth(x, self.Mth, self.Mz, self.Mr)  # noqa: E731
        dr = np.moveaxis(cheb_derivative(np.moveaxis(orth, -1, 0), 1), 0, -1)
        return (d(orth), self.rf.df * d(dr), d(1j * self.m_odd[:, None, None] * orth),
                d(1j * self.kz_odd[None, :, None] * orth))

    def scalar_product_basis(self, v, S):
        c = 2.0 * math.pi * self.Lz * fourier_forward(fourier_forward(v, 0, self.Ntheta), 1, self.Nz)
        s = np.moveaxis(cheb_scalar_product(np.moveaxis(c, -1, 0), self.Nr), 0, -1) / self.rf.df
        return self.mask(s @ S.T)

    def _coef(self, v):
        c = fourier_forward(fourier_forward(v, 0, self.Ntheta), 1, self.Nz)
        return np.moveaxis(cheb_forward(np.moveaxis(c, -1, 0), self.Nr), 0, -1)

    def _std(self, chat, which):
        syn = lambda x: self._synth(x, self.Ntheta, self.Nz, self.Nr)  # noqa: E731
        if which == "t":
            return syn(1j * self.m_odd[:, None, None] * chat)
        if which == "z":
            return syn(1j * self.kz_odd[None, :, None] * chat)
        return self.rf.df * syn(np.moveaxis(cheb_derivative(np.moveaxis(chat, -1, 0), 1), 0, -1))

    def nonlinear(self, x):
        """``:2141-2190``."""
        F = [self._phys_f(c, f) for f, c in enumerate(x)]
        (ur, urr, urt, urz), (ut, utr, utt, utz), (uz, uzr, uzt, uzz) = F[:3]
        (br, brr, brt, brz), (bt, btr, btt, btz), (bz, bzr, bzt, bzz) = F[3:]
        invr = (1.0 / self.r_pad)[None, None, :]
        au_r = ur * urr + (ut * invr) * urt + uz * urz - ut * ut * invr
        au_t = ur * utr + (ut * invr) * utt + uz * utz + ur * ut * invr
        au_z = ur * uzr + (ut * invr) * uzt + uz * uzz
        lb_r = br * brr + (bt * invr) * brt + bz * brz - bt * bt * invr
        lb_t = br * btr + (bt * invr) * btt + bz * btz + br * bt * invr
        lb_z = br * bzr + (bt * invr) * bzt + bz * bzz
        std = self._dealias_to_standard
        nu_r, nu_t, nu_z = std(au_r - lb_r), std(au_t - lb_t), std(au_z - lb_z)
        eps_r, eps_t, eps_z = std(ut * bz - uz * bt), std(uz * br - ur * bz), std(ur * bt - ut * br)
        er, et, ez = self._coef(eps_r), self._coef(eps_t), self._coef(eps_z)
        inv_s = 1.0 / self.rf.r[None, None, :]
        nb_r = -inv_s * self._std(ez, "t") + self._std(et, "z")
        nb_t = -self._std(er, "z") + self._std(ez, "r")
        nb_z = -self._std(et, "r") - eps_t * inv_s + inv_s * self._std(er, "t")
        SD = self.rf.SD
        return (self.scalar_product_basis(nu_r, SD), self.scalar_product_basis(nu_t, SD),
                self.scalar_product_basis(nu_z, SD), self.scalar_product_basis(nb_r, SD),
                self.scalar_product_basis(nb_t, self.SBT), self.scalar_product_basis(nb_z, self.SBZ))

    def _modes(self, u):
        nm = self.Ntheta * self.Nz
        return np.concatenate([c.reshape(nm, self.Nd) for c in u], axis=1)

    def _unmodes(self, x):
        n = self.Nd
        return tuple(x[:, i * n:(i + 1) * n].reshape(self.Ntheta, self.Nz, n) for i in range(6))

    def solve_rhs(self, rhs_x, lu):
        n, nm = self.Nd, self.Ntheta * self.Nz
        r6 = self._modes(rhs_x)
        rhs = np.concatenate([r6[:, :3 * n], np.zeros((nm, n), dtype=complex), r6[:, 3 * n:]], axis=1)
        rhs[0, 3 * n] = 0.0
        sol = np.stack([sla.lu_solve(lu[q], rhs[q]) for q in range(nm)])
        x = np.concatenate([sol[:, :3 * n], sol[:, 4 * n:]], axis=1)
        return self._unmodes(x), sol[:, 3 * n:4 * n].reshape(self.Ntheta, self.Nz, n)

    def _pin(self, limp):
        limp = limp.copy()
        row = 3 * self.Nd
        limp[0, row, :] = 0.0
        limp[0, row, row] = 1.0
        return limp

    # ------------------------------------------------------------------ states
    def zero_state(self) -> TCState:
        sh = (self.Ntheta, self.Nz, self.Nd)
        z = tuple(np.zeros(sh, dtype=complex) for _ in range(6))
        return TCState(u=z, p=np.zeros(sh, dtype=complex), nonlinear_old=tuple(np.zeros_like(a) for a in z))

    def forward_basis(self, v, S):
        c = fourier_forward(fourier_forward(np.asarray(v, dtype=complex), 0, self.Ntheta), 1, self.Nz)
        s = np.moveaxis(cheb_scalar_product(np.moveaxis(c, -1, 0), self.Nr), 0, -1) @ S.T
        M = S @ np.diag(np.where(np.arange(self.Nr) == 0, math.pi, math.pi / 2)) @ S.T
        return np.linalg.solve(M, s.reshape(-1, self.Nd).T).T.reshape(s.shape)

    def seed_linear_eigenmode(self, m=1, kz_mode=1, amp=1.0e-6, which=0):
        """``:2192-2230``: no pivot phase for m != 0."""
        kz = 2.0 * math.pi * int(kz_mode) / self.Lz
        lin = TCMRILinearOracle(self.base, B0=self.B0, nu=self.nu, eta_mag=self.eta, N=self.Nr)
        w, vecs = lin.eigs(int(m), kz, n_return=which + 1)
        vec = vecs[:, which]
        if int(m) == 0:
            vec = positive_pivot_phase(vec)
        n = lin.n
        mpos, mneg = int(m) % self.Ntheta, (-int(m)) % self.Ntheta
        kpos, kneg = int(kz_mode) % self.Nz, (-int(kz_mode)) % self.Nz
        comps = []
        for blk in (0, 1, 2, 4, 5, 6):
            block = vec[blk * n:(blk + 1) * n] * amp
            arr = np.zeros((self.Ntheta, self.Nz, self.Nd), dtype=complex)
            if mpos == mneg and kpos == kneg:
                arr[mpos, kpos, :n] = np.real(block)
            else:
                arr[mpos, kpos, :n] = 0.5 * block
                arr[mneg, kneg, :n] = 0.5 * np.conj(block)
            comps.append(arr)
        st = self.zero_state()
        return replace(st, u=tuple(comps), nonlinear_old=tuple(np.zeros_like(a) for a in comps)), complex(w[which])

    # ------------------------------------------------------------- diagnostics
    def backward_f(self, c, f):
        return self._synth(self._orth(c, f), self.Ntheta, self.Nz, self.Nr)

    def energy_parts(self, s):
        r = self.rf.r[None, None, :]
        F = self.fields_physical(s)
        ek = self._integrate(0.5 * sum(np.real(np.conj(v) * v) for v in F[:3]) * r)
        em = self._integrate(0.5 * sum(np.real(np.conj(v) * v) for v in F[3:]) * r)
        return ek, em

    def _div(self, s, fr, ft, fz):
        """``:2232-2245``: d/dr + 1/r of the radial, (1/r) d/dtheta of the
        azimuthal and d/dz of the axial component."""
        syn = lambda x: self._synth(x, self.Ntheta, self.Nz, self.Nr)  # noqa: E731
        orth_r = self._orth(s.u[fr], fr)
        d_r = self.rf.df * syn(np.moveaxis(cheb_derivative(np.moveaxis(orth_r, -1, 0), 1), 0, -1))
        r = self.rf.r[None, None, :]
        return (d_r + syn(orth_r) / r + syn(1j * self.m_odd[:, None, None] * self._orth(s.u[ft], ft)) / r
                + syn(1j * self.kz_odd[None, :, None] * self._orth(s.u[fz], fz)))

    def diagnostics(self, s) -> dict:
        ek, em = self.energy_parts(s)
        ends = npcheb.chebvander(np.array([-1.0, 1.0]), self.Nr - 1) @ self.rf.SD.T
        wall = max(float(np.max(np.abs(fourier_backward(fourier_backward(c @ ends.T, 0, self.Ntheta), 1, self.Nz))))
                   for c in s.u[:3])
        return {"Ekin": ek, "Emag": em, "E": ek + em, "divu": float(np.max(np.abs(self._div(s, 0, 1, 2)))),
                "divb": float(np.max(np.abs(self._div(s, 3, 4, 5)))), "wall_u": wall}
