"""CPU FP64 oracle of the insulating-wall rotating-shear plane-Couette MHD step.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  A numpy restatement of
``PlaneCouetteMRIShearpyInsulatingJax``
(``examples/pcf_mhd_mri_shearpy_jax.py:516-836``) over its parents
``PlaneCouetteMRIShearpyJax`` (``:30-513``) and ``PlaneCouetteMHDJax``
(``examples/pcf_mhd_jax.py``), for SBDF3 with the IMEXRK3 start-up:

* base flow U_b = -S x (shear rate S), Coriolis forcing (-2 Omega u1, 2 Omega u0, 0)
  and a uniform background field B0 added to B (``:86-92,149-176``);
* A_x in the Dirichlet space TD (wall gauge), A_y, A_z and B, J in the
  orthogonal Chebyshev space TC (``:548-557``);
* the A_y/A_z subsystem per Fourier mode rotated to P = cy Ay + cz Az,
  Q = cz Ay - cy Az (cy = ky/k, cz = kz/k; (1, 0) at k = 0) and solved with the
  radial operator h - gamma eta (D2 - k^2 h) whose last two rows are replaced by
  the vacuum-matching rows P'(+-1) = 0 and Q'(-1) - k Q(-1) = 0,
  Q'(1) + k Q(1) = 0 (``:559-631,661-712``), by a pivoted dense LU per mode as
  the reference does (``jax.scipy.linalg.lu_factor``);
* diagnostics: the hydrodynamic set, Emag and divB_L2 of the fluctuation field,
  Emag_total with B0, and the insulating boundary residual (``:802-836``).

Pins (``tests/test_mhd_insulating_oracle.py``): the vacuum-matching residual
after steps (< 1e-10, the reference's own test threshold,
``tests/couette/test_pcf_sbdf3_jax.py:164-167``); div B = 0; the
hydrodynamic oracle recovered exactly at zero field with S = -1, Omega = 0;
the P and Q eigenmodes of the bordered resistive operator decaying by the scalar
IMEXRK3/SBDF3 recurrence (no flow, no shear: the EMF vanishes); and B = curl A
against nodal derivatives.
"""

from __future__ import annotations

import math

import numpy as np

from .kmm import cheb_norm_squared, fourier_forward
from .mhd import MHDOracleState, RotatingShearMHDOracle


def _cheb_D(n: int) -> np.ndarray:
    """Chebyshev derivative matrix: T_b' = sum_a D[a, b] T_a (``Chebyshev.py:320-355``)."""
    D = np.zeros((n, n))
    for a in range(n):
        for b in range(a + 1, n, 2):
            D[a, b] = b if a == 0 else 2 * b
    return D


class InsulatingMHDOracle(RotatingShearMHDOracle):
    """``PlaneCouetteMRIShearpyInsulatingJax`` for SBDF3 / rotational / optimized:
    the rotating-shear oracle (base flow, Coriolis, B0) with the insulating A
    spaces and wall rows."""

    def __init__(self, N=(17, 16, 16), domain=((-1.0, 1.0), (0.0, 4.0 * math.pi), (0.0, 2.0 * math.pi)),
                 Re: float = 400.0, Rm: float | None = None, omega: float = 1.0, shear_rate: float = 1.0,
                 background_b=(0.0, 0.0, 0.025), dt: float = 0.01, padding_factor=(1.0, 1.5, 1.5),
                 perturbation_amplitude: float = 0.05, magnetic_amplitude: float = 0.0,
                 time_integrator: str = "SBDF3") -> None:
        super().__init__(N=N, domain=domain, Re=Re, Rm=Rm, omega=omega, shear_rate=shear_rate,
                         background_b=background_b, dt=dt, padding_factor=padding_factor,
                         perturbation_amplitude=perturbation_amplitude, magnetic_amplitude=magnetic_amplitude,
                         time_integrator=time_integrator)
        n = self.N0
        D = _cheb_D(n)
        self.CD = D
        self.M1 = np.diag(cheb_norm_squared(n))
        self.D2C = self.M1 @ D @ D          # (T_i, T_j'')_w, shearpy :567-572
        j = np.arange(n)
        self.d0 = np.stack([(-1.0) ** j, np.ones(n)])          # T_j(-1), T_j(1)
        self.d1 = np.stack([(-1.0) ** (j + 1) * j * j, 1.0 * j * j])  # T_j'(-1), T_j'(1)
        K1 = np.broadcast_to(self.K1[0], (self.N[1], self.N[2]))
        K2 = np.broadcast_to(self.K2[0], (self.N[1], self.N[2]))
        kmag = np.sqrt(K1**2 + K2**2)
        knz = np.where(kmag > 0, kmag, 1.0)
        self.kmag = kmag
        self.cy = np.where(kmag > 0, K1 / knz, 1.0)
        self.cz = np.where(kmag > 0, K2 / knz, 0.0)
        self._bordered = {}

    def initial_state(self):
        """``initial_state`` (shearpy :99-147), the default channel-mode seed with
        the ``ax_yz`` magnetic seed; A projected onto (TD, TC, TC)."""
        flow, ax = self.shearpy_seed()
        z = self.zero_A()
        A = (self.mask(self.tp_forward(ax.astype(complex), self.SD, self.MD1)), z[1], z[2])
        return self.reset_mhd_history(MHDOracleState(flow=flow, A=A))

    # ------------------------------------------------------------ spaces
    def b_stencils(self):
        """``b_coeff_spaces`` = TC^3 (shearpy :548-557)."""
        return (None, None, None)

    def a_stencils(self):
        """``a_coeff_spaces`` = (TD, TC, TC)."""
        return (self.SD, None, None)

    def zero_A(self):
        shp = (self.N[1], self.N[2])
        return (np.zeros((self.ND,) + shp, complex), np.zeros((self.N0,) + shp, complex),
                np.zeros((self.N0,) + shp, complex))

    def B_from_A(self, A):
        """``update_B_from_A`` with A in (TD, TC, TC), B in TC^3."""
        AO = (self.to_orth(A[0], self.SD), A[1], A[2])
        bx = self._dcoef(AO[2], 0, 1, 0) - self._dcoef(AO[1], 0, 0, 1)
        by = self._dcoef(AO[0], 0, 0, 1) - self._dcoef(AO[2], 1, 0, 0)
        bz = self._dcoef(AO[1], 1, 0, 0) - self._dcoef(AO[0], 0, 1, 0)
        return (self.mask(bx), self.mask(by), self.mask(bz))

    def J_from_B(self, B):
        """``update_J_from_B`` with B, J in TC^3 (no projection)."""
        jx = self._dcoef(B[2], 0, 1, 0) - self._dcoef(B[1], 0, 0, 1)
        jy = self._dcoef(B[0], 0, 0, 1) - self._dcoef(B[2], 1, 0, 0)
        jz = self._dcoef(B[1], 1, 0, 0) - self._dcoef(B[0], 0, 1, 0)
        return (self.mask(jx), self.mask(jy), self.mask(jz))

    def backward_B(self, B, padded=True):
        return tuple(self.tp_backward(b, None, padded) for b in B)

    def backward_J(self, J, padded=True):
        return tuple(self.tp_backward(j, None, padded) for j in J)

    def tp_forward_C(self, u):
        """TCp.forward with truncation: Chebyshev coefficients (scalar products / h)."""
        from .kmm import cheb_scalar_product

        L = cheb_scalar_product(u, self.N0) / self.h[:, None, None]
        z = fourier_forward(L, 2, self.N[2])
        return fourier_forward(z, 1, self.N[1])

    # ------------------------------------------------------------ physics hooks
    def forward_emf(self, emf):
        """``_A_forward_emf`` (shearpy :654-659): TD for x, TC for y and z."""
        return (self.mask(self.tp_forward_D(emf[0])), self.mask(self.tp_forward_C(emf[1])),
                self.mask(self.tp_forward_C(emf[2])))

    def _radial(self, Mat, c):
        return np.tensordot(Mat, c, axes=(1, 0))

    def A_mass(self, A):
        """``_A_mass_rhs`` (shearpy :636-641)."""
        return (self.MA_apply(A[0]), self._radial(self.M1, A[1]), self._radial(self.M1, A[2]))

    def A_lap(self, A):
        """``_A_eta_lap`` (shearpy :643-652)."""
        k2 = (self.kmag**2)[None]

        def lap(c):
            return self.eta * (self._radial(self.D2C, c) - k2 * self._radial(self.M1, c))

        return (self.LA_apply(A[0]), lap(A[1]), lap(A[2]))

    def _bordered_mats(self, gamma):
        """``bordered_factors`` (shearpy :599-613): per-mode P and Q systems."""
        key = round(gamma, 17)
        if key not in self._bordered:
            n = self.N0
            k = self.kmag.reshape(-1)
            base = self.M1[None] - gamma * self.eta * (self.D2C[None] - (k**2)[:, None, None] * self.M1[None])
            sp = base.copy()
            sp[:, n - 2, :] = self.d1[0]
            sp[:, n - 1, :] = self.d1[1]
            sq = base.copy()
            sq[:, n - 2, :] = self.d1[0][None] - k[:, None] * self.d0[0][None]
            sq[:, n - 1, :] = self.d1[1][None] + k[:, None] * self.d0[1][None]
            self._bordered[key] = (sp, sq)
        return self._bordered[key]

    def _bordered_solve(self, mats, rhs):
        """``_A_bordered_solve`` (shearpy :661-670): pivoted dense LU per mode."""
        n = self.N0
        modes = self._modes(rhs).copy()
        modes[:, n - 2:] = 0.0
        sol = np.linalg.solve(mats, modes[:, :, None])[:, :, 0]
        return self._unmodes(sol)

    def A_solve_gamma(self, gamma, rhs):
        """``_A_solve`` (shearpy :698-712)."""
        ax = self.A_solve(self._A_fac(gamma), (rhs[0],))[0]
        sp, sq = self._bordered_mats(gamma)
        cy, cz = self.cy[None], self.cz[None]
        p = self._bordered_solve(sp, cy * rhs[1] + cz * rhs[2])
        q = self._bordered_solve(sq, cz * rhs[1] - cy * rhs[2])
        return (ax, self.mask(cy * p + cz * q), self.mask(cz * p - cy * q))

    # ------------------------------------------------------------ diagnostics
    def insulating_bc_residual(self, s) -> float:
        """``insulating_bc_residual`` (shearpy :802-830)."""
        B = self.B_from_A(s.A)
        w = lambda rows, c: np.tensordot(rows, c, axes=(1, 0))  # noqa: E731
        bx0, bx1 = w(self.d0, B[0]), w(self.d1, B[0])
        by0, bz0 = w(self.d0, B[1]), w(self.d0, B[2])
        k = self.kmag[None]
        knz = np.where(k > 0, k, 1.0)
        K1 = np.broadcast_to(self.K1[0], k.shape[1:])[None]
        K2 = np.broadcast_to(self.K2[0], k.shape[1:])[None]
        sign = np.array([-1.0, 1.0])[:, None, None]
        res = (bx1 + sign * k * bx0, by0 + sign * (1j * K1 / knz) * bx0, bz0 + sign * (1j * K2 / knz) * bx0)
        return float(max(np.max(np.abs(r)) for r in res))

    def diagnostics(self, s) -> dict:
        out = super().diagnostics(s)
        out["insulating_bc_residual"] = self.insulating_bc_residual(s)
        return out

    def divB(self, B):
        return (self.tp_backward(B[0], None, padded=False, kx=1) + self.tp_backward(B[1], None, padded=False, k1=1)
                + self.tp_backward(B[2], None, padded=False, k2=1))
