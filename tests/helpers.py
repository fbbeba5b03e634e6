"""Shared helpers for the parity tests (oracle <-> product state conversion)."""

from __future__ import annotations

import math

import numpy as np

from jaxfun_amd.pcf import KMMState
from oracle.kmm import OracleState, PCFOracle


def to_oracle(s: KMMState) -> OracleState:
    return OracleState(u=tuple(np.array(a) for a in s.u), g=np.array(s.g),
                       nonlinear_old=tuple(np.array(a) for a in s.nonlinear_old),
                       nonlinear_older=tuple(np.array(a) for a in s.nonlinear_older),
                       solution_old=tuple(np.array(a) for a in s.solution_old),
                       solution_older=tuple(np.array(a) for a in s.solution_older),
                       history_steps=float(s.history_steps), have_old=float(s.have_old),
                       previous_dt=float(s.previous_dt))


def leaves(s) -> list:
    out = list(s.u) + [s.g]
    for grp in (s.nonlinear_old, s.nonlinear_older, s.solution_old, s.solution_older):
        out.extend(grp)
    return out


LEAF_NAMES = ["u0", "u1", "u2", "g"] + [f"{t}.{c}" for t in ("nl_old", "nl_older", "sol_old", "sol_older")
                                        for c in ("u", "g", "v", "w")]


def max_rel_errors(a, b) -> dict:
    """Per-leaf max |a-b| / scale, b the reference.

    scale = max(max |b_leaf|, 1e-3 * max |b| over the leaf's group), the groups
    being the state (u0, u1, u2, g) and each history tuple.  The floor keeps a
    leaf that is pure round-off in both (e.g. the mean-mode nonlinear rows of a
    symmetric initial field) from reading as an O(1) relative error.
    """
    la, lb = leaves(a), leaves(b)
    groups = [range(0, 4)] + [range(4 + 4 * k, 8 + 4 * k) for k in range(4)]
    out = {}
    for grp in groups:
        gmax = max((float(np.max(np.abs(np.asarray(lb[i])))) for i in grp), default=0.0)
        for i in grp:
            x, y = np.asarray(la[i]), np.asarray(lb[i])
            scale = max(float(np.max(np.abs(y))), 1e-3 * gmax)
            err = float(np.max(np.abs(x - y)))
            out[LEAF_NAMES[i]] = err / scale if scale > 0 else err
    return out


def oracle_for(solver) -> PCFOracle:
    return PCFOracle(N=solver.N, domain=solver.domain, Re=solver.Re, U_wall=solver.U_wall, dt=solver.dt,
                     padding_factor=solver.padding_factor,
                     perturbation_amplitude=solver.perturbation_amplitude,
                     time_integrator=getattr(solver, "time_integrator", "SBDF3"))


# ------------------------------------------------------------------ MHD
MHD_A_GROUPS = ("A", "nonlinear_A_old", "nonlinear_A_older", "A_old", "A_older")
MHD_LEAF_NAMES = LEAF_NAMES + [f"{g}.{c}" for g in MHD_A_GROUPS for c in "xyz"]


def to_mhd_oracle(s):
    from oracle.mhd import MHDOracleState

    return MHDOracleState(flow=to_oracle(s.flow), **{g: tuple(np.array(a) for a in getattr(s, g))
                                                      for g in MHD_A_GROUPS})


def mhd_leaves(s) -> list:
    out = leaves(s.flow)
    for g in MHD_A_GROUPS:
        out.extend(getattr(s, g))
    return out


def mhd_max_rel_errors(a, b) -> dict:
    """max_rel_errors over the flow leaves plus each A group (scale floored at
    1e-3 of the group's largest entry)."""
    out = max_rel_errors(a.flow, b.flow)
    la, lb = mhd_leaves(a), mhd_leaves(b)
    for k in range(len(MHD_A_GROUPS)):
        grp = range(20 + 3 * k, 23 + 3 * k)
        gmax = max(float(np.max(np.abs(np.asarray(lb[i])))) for i in grp)
        for i in grp:
            x, y = np.asarray(la[i]), np.asarray(lb[i])
            scale = max(float(np.max(np.abs(y))), 1e-3 * gmax)
            err = float(np.max(np.abs(x - y)))
            out[MHD_LEAF_NAMES[i]] = err / scale if scale > 0 else err
    return out


def mhd_oracle_for(solver):
    from oracle.mhd import MHDOracle

    return MHDOracle(N=solver.N, domain=solver.domain, Re=solver.Re, Rm=solver.Rm, U_wall=solver.U_wall,
                     dt=solver.dt, padding_factor=solver.padding_factor,
                     perturbation_amplitude=solver.perturbation_amplitude,
                     magnetic_amplitude=solver.magnetic_amplitude, background_b=solver.background_b,
                     time_integrator=solver.time_integrator)


def shear_oracle_for(solver):
    from oracle.mhd import RotatingShearMHDOracle

    return RotatingShearMHDOracle(N=solver.N, domain=solver.domain, Re=solver.Re, Rm=solver.Rm, omega=solver.omega,
                                  shear_rate=solver.shear_rate, background_b=solver.background_b, dt=solver.dt,
                                  padding_factor=solver.padding_factor,
                                  perturbation_amplitude=solver.perturbation_amplitude,
                                  magnetic_amplitude=solver.magnetic_amplitude,
                                  time_integrator=solver.time_integrator)


def insulating_oracle_for(solver):
    from oracle.mhd_insulating import InsulatingMHDOracle

    return InsulatingMHDOracle(N=solver.N, domain=solver.domain, Re=solver.Re, Rm=solver.Rm, omega=solver.omega,
                               shear_rate=solver.shear_rate, background_b=solver.background_b, dt=solver.dt,
                               padding_factor=solver.padding_factor,
                               perturbation_amplitude=solver.perturbation_amplitude,
                               magnetic_amplitude=solver.magnetic_amplitude,
                               time_integrator=solver.time_integrator)


# ------------------------------------------------------------------ host-only geometry
class HostSolver:
    """The host-side geometry of PlaneCouetteFluctuationHIP without a GPU handle:
    enough for the seeded initial conditions (``jaxfun_amd.ic``), which need
    only the closed-form setup matrices of the library."""

    def __init__(self, N, Ly=4 * math.pi, Lz=2 * math.pi):
        from jaxfun_amd import _native as nat
        from jaxfun_amd.pcf import PlaneCouetteFluctuationHIP as S

        self.N = tuple(N)
        self.Ly, self.Lz = Ly, Lz
        self.domain = ((-1.0, 1.0), (0.0, Ly), (0.0, Lz))
        c = nat.PcfConfig()
        c.abi_version = nat.ABI_VERSION
        c.n0, c.ny, c.nz = self.N
        c.lx[0], c.lx[1] = -1.0, 1.0
        c.ly, c.lz = Ly, Lz
        c.nu, c.u_wall, c.dt = 1 / 400, 1.0, 1e-3
        c.padding_y = c.padding_z = 1.5
        c.precision = nat.PREC_FP64
        c.time_integrator = nat.TI_SBDF3
        c.nonlinear_form = nat.NL_ROTATIONAL
        self._cfg = c
        self.matrix = lambda name: nat.setup_matrix(self._cfg, name)
        self.wavenumbers = S.wavenumbers.__get__(self)
        self.mask_nyquist = S.mask_nyquist.__get__(self)
        self._with_zero_history = S._with_zero_history.__get__(self)
        self.nb, self.nd = self.N[0] - 4, self.N[0] - 2
