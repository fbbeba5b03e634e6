"""Plane-Couette MHD on the MI355X: drop-ins for the reference's vector-potential solvers.

``PlaneCouetteMHDHIP`` mirrors ``PlaneCouetteMHDJax``
(reference ``examples/pcf_mhd_jax.py:69-789``): the KMM flow of
``PlaneCouetteFluctuationHIP`` coupled to a magnetic vector potential
A = (Ax, Ay, Az) in the Dirichlet space TD^3 (conducting walls), with
B = curl A, J = curl B, the Lorentz force J x B in the flow forcing and the
induction EMF u_tot x B driving A, advanced on one SBDF3 timeline with an
IMEXRK3 start-up.  ``PlaneCouetteMRIShearpyHIP`` and
``PlaneCouetteMRIShearpyInsulatingHIP`` mirror the rotating-shear classes of
``examples/pcf_mhd_mri_shearpy_jax.py`` (conducting and insulating walls), the
two solvers the reference's vector-potential runner constructs.  The state is
``MHDState`` with the reference's fields in its full-complex layout; every
step runs in libpcfdns.so.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, replace

import numpy as np

from . import _native as nat
from .pcf import KMMState, PlaneCouetteFluctuationHIP, state_leaves

_A_GROUPS = (("A", ""), ("nonlinear_A_old", "NL_A_OLD_"), ("nonlinear_A_older", "NL_A_OLDER_"),
             ("A_old", "A_OLD_"), ("A_older", "A_OLDER_"))


@dataclass(frozen=True)
class MHDState:
    """Coefficient-space MHD state (reference ``examples/pcf_mhd_jax.py:34-66``).

    ``flow`` is the ``KMMState``; ``A`` and its histories are triples of
    complex128 (n0-2, Ny, Nz) arrays in TD.
    """

    flow: KMMState
    A: tuple
    nonlinear_A_old: tuple | None = None
    nonlinear_A_older: tuple | None = None
    A_old: tuple | None = None
    A_older: tuple | None = None


class PlaneCouetteMHDHIP(PlaneCouetteFluctuationHIP):
    """Conducting-wall plane-Couette MHD whose step runs in HIP on one GPU.

    Constructor follows ``PlaneCouetteMHDJax.__init__``
    (``examples/pcf_mhd_jax.py:76-114``): ``Rm`` defaults to ``Re`` and the
    magnetic diffusivity is ``eta = U_wall / Rm``.  ``background_b`` is the
    uniform field B0 of the reference's rotating-shear subclass
    (``examples/pcf_mhd_mri_shearpy_jax.py:149-151``); the conducting-wall
    class has none.  Built options are those of ``PlaneCouetteFluctuationHIP``.
    """

    def __init__(self, N=(17, 16, 16), domain=((-1.0, 1.0), (0.0, 4.0 * math.pi), (0.0, 2.0 * math.pi)),
                 Re: float = 400.0, Rm: float | None = None, U_wall: float = 1.0, dt: float = 0.01,
                 family: str = "C", padding_factor=(1.0, 1.5, 1.5), perturbation_amplitude: float = 0.1,
                 magnetic_amplitude: float = 0.05, time_integrator: str | None = None,
                 nonlinear_form: str = "gradient", coefficient_path: str = "transform",
                 solve_batching: str = "batched", background_b=(0.0, 0.0, 0.0), precision: str = "fp64",
                 device: int = 0) -> None:
        self.Rm = float(Re if Rm is None else Rm)
        if not self.Rm > 0:
            raise ValueError("Rm must be positive")
        if time_integrator not in {"SBDF3", "IMEXRK3", "IMEXRK222", "CNAB2", None}:
            raise ValueError("time_integrator must be one of {'IMEXRK222', 'IMEXRK3', 'CNAB2', 'SBDF3'}")
        if time_integrator is None:  # the reference's default ARS stepper
            time_integrator = "IMEXRK222"
        if time_integrator not in {"SBDF3", "IMEXRK222", "CNAB2"}:
            raise NotImplementedError("the MHD step is built for SBDF3, IMEXRK222 and CNAB2")
        self.eta = float(U_wall) / self.Rm
        self.magnetic_amplitude = float(magnetic_amplitude)
        self.background_b = tuple(float(v) for v in background_b)
        super().__init__(N=N, domain=domain, Re=Re, U_wall=U_wall, dt=dt, family=family,
                         padding_factor=padding_factor, perturbation_amplitude=perturbation_amplitude,
                         time_integrator=time_integrator, nonlinear_form=nonlinear_form,
                         coefficient_path=coefficient_path, solve_batching=solve_batching, precision=precision,
                         device=device)

    def _configure(self, cfg) -> None:
        cfg.mhd = 1
        cfg.eta = self.eta
        for i in range(3):
            cfg.b0[i] = self.background_b[i]

    # ------------------------------------------------------------ states
    def _zero_A(self):
        return tuple(np.zeros((self._a_rows(i), self.N[1], self.N[2]), complex) for i in range(3))

    def _new_state(self, flow: KMMState, A) -> MHDState:
        """``_new_mhd_state`` / ``_ensure_mhd_history`` (``pcf_mhd_jax.py:217-251``):
        fresh flow and A histories on one timeline."""
        flow = self._with_zero_history(flow.u, flow.g)
        z = self._zero_A()
        return MHDState(flow=flow, A=tuple(np.asarray(a, complex) for a in A), nonlinear_A_old=z,
                        nonlinear_A_older=z, A_old=z, A_older=z)

    def zero_state(self) -> MHDState:
        return self._new_state(super().zero_state(), self._zero_A())

    def initial_state(self) -> MHDState:
        """``initial_state`` (``pcf_mhd_jax.py:253-267``): the flow seed plus
        Ax = a (1-x^2) sin(ky y) sin(kz z) / kz with ky, kz the fundamental
        wavenumbers (exact two-term Fourier expansions, projected onto TD)."""
        flow = super().initial_state()
        n0, n1, n2 = self.N
        A = list(self._zero_A())
        amp = self.magnetic_amplitude
        if amp != 0.0:
            kz = 2.0 * math.pi / self.Lz
            xn = np.cos(np.pi + (2 * np.arange(n0) + 1) * np.pi / (2 * n0))
            prof = amp * (1.0 - xn**2) / kz
            cy = np.zeros(n1, complex)
            cz = np.zeros(n2, complex)
            cy[1], cy[-1] = 1 / 2j, -1 / 2j
            cz[1], cz[-1] = 1 / 2j, -1 / 2j
            A[0] = self.mask_nyquist(np.einsum("i,jk->ijk", self.matrix("FD") @ prof, np.outer(cy, cz)))
        return self._new_state(flow, A)

    # ------------------------------------------------------------ device I/O
    def _device_mhd_history(self, state: MHDState) -> MHDState:
        """``_ensure_mhd_history`` (``pcf_mhd_jax.py:217-248``): under SBDF3 the
        flow history follows ``_ensure_flow_history`` (restarted when a flow
        level is missing) and a complete A history is kept as loaded; a
        missing A level restarts flow and A together.  CNAB2 starts a missing
        ``nonlinear_A_old`` at zero (the flow rule of ``_device_history``
        applies to ``flow``).  Levels the step does not read travel as zeros."""
        flow = state.flow
        if self.time_integrator == "SBDF3":
            if all(getattr(state, name) is not None for name, _ in _A_GROUPS[1:]):
                return replace(state, flow=self._device_history(flow))
            return self._new_state(flow, state.A)
        z = self._zero_A()
        return replace(state, flow=self._device_history(flow),
                       **{name: z for name, _ in _A_GROUPS[1:] if getattr(state, name) is None})

    def load(self, state: MHDState) -> None:
        state = self._device_mhd_history(state)
        self._begin_load(state)
        super().load(state.flow)
        for name, tag in _A_GROUPS:
            for comp, arr in zip("XYZ", getattr(state, name)):
                self._set(f"A{comp}" if not tag else f"{tag}{comp}", arr)
        self._end_load()
        self._loaded = None

    def _result(self, state: MHDState, device: bool = False) -> MHDState:
        """The device state after steps from ``state``: SBDF3 with its own
        rings; IMEXRK222 with every history field of ``state`` passed through
        (``_step_imex_mhd`` returns ``replace(state, ...)``,
        ``examples/pcf_mhd_jax.py:482-486``); CNAB2 with the step's nonlinear
        rows of flow and A as the only levels (``:383-387``)."""
        st = self.fetch_device() if device else self.fetch()
        if self.time_integrator == "IMEXRK222":
            st = replace(state, flow=replace(state.flow, u=st.flow.u, g=st.flow.g), A=st.A)
        elif self.time_integrator == "CNAB2":
            f = st.flow
            st = MHDState(flow=KMMState(u=f.u, g=f.g, nonlinear_old=f.nonlinear_old, have_old=f.have_old,
                                        previous_dt=f.previous_dt),
                          A=st.A, nonlinear_A_old=st.nonlinear_A_old)
        self._loaded = st
        return st

    def _a_rows(self, comp: int) -> int:
        """Rows of A component comp (Ax in TD; Ay, Az in TD for conducting walls)."""
        return self.nd

    def fetch(self) -> MHDState:
        flow = super().fetch()
        groups = {}
        for name, tag in _A_GROUPS:
            groups[name] = tuple(self._get(f"A{c}" if not tag else f"{tag}{c}", self._a_rows(i))
                                 for i, c in enumerate("XYZ"))
        st = MHDState(flow=flow, **groups)
        self._loaded = st
        return st

    # ------------------------------------------------------------ reference API
    def diagnostics(self, state: MHDState | None = None) -> dict:
        """``diagnostics`` (``pcf_mhd_jax.py:776-789``): the flow diagnostics plus
        ``divB_L2`` and ``Emag``."""
        out = super().diagnostics(state)
        m = np.zeros(4)
        nat.check(self._lib.pcf_mhd_diagnostics(self._h, m.ctypes.data_as(C.POINTER(C.c_double))))
        self._mhd_diag = m
        out["divB_L2"] = float(m[1])
        out["Emag"] = float(m[0])
        out["Emag_total"] = float(m[2])
        return out

    def vjp(self, state: MHDState, steps: int, cotangent=None, checkpoint_every: int = 20):
        """Run ``steps`` SBDF3 steps from ``state`` and pull a cotangent back to
        every leaf of it (flow, A and their histories): the reverse mode of the
        coupled flow / vector-potential step that the reference takes with
        ``jax.value_and_grad`` (``production/autograd_cost_benchmark.py:56-61,228-330``
        registers ``PlaneCouetteMRIShearpyJax`` with SBDF3).

        ``cotangent`` is ``((u0, u1, u2, g), (Ax, Ay, Az))`` of the final state
        in jax's convention (directional derivative ``Re sum(c * d)``); ``None``
        differentiates Epert + Emag of the final state.  Conducting walls
        (A in TD^3), FP64.  Returns ``(final_state, gradient)`` with the
        gradient an ``MHDState``."""
        if self.time_integrator != "SBDF3":
            raise NotImplementedError("the reverse mode covers the SBDF3 step only")
        if self.precision != "fp64":
            raise NotImplementedError("the MHD reverse mode runs in FP64")
        self._ensure_loaded(state)
        self._loaded = None
        seed = nat.SEED_ENERGY
        if cotangent is not None:
            flow_cot, a_cot = cotangent
            for name, arr in list(zip(("U0", "U1", "U2", "G"), flow_cot)) + list(zip(("AX", "AY", "AZ"), a_cot)):
                a = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128))
                nat.check(self._lib.pcf_set_cotangent(self._h, nat.FIELD_ID[name],
                                                      a.ctypes.data_as(C.POINTER(C.c_double))))
            seed = nat.SEED_COTANGENT
        nat.check(self._lib.pcf_vjp(self._h, int(steps), int(checkpoint_every), seed))
        grad = self._mhd_gradient()
        return self.fetch(), grad

    def value_and_grad(self, state: MHDState, steps: int, checkpoint_every: int = 20):
        """Final Epert + Emag and its gradient with respect to the leaves of ``state``."""
        final, grad = self.vjp(state, steps, None, checkpoint_every)
        d = PlaneCouetteMHDHIP.diagnostics(self)
        return d["Epert"] + d["Emag"], grad, final

    def _mhd_gradient(self) -> MHDState:
        flow = self._gradient()
        groups = {}
        for name, tag in _A_GROUPS:
            out = []
            for i, c in enumerate("XYZ"):
                arr = np.empty((self._a_rows(i), self.N[1], self.N[2]), np.complex128)
                field = f"A{c}" if not tag else f"{tag}{c}"
                nat.check(self._lib.pcf_get_gradient(self._h, nat.FIELD_ID[field],
                                                     arr.ctypes.data_as(C.POINTER(C.c_double))))
                out.append(arr)
            groups[name] = tuple(out)
        return MHDState(flow=flow, **groups)

    @staticmethod
    def benchmark_state_fields(state: MHDState):
        """``benchmark_state_fields`` (``pcf_mhd_jax.py:754-758``)."""
        return (*state.flow.u, state.flow.g, *state.A)


class PlaneCouetteMRIShearpyHIP(PlaneCouetteMHDHIP):
    """Rotating-shear plane-Couette MHD with conducting walls.

    Mirrors ``PlaneCouetteMRIShearpyJax``
    (``examples/pcf_mhd_mri_shearpy_jax.py:30-514``): base flow U_b = -S x,
    Coriolis force of rotation rate ``omega``, uniform background field
    ``background_b``, A in TD^3 as in ``PlaneCouetteMHDHIP``.  Viscosity 1/Re
    and magnetic diffusivity 1/Rm (U_wall = 1 as in the reference).
    """

    def __init__(self, N=(17, 16, 16), domain=((-1.0, 1.0), (0.0, 4.0 * math.pi), (0.0, 2.0 * math.pi)),
                 Re: float = 400.0, Rm: float | None = None, omega: float = 1.0, shear_rate: float = 1.0,
                 background_b=(0.0, 0.0, 0.025), dt: float = 0.01, family: str = "C",
                 padding_factor=(1.0, 1.5, 1.5), perturbation_amplitude: float = 0.05,
                 magnetic_amplitude: float = 0.0, magnetic_seed: str = "ax_yz",
                 solenoidal_velocity_seed: bool = False, time_integrator: str | None = None,
                 nonlinear_form: str = "gradient", coefficient_path: str = "transform",
                 solve_batching: str = "batched", precision: str = "fp64", device: int = 0) -> None:
        self.omega = float(omega)
        self.shear_rate = float(shear_rate)
        self.magnetic_seed = str(magnetic_seed)
        self.solenoidal_velocity_seed = bool(solenoidal_velocity_seed)
        if self.magnetic_seed not in {"ax_yz", "sinusoidal_bz_x"}:
            raise ValueError("magnetic_seed must be one of {ax_yz, sinusoidal_bz_x}")
        if tuple(float(v) for v in domain[0]) != (-1.0, 1.0):
            raise ValueError("wall-normal domain must be (-1, 1)")
        super().__init__(N=N, domain=domain, Re=Re, Rm=Rm, U_wall=1.0, dt=dt, family=family,
                         padding_factor=padding_factor, perturbation_amplitude=perturbation_amplitude,
                         magnetic_amplitude=magnetic_amplitude, time_integrator=time_integrator,
                         nonlinear_form=nonlinear_form, coefficient_path=coefficient_path,
                         solve_batching=solve_batching, background_b=background_b, precision=precision,
                         device=device)
        self.q_shear = self.shear_rate / self.omega if self.omega != 0 else math.inf
        self.kappa2 = 2.0 * self.omega * (2.0 * self.omega - self.shear_rate)

    def _configure(self, cfg) -> None:
        super()._configure(cfg)
        cfg.omega = self.omega
        cfg.u_wall = -self.shear_rate  # base flow U_b = -S x (shearpy :89-92)

    def _a_projection(self, comp: int) -> np.ndarray:
        """Nodal values -> coefficients of A component comp (Galerkin projection onto TD)."""
        return self.matrix("FD")

    def initial_state(self) -> MHDState:
        """``initial_state`` (``examples/pcf_mhd_mri_shearpy_jax.py:99-147``): the
        channel-mode velocity seed and the ``ax_yz`` or ``sinusoidal_bz_x`` field,
        projected with the library's Galerkin matrices (every term separable)."""
        n0, n1, n2 = self.N
        xn = np.cos(np.pi + (2 * np.arange(n0) + 1) * np.pi / (2 * n0))
        wall = 1.0 - xn**2
        FB, FD = self.matrix("FB"), self.matrix("FD")
        amp = self.perturbation_amplitude

        def four(kind_y, my, kind_z, mz):
            cy = np.zeros(n1, complex)
            cz = np.zeros(n2, complex)
            for c, kind, mm in ((cy, kind_y, my), (cz, kind_z, mz)):
                if kind == "one":
                    c[0] = 1.0
                elif kind == "sin":
                    c[mm], c[-mm] = 1 / 2j, -1 / 2j
                else:
                    c[mm], c[-mm] = 0.5, 0.5
            return np.outer(cy, cz)

        def proj(F, terms):
            out = 0
            for prof, fc in terms:
                out = out + np.einsum("i,jk->ijk", F @ prof, fc)
            return self.mask_nyquist(out)

        t0, t1, t2 = [], [], []
        kz1 = 2.0 * math.pi / self.Lz
        if amp != 0.0:
            if self.solenoidal_velocity_seed:
                wall_gradient = -4.0 * xn * (1.0 - xn**2)
                for hm in (1, 2, 3):
                    t0.append((amp * wall**2, four("one", 0, "cos", hm)))
                    t1.append((amp * wall**2, four("one", 0, "sin", hm)))
                    t2.append((-(amp * wall_gradient / (hm * kz1)), four("one", 0, "sin", hm)))
            else:
                for hm in (1, 2, 3):
                    t0.append((amp * wall, four("one", 0, "cos", hm)))
                    t1.append((amp * wall, four("one", 0, "sin", hm)))
                t0.append((0.1 * amp * wall, four("sin", 1, "cos", 1)))
                t1.append((0.1 * amp * wall, four("cos", 1, "sin", 1)))
                t2.append((0.1 * amp * wall, four("sin", 2, "cos", 2)))
        z = np.zeros((n0, n1, n2), complex)
        u0 = proj(FB, t0) if t0 else z[:-4].copy()
        u1 = proj(FD, t1) if t1 else z[:-2].copy()
        u2 = proj(FD, t2) if t2 else z[:-2].copy()
        ky, kz = self.wavenumbers()
        g = self.mask_nyquist(1j * ky[None] * u2 - 1j * kz[None] * u1)
        flow = KMMState(u=(u0, u1, u2), g=g)
        A = list(self._zero_A())
        mag = self.magnetic_amplitude
        if mag != 0.0 and self.magnetic_seed == "ax_yz":
            A[0] = proj(self._a_projection(0), [(mag * wall / kz1, four("sin", 1, "sin", 1))])
        elif mag != 0.0:
            # zero-net-flux vertical field: Ay = -(mag / pi)(cos(pi x) + 1), zero at both walls
            prof = -(mag / np.pi) * (np.cos(np.pi * xn) + 1.0)
            A[1] = proj(self._a_projection(1), [(prof, four("one", 0, "one", 0))])
        return self._new_state(flow, A)

    def diagnostics(self, state: MHDState | None = None) -> dict:
        """``PlaneCouetteMRIShearpyJax.diagnostics``
        (``examples/pcf_mhd_mri_shearpy_jax.py:431-513``): the flow and
        magnetic diagnostics, the electromagnetic health set (``:355-429``),
        Reynolds and Maxwell stresses and alpha, the mean-field split, viscous
        and resistive dissipation (``:178-201``), the kz = 1 channel amplitudes
        (``:203-215``) and the rotation numbers (``:94-97``), in the
        reference's key order.  The reductions run on the device
        (``pcf_mhd_extended_diagnostics``); the derived keys are scalar
        arithmetic on them, as in the reference."""
        base = super().diagnostics(state)
        x = np.zeros(nat.XDIAG_COUNT)
        nat.check(self._lib.pcf_mhd_extended_diagnostics(self._h, x.ctypes.data_as(C.POINTER(C.c_double))))
        d = dict(zip(nat.XDIAG_KEYS, (float(v) for v in x)))
        volume = 2.0 * self.Ly * self.Lz
        mean_b = (d["mean_bx"], d["mean_by"], d["mean_bz"])
        trace = (d["mean_bx_trace"], d["mean_by_trace"], d["mean_bz_trace"])
        em = {k: d[k] for k in ("electric_ideal_l2", "electric_resistive_l2", "electric_total_l2",
                                "divergence_e_l2", "divergence_e_ideal_l2", "divergence_e_resistive_l2",
                                "divergence_a_l2", "divergence_a_wall_linf",
                                "electric_ideal_wall_tangential_linf", "electric_resistive_wall_tangential_linf",
                                "electric_wall_tangential_linf")}
        em["faraday_mean_bx_tendency"] = 0.0
        em["faraday_mean_by_tendency"] = d["faraday_mean_by_tendency"]
        em["faraday_mean_bz_tendency"] = d["faraday_mean_bz_tendency"]
        em.update({k: d[k] for k in ("mean_bx_trace", "mean_by_trace", "mean_bz_trace")})
        em["mean_b_trace_mismatch_linf"] = max(abs(a - b) for a, b in zip(mean_b, trace))
        for label in ("ideal", "resistive", "total"):
            for axis in "xyz":
                em[f"mean_e{axis}_{label}"] = d[f"mean_e{axis}_{label}"]
        reynolds, maxwell = d["reynolds_stress"], d["maxwell_stress"]
        transport = reynolds + maxwell
        sh2 = self.shear_rate**2  # (S x_half_width)^2 with x in (-1, 1)
        mag_energy_mean = volume * sum(v * v for v in mean_b)
        kz1_u, kz1_b = d["channel_kz1_velocity_rms"], d["channel_kz1_magnetic_rms"]
        out = {
            **base,
            **em,
            "Emag_total": d["Emag_total"],
            "bmax": d["bmax"],
            "bmax_total": d["bmax_total"],
            "reynolds_xy": reynolds,
            "maxwell_xy": maxwell,
            "reynolds_stress": reynolds,
            "maxwell_stress": maxwell,
            "transport_xy": transport,
            "total_stress": transport,
            "alpha_Sh": transport / sh2 if sh2 > 0.0 else math.nan,
            "mean_bx": mean_b[0],
            "mean_by": mean_b[1],
            "mean_bz": mean_b[2],
            "mag_energy_fluct_total": d["mag_energy_fluct_total"],
            "mag_energy_mean": mag_energy_mean,
            "mag_energy_fluct": d["Emag_total"] - mag_energy_mean,
            "dissipation_kinetic": d["dissipation_kinetic"],
            "dissipation_magnetic": d["dissipation_magnetic"],
            "q_shear": self.q_shear,
            "kappa2": self.kappa2,
            "channel_kz1_velocity_rms": kz1_u,
            "channel_kz1_magnetic_rms": kz1_b,
            "channel_kz1_total_rms": math.sqrt(kz1_u**2 + kz1_b**2),
        }
        vA2 = self.background_b[2] ** 2
        if vA2 > 0.0:
            out["alpha"] = out["alpha_B0"] = transport / vA2
        return out


class PlaneCouetteMRIShearpyInsulatingHIP(PlaneCouetteMRIShearpyHIP):
    """Rotating-shear plane-Couette MHD with insulating (vacuum-matched) walls.

    Mirrors ``PlaneCouetteMRIShearpyInsulatingJax``
    (``examples/pcf_mhd_mri_shearpy_jax.py:516-836``): as
    ``PlaneCouetteMRIShearpyHIP``, with A_x in TD and A_y, A_z in the orthogonal
    Chebyshev space TC with per-mode vacuum-matching wall rows.
    """

    def _configure(self, cfg) -> None:
        super()._configure(cfg)
        cfg.walls = nat.WALLS_INSULATING

    def _a_rows(self, comp: int) -> int:
        return self.nd if comp == 0 else self.N[0]

    def vjp(self, *args, **kwargs):
        raise NotImplementedError("the MHD reverse mode covers conducting walls (A in TD^3) only")

    def value_and_grad(self, *args, **kwargs):
        raise NotImplementedError("the MHD reverse mode covers conducting walls (A in TD^3) only")

    def _a_projection(self, comp: int) -> np.ndarray:
        """TD projection for A_x; TC coefficients (scalar products / h) for A_y, A_z."""
        if comp == 0:
            return self.matrix("FD")
        h = np.full(self.N[0], np.pi / 2)
        h[0] = np.pi
        return np.diag(1.0 / h) @ self.matrix("WQ")

    def diagnostics(self, state: MHDState | None = None) -> dict:
        """Adds ``insulating_bc_residual``
        (``examples/pcf_mhd_mri_shearpy_jax.py:802-836``)."""
        out = super().diagnostics(state)
        out["insulating_bc_residual"] = float(self._mhd_diag[3])
        return out


def mhd_state_leaves(state: MHDState) -> list:
    """Leaves in the reference pytree order (``pcf_mhd_jax.py:46-55``)."""
    out = state_leaves(state.flow)
    for name, _ in _A_GROUPS:
        out.extend(getattr(state, name))
    return out


def mhd_state_checksum(state: MHDState) -> dict:
    """Full-state checksum as ``production/benchmark.py:234-260`` computes it."""
    real_sum = imag_sum = norm2 = 0.0
    count = 0
    leaves = mhd_state_leaves(state)
    for v in leaves:
        v = np.asarray(v)
        real_sum += float(np.sum(np.real(v), dtype=np.float64))
        imag_sum += float(np.sum(np.imag(v), dtype=np.float64))
        norm2 += float(np.sum(np.abs(v) ** 2, dtype=np.float64))
        count += int(v.size)
    return {"leaf_count": len(leaves), "element_count": count, "real_sum": real_sum,
            "imag_sum": imag_sum, "l2_norm": math.sqrt(norm2)}
