"""Plane-Couette fluctuation DNS on the MI355X: drop-in for the reference solver.

``PlaneCouetteFluctuationHIP`` mirrors ``PlaneCouetteFluctuationJax``
(reference ``examples/pcf_fluctuations_jax.py:29-161`` over
``examples/channelflow_kmm.py:104-1308``): the same constructor keywords, the
same ``KMMState`` fields in the same full-complex layout, and the same
``initial_state / step / solve / set_dt / diagnostics`` methods, so a runner
that builds the reference solver from a problem spec
(``production/oracles.py:1442-1480``) and drives ``solver.solve(state, steps)``
(``production/oracles.py:3110-3123``) can construct this class instead.

Every step runs in libpcfdns.so (hand-written HIP kernels); this module only
moves state across the boundary and builds initial conditions from the
library's closed-form setup matrices.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, replace

import numpy as np

from . import _native as nat

_PRECISIONS = {"fp64": nat.PREC_FP64, "float64": nat.PREC_FP64, "mixed": nat.PREC_MIXED}


def _is_torch(x) -> bool:
    return type(x).__module__.split(".")[0] == "torch"


def _real(x):
    if _is_torch(x):
        return x.real if x.is_complex() else x
    return np.real(x)


def _leaves(state) -> list:
    """Array leaves of a (possibly nested) state dataclass."""
    out = []
    for v in vars(state).values():
        if isinstance(v, (tuple, list)):
            out.extend(v)
        elif hasattr(v, "__dataclass_fields__"):
            out.extend(_leaves(v))
        elif v is not None and not isinstance(v, (int, float)):
            out.append(v)
    return out


def to_host(state):
    """A copy of a state (KMMState or MHDState) whose torch tensor leaves are
    numpy arrays on the host; numpy leaves are kept as they are."""

    def conv(v):
        if _is_torch(v):
            return v.detach().cpu().numpy()
        if isinstance(v, tuple):
            return tuple(conv(x) for x in v)
        if hasattr(v, "__dataclass_fields__"):
            return to_host(v)
        return v

    return replace(state, **{k: conv(v) for k, v in vars(state).items()})


@dataclass(frozen=True)
class RolloutCacheInfo:
    """Graph-cache counters in the fields of the reference's
    ``ScanRolloutCacheInfo`` (``src/jaxfun/integrators/cnab2.py:23-33``)."""

    generation: int
    max_entries: int
    live_entries: int
    step_counts: tuple
    hits: int
    misses: int
    evictions: int


@dataclass(frozen=True)
class KMMState:
    """Coefficient-space KMM state (reference ``channelflow_kmm.py:50-101``).

    ``u = (u0, u1, u2)`` with u0 in the clamped space TB, u1/u2 and g in the
    Dirichlet space TD, all complex128 of shape (n, Ny, Nz) in FFT order.
    History tuples are ``(Nu, Ng, Nv00, Nw00)`` and ``(u0, g, v00, w00)``.
    """

    u: tuple
    g: np.ndarray
    nonlinear_old: tuple | None = None
    nonlinear_older: tuple | None = None
    solution_old: tuple | None = None
    solution_older: tuple | None = None
    history_steps: float = 0.0
    have_old: float = 0.0
    previous_dt: float = 0.0


class PlaneCouetteFluctuationHIP:
    """KMM SBDF3 plane-Couette solver whose step runs in HIP on one GPU.

    Constructor follows ``PlaneCouetteFluctuationJax.__init__``
    (``examples/pcf_fluctuations_jax.py:35-75``), defaults included.  Built
    options: Chebyshev family; SBDF3 with IMEXRK3 startup, IMEXRK3 alone,
    IMEXRK222 or CNAB2 (``time_integrator=None`` selects IMEXRK222 as in the
    reference); the rotational nonlinear form; both coefficient paths (they
    form identical rows).  The reference's default ``nonlinear_form="gradient"``
    and the Legendre family raise ``NotImplementedError``; invalid values raise
    ``ValueError`` as in ``examples/channelflow_kmm.py:140-187``.
    """

    def __init__(
        self,
        N=(17, 16, 16),
        domain=((-1.0, 1.0), (0.0, 4.0 * math.pi), (0.0, 2.0 * math.pi)),
        Re: float = 600.0,
        U_wall: float = 1.0,
        dt: float = 0.01,
        family: str = "C",
        padding_factor=(1.0, 1.5, 1.5),
        perturbation_amplitude: float = 0.05,
        timestepper=None,
        time_integrator: str | None = None,
        nonlinear_form: str = "gradient",
        coefficient_path: str = "transform",
        solve_batching: str = "batched",
        precision: str = "fp64",
        device: int = 0,
    ) -> None:
        nonlinear_form = str(nonlinear_form).lower()
        if nonlinear_form not in {"gradient", "rotational"}:
            raise ValueError("nonlinear_form must be 'gradient' or 'rotational'")
        coefficient_path = str(coefficient_path).lower()
        if coefficient_path not in {"optimized", "transform"}:
            raise ValueError("coefficient_path must be 'optimized' or 'transform'")
        solve_batching = str(solve_batching).lower()
        if solve_batching not in {"batched", "separate"}:
            raise ValueError("solve_batching must be 'batched' or 'separate'")
        if time_integrator is not None and time_integrator not in {"IMEXRK222", "IMEXRK3", "CNAB2", "SBDF3"}:
            raise ValueError("time_integrator must be one of {'IMEXRK222', 'IMEXRK3', 'CNAB2', 'SBDF3'}")
        fam = str(family).upper()
        if not (fam.startswith("C") or fam.startswith("L")):
            raise ValueError("family must be 'L' or 'C'")
        if fam.startswith("L"):
            raise NotImplementedError("only the Chebyshev family is built")
        if time_integrator is None:  # the reference's default ARS stepper (channelflow_kmm.py:160-172)
            time_integrator = "IMEXRK222"
        if nonlinear_form != "rotational":
            raise NotImplementedError("only the rotational nonlinear form is built; pass "
                                      "nonlinear_form='rotational'")
        # coefficient_path "transform" (channelflow_kmm.py:785-803,831-847) forms the
        # same rows as "optimized" (:810-829,855-870): every quadrature on the
        # unpadded grid is exact for these degrees, so both run the same kernels.
        if str(precision).lower() not in _PRECISIONS:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if tuple(float(v) for v in domain[0]) != (-1.0, 1.0):
            raise ValueError("wall-normal domain must be (-1, 1)")

        self.N = tuple(int(n) for n in N)
        self.domain = domain
        self.Re = float(Re)
        self.U_wall = float(U_wall)
        self.nu = self.U_wall / self.Re
        self.dt = float(dt)
        self.family = "C"
        self.padding_factor = tuple(float(p) for p in padding_factor)
        self.perturbation_amplitude = float(perturbation_amplitude)
        self.time_integrator = time_integrator
        self.nonlinear_form = nonlinear_form
        self.coefficient_path = coefficient_path
        self.solve_batching = solve_batching
        self.precision = str(precision).lower()
        self.Ly = float(domain[1][1] - domain[1][0])
        self.Lz = float(domain[2][1] - domain[2][0])

        cfg = nat.PcfConfig()
        cfg.abi_version = nat.ABI_VERSION
        cfg.n0, cfg.ny, cfg.nz = self.N
        cfg.lx[0], cfg.lx[1] = -1.0, 1.0
        cfg.ly, cfg.lz = self.Ly, self.Lz
        cfg.nu = self.nu
        cfg.u_wall = self.U_wall
        cfg.dt = self.dt
        cfg.padding_y, cfg.padding_z = self.padding_factor[1], self.padding_factor[2]
        cfg.precision = _PRECISIONS[self.precision]
        cfg.time_integrator = nat.TIME_INTEGRATORS[self.time_integrator]
        cfg.nonlinear_form = nat.NL_ROTATIONAL
        cfg.device = int(device)
        self._configure(cfg)
        self._cfg = cfg
        self._lib = nat.load()
        handle = C.c_void_p()
        nat.check(self._lib.pcf_create(C.byref(cfg), C.byref(handle)))
        self._h = handle
        self._loaded = None  # the state object whose values are resident on the device
        self._pending = []  # device leaves whose copies the library stream has not finished
        self._fetch_on_device = False

    def _configure(self, cfg) -> None:
        """Hook for subclasses to fill the rest of the library configuration."""

    # ------------------------------------------------------------ lifecycle
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.pcf_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ geometry
    @property
    def nb(self) -> int:
        return self.N[0] - 4

    @property
    def nd(self) -> int:
        return self.N[0] - 2

    def matrix(self, name: str) -> np.ndarray:
        return nat.setup_matrix(self._cfg, name)

    def wavenumbers(self):
        """Scaled (ky, kz) grids in FFT order (``tensorproductspace.py:253-271``)."""
        def k(n, L):
            idx = np.arange(n)
            kk = np.where(idx < (n + 1) // 2, idx, idx - n).astype(float)
            df = 2 * math.pi / L if abs(L - 2 * math.pi) > 1e-12 else 1.0
            return kk * df
        return k(self.N[1], self.Ly)[:, None], k(self.N[2], self.Lz)[None, :]

    def mask_nyquist(self, c: np.ndarray) -> np.ndarray:
        c = np.array(c, copy=True)
        c[..., self.N[1] // 2, :] = 0
        c[..., :, self.N[2] // 2] = 0
        return c

    # ------------------------------------------------------------ states
    def _with_zero_history(self, u, g) -> KMMState:
        """``_reset_flow_history`` (``channelflow_kmm.py:637-652``)."""
        zu = np.zeros(tuple(u[0].shape), complex)
        zg = np.zeros(tuple(g.shape), complex)
        zv = np.zeros(self.nd)
        z = (zu, zg, zv, zv)
        return KMMState(u=tuple(u), g=g, nonlinear_old=z, nonlinear_older=z, solution_old=z,
                        solution_older=z, history_steps=0.0, have_old=0.0, previous_dt=0.0)

    def zero_state(self) -> KMMState:
        n1, n2 = self.N[1], self.N[2]
        u = (np.zeros((self.nb, n1, n2), complex), np.zeros((self.nd, n1, n2), complex),
             np.zeros((self.nd, n1, n2), complex))
        return self._with_zero_history(u, np.zeros((self.nd, n1, n2), complex))

    def state_from_separable(self, profiles, fourier):
        """Initial state from separable physical fields (host setup, closed form).

        ``profiles[i]`` are wall-normal nodal values and ``fourier[i]`` the
        (Ny, Nz) full-complex Fourier coefficients of component i; the result
        equals ``state_from_physical`` (``channelflow_kmm.py:680-686``) of the
        product fields, because the Galerkin projection acts on the wall-normal
        factor only.
        """
        FB, FD = self.matrix("FB"), self.matrix("FD")
        u0 = self.mask_nyquist(np.einsum("i,jk->ijk", FB @ profiles[0], fourier[0]))
        u1 = self.mask_nyquist(np.einsum("i,jk->ijk", FD @ profiles[1], fourier[1]))
        u2 = self.mask_nyquist(np.einsum("i,jk->ijk", FD @ profiles[2], fourier[2]))
        ky, kz = self.wavenumbers()
        g = self.mask_nyquist(1j * ky[None] * u2 - 1j * kz[None] * u1)
        return self._with_zero_history((u0, u1, u2), g)

    def initial_state(self) -> KMMState:
        """Analytic fluctuation IC (``pcf_fluctuations_jax.py:77-97``).

        u0 = a (1-x^2) sin(2 pi y/Ly) cos(2 pi z/Lz), u1 = a (1-x^2) cos(.) sin(.),
        u2 = a (1-x^2) sin(4 pi y/Ly) cos(4 pi z/Lz).  The Fourier factors have
        exact two-term expansions; on the reference's discrete grid
        (fft(norm='forward')) they are the coefficients below.
        """
        a = self.perturbation_amplitude
        x = self.matrix("VB")  # only for the node count; nodes come from the library
        n0, n1, n2 = self.N
        xn = np.cos(np.pi + (2 * np.arange(n0) + 1) * np.pi / (2 * n0))
        prof = a * (1.0 - xn**2)

        def coeffs(kind_y, my, kind_z, mz):
            # sin(2 pi m t/L) -> (e^{i m} - e^{-i m}) / 2i ; cos -> (e^{i m} + e^{-i m}) / 2
            cy = np.zeros(n1, complex)
            cz = np.zeros(n2, complex)
            if kind_y == "sin":
                cy[my], cy[-my] = 1 / 2j, -1 / 2j
            else:
                cy[my], cy[-my] = 0.5, 0.5
            if kind_z == "sin":
                cz[mz], cz[-mz] = 1 / 2j, -1 / 2j
            else:
                cz[mz], cz[-mz] = 0.5, 0.5
            return np.outer(cy, cz)

        del x
        f0 = coeffs("sin", 1, "cos", 1)
        f1 = coeffs("cos", 1, "sin", 1)
        f2 = coeffs("sin", 2, "cos", 2)
        return self.state_from_separable((prof, prof, prof), (f0, f1, f2))

    # ------------------------------------------------------------ device I/O
    def _set(self, name, arr):
        if _is_torch(arr):
            return self._set_device(name, arr)
        arr = np.ascontiguousarray(arr)
        if np.iscomplexobj(arr):
            arr = arr.astype(np.complex128, copy=False)
        else:
            arr = arr.astype(np.float64, copy=False)
        nat.check(self._lib.pcf_set_field(self._h, nat.FIELD_ID[name],
                                          arr.ctypes.data_as(C.POINTER(C.c_double))))

    def _torch_device(self):
        import torch

        return torch.device("cuda", int(self._cfg.device))

    def _set_device(self, name, t):
        """A torch tensor leaf straight from device memory (pcf_set_field_device):
        cast to complex128 / float64 on this solver's GPU; kept alive until the
        library stream has read it."""
        import torch

        t = t.to(device=self._torch_device(), dtype=torch.complex128 if t.is_complex() else torch.float64)
        t = t.contiguous()
        self._pending.append(t)
        nat.check(self._lib.pcf_set_field_device(self._h, nat.FIELD_ID[name], C.c_void_p(t.data_ptr())))

    def _get(self, name, rows, vector=False):
        if self._fetch_on_device:
            import torch

            shape = (rows,) if vector else (rows, self.N[1], self.N[2])
            out = torch.empty(shape, dtype=torch.float64 if vector else torch.complex128,
                              device=self._torch_device())
            nat.check(self._lib.pcf_get_field_device(self._h, nat.FIELD_ID[name], C.c_void_p(out.data_ptr())))
            return out
        if vector:
            out = np.empty(rows, np.float64)
        else:
            out = np.empty((rows, self.N[1], self.N[2]), np.complex128)
        nat.check(self._lib.pcf_get_field(self._h, nat.FIELD_ID[name],
                                          out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def _device_history(self, state: KMMState) -> KMMState:
        """The history the step reads, as ``_ensure_flow_history``
        (``channelflow_kmm.py:612-635``) leaves it: SBDF3 restarts unless all
        four levels are present; CNAB2 without ``nonlinear_old`` starts with
        zero rows, ``have_old = 0``, ``previous_dt = 0``.  Levels the step does
        not read travel as zeros."""
        groups = ("nonlinear_old", "nonlinear_older", "solution_old", "solution_older")
        if self.time_integrator == "SBDF3":
            if any(getattr(state, k) is None for k in groups):
                return self._with_zero_history(state.u, state.g)
            return state
        z = self._with_zero_history(state.u, state.g)
        if self.time_integrator == "CNAB2" and state.nonlinear_old is None:
            state = replace(state, have_old=0.0, previous_dt=0.0)
        return replace(state, **{k: getattr(z, k) for k in groups if getattr(state, k) is None})

    def load(self, state: KMMState) -> None:
        """Copy a state (with its history) to the device.  Leaves may be numpy
        arrays (host copies) or torch tensors (device-to-device copies on the
        library stream, after the producing torch stream has finished)."""
        state = self._device_history(state)
        self._begin_load(state)
        self._set("U0", state.u[0])
        self._set("U1", state.u[1])
        self._set("U2", state.u[2])
        self._set("G", state.g)
        for tag, grp in (("NL_OLD", state.nonlinear_old), ("NL_OLDER", state.nonlinear_older),
                         ("SOL_OLD", state.solution_old), ("SOL_OLDER", state.solution_older)):
            for comp, arr in zip("UGVW", grp):
                self._set(f"{tag}_{comp}", _real(arr) if comp in "VW" else arr)
        nat.check(self._lib.pcf_set_history(self._h, float(state.history_steps),
                                            float(state.have_old), float(state.previous_dt)))
        self._end_load()
        self._loaded = None

    def _begin_load(self, state) -> None:
        if any(_is_torch(leaf) for leaf in _leaves(state)):
            import torch

            torch.cuda.current_stream(self._torch_device()).synchronize()

    def _end_load(self) -> None:
        if self._pending:
            nat.check(self._lib.pcf_synchronize(self._h))
            self._pending.clear()

    def fetch(self) -> KMMState:
        """Copy the device state back (full-complex reference layout)."""
        nb, nd = self.nb, self.nd
        u = (self._get("U0", nb), self._get("U1", nd), self._get("U2", nd))
        g = self._get("G", nd)
        groups = []
        for tag in ("NL_OLD", "NL_OLDER", "SOL_OLD", "SOL_OLDER"):
            groups.append((self._get(f"{tag}_U", nb), self._get(f"{tag}_G", nd),
                           self._get(f"{tag}_V", nd, True), self._get(f"{tag}_W", nd, True)))
        hs, ho, pd = C.c_double(), C.c_double(), C.c_double()
        nat.check(self._lib.pcf_get_history(self._h, C.byref(hs), C.byref(ho), C.byref(pd)))
        st = KMMState(u=u, g=g, nonlinear_old=groups[0], nonlinear_older=groups[1],
                      solution_old=groups[2], solution_older=groups[3],
                      history_steps=hs.value, have_old=ho.value, previous_dt=pd.value)
        self._loaded = st
        return st

    def fetch_device(self):
        """The device state as a state whose leaves are torch CUDA tensors
        (complex128 fields, float64 mean-mode vectors) copied on the device
        (pcf_get_field_device); ready on every torch stream on return."""
        self._fetch_on_device = True
        try:
            st = self.fetch()
        finally:
            self._fetch_on_device = False
        nat.check(self._lib.pcf_synchronize(self._h))
        return st

    def _ensure_loaded(self, state: KMMState) -> None:
        if self._loaded is not state:
            self.load(state)

    def advance(self, steps: int) -> None:
        """Advance the device-resident state (no host transfer)."""
        self._loaded = None
        nat.check(self._lib.pcf_step(self._h, int(steps)))

    def advance_steady(self, steps: int, graph_block: int = 0) -> None:
        """Exactly ``steps`` steady SBDF3 steps.  With ``graph_block`` they replay
        HIP graphs: whole blocks of ``graph_block`` steps, then one graph of the
        remainder (graphs are cached per ring phase and length)."""
        self._loaded = None
        steps = int(steps)
        if graph_block:
            q, r = divmod(steps, int(graph_block))
            if q:
                nat.check(self._lib.pcf_step_steady_graph(self._h, int(graph_block), q))
            if r:
                nat.check(self._lib.pcf_step_steady_graph(self._h, r, 1))
        else:
            nat.check(self._lib.pcf_step_steady(self._h, steps))

    def prepare_steady(self, steps: int, graph_block: int) -> None:
        """Capture, without running anything, the graphs that
        ``advance_steady(steps, graph_block)`` replays from the current ring
        phase; with ``graph_block`` a multiple of 12 the remainder graph starts
        at the same phase, so the whole run then replays cached graphs."""
        q, r = divmod(int(steps), int(graph_block))
        if graph_block % 12:
            raise ValueError("graph_block must be a multiple of 12 steps")
        if q:
            nat.check(self._lib.pcf_step_steady_graph(self._h, int(graph_block), 0))
        if r:
            nat.check(self._lib.pcf_step_steady_graph(self._h, r, 0))

    def synchronize(self) -> None:
        nat.check(self._lib.pcf_synchronize(self._h))

    def rollout_cache_info(self) -> RolloutCacheInfo:
        """``KMM.rollout_cache_info`` (``channelflow_kmm.py:1255-1257``): the
        counters of the handle's HIP-graph caches in the fields of the
        reference's ``ScanRolloutCacheInfo`` (``src/jaxfun/integrators/cnab2.py:23-33``).
        The caches keep every graph until the stream changes, so
        ``max_entries`` is -1 (unbounded) and ``evictions`` counts the graphs
        dropped then."""
        out = (C.c_int64 * 69)()
        n = nat.check(self._lib.pcf_graph_cache_info(self._h, out, 69))
        return RolloutCacheInfo(generation=int(out[0]), max_entries=-1, live_entries=int(out[1]),
                                step_counts=tuple(int(v) for v in out[5:5 + min(n, 64)]), hits=int(out[2]),
                                misses=int(out[3]), evictions=int(out[4]))

    def solve_with_cadence(self, state, steps: int, cadence, *, block_size: int = 1, on_diagnostics=None,
                           on_snapshot=None, on_checkpoint=None, should_stop=None, t0: float = 0.0,
                           tstep0: int = 0):
        """``KMM.solve_with_cadence`` (``channelflow_kmm.py:1264-1291``) with the
        loop of ``jaxfun.io.run_with_cadence`` (``src/jaxfun/io/__init__.py:117-186``):
        blocks of at most ``block_size`` steps that end on every step a
        cadence of ``cadence`` (``diagnostics_every``, ``snapshot_every``,
        ``checkpoint_every``) falls due; after each block the due callbacks
        in that order, then ``should_stop(t, tstep, state)``, which may end
        the run early."""
        if steps < 0:
            raise ValueError("steps must be non-negative")
        if block_size < 1:
            raise ValueError("block_size must be positive")
        every = {k: getattr(cadence, k, None) for k in ("diagnostics_every", "snapshot_every", "checkpoint_every")}
        periods = [int(e) for e in every.values() if e is not None and e > 0]

        def due(tstep, e):
            return e is not None and e > 0 and tstep % e == 0

        out = state
        tstep = int(tstep0)
        final = tstep + int(steps)
        while tstep < final:
            n = min([block_size, final - tstep] + [max(1, (tstep // e + 1) * e - tstep) for e in periods])
            out = self.solve(out, n)
            tstep += n
            t = float(t0) + (tstep - int(tstep0)) * float(self.dt)
            if due(tstep, every["diagnostics_every"]) and on_diagnostics:
                on_diagnostics(t, tstep, self.diagnostics(out))
            if due(tstep, every["snapshot_every"]) and on_snapshot:
                on_snapshot(t, tstep, out)
            if due(tstep, every["checkpoint_every"]) and on_checkpoint:
                on_checkpoint(t, tstep, out)
            if should_stop is not None and bool(should_stop(t, tstep, out)):
                break
        return out

    # ------------------------------------------------------------ reference API
    def step(self, state: KMMState, dt=None, factor_args=None) -> KMMState:
        """One step (``channelflow_kmm.py:1123-1137``); state in, new state out.

        A ``dt`` other than the solver's runs this one step at ``dt`` with its
        implicit systems factorised at ``dt`` (the reference pairs a runtime dt
        with ``factor_args`` built for it); SBDF3 refuses it."""
        if dt is not None and not math.isclose(float(dt), self.dt, rel_tol=8 * np.finfo(float).eps):
            keep = self.dt
            self.set_dt(float(dt))
            try:
                return self.step(state)
            finally:
                self.set_dt(keep)
        self._ensure_loaded(state)
        self.advance(1)
        return self._result(state)

    def solve(self, state: KMMState, steps: int) -> KMMState:
        """``KMM.solve`` (``channelflow_kmm.py:1231-1247``)."""
        self._ensure_loaded(state)
        self.advance(int(steps))
        return self._result(state)

    def solve_device(self, state: KMMState, steps: int) -> KMMState:
        """``solve`` with the result's leaves as torch CUDA tensors: the state
        stays in device memory (the reference's jax arrays do too), and
        passing the result back in reuses what is resident."""
        self._ensure_loaded(state)
        self.advance(int(steps))
        return self._result(state, device=True)

    def _result(self, state: KMMState, device: bool = False) -> KMMState:
        """The device state after steps from ``state``, with the history fields
        each integrator leaves: SBDF3 its own rings; IMEXRK3 those of ``state``
        unchanged (``replace``, ``channelflow_kmm.py:935``); IMEXRK222 only
        ``nonlinear_old``, ``have_old`` and ``previous_dt`` of ``state``
        (``:1194-1200``), the device's history slots being scratch there;
        CNAB2 the step's nonlinear rows as ``nonlinear_old`` with ``have_old``
        and ``previous_dt`` (``:1112-1117``)."""
        st = self.fetch_device() if device else self.fetch()
        if self.time_integrator == "IMEXRK3":
            st = replace(state, u=st.u, g=st.g)
        elif self.time_integrator == "IMEXRK222":
            st = KMMState(u=st.u, g=st.g, nonlinear_old=state.nonlinear_old, have_old=state.have_old,
                          previous_dt=state.previous_dt)
        elif self.time_integrator == "CNAB2":
            st = KMMState(u=st.u, g=st.g, nonlinear_old=st.nonlinear_old, have_old=st.have_old,
                          previous_dt=st.previous_dt)
        self._loaded = st
        return st

    def set_dt(self, dt: float) -> None:
        """``KMM.set_dt`` (``channelflow_kmm.py:1207-1222``): SBDF3 is fixed-step
        (``NotImplementedError``); the other integrators take the new dt for
        the following steps."""
        nat.check(self._lib.pcf_set_dt(self._h, float(dt)))
        if self.time_integrator != "SBDF3":
            self.dt = float(dt)

    def diagnostics(self, state: KMMState | None = None) -> dict:
        """``PlaneCouetteFluctuationJax.diagnostics`` (``pcf_fluctuations_jax.py:131-161``)."""
        if state is not None:
            self._ensure_loaded(state)
        out = np.zeros(8)
        nat.check(self._lib.pcf_diagnostics(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        keys = ("Epert", "Etot", "divL2", "u_top", "u_bot", "wall_shear_top", "wall_shear_bot",
                "mean_shear")
        return {k: float(v) for k, v in zip(keys, out)}

    def _physical(self, state=None) -> np.ndarray:
        """Physical fluctuation fields of ``state`` (or of the resident state)
        on the unpadded grid, (nf, n0, ny, nz) real, from the device transforms."""
        if state is not None:
            self._ensure_loaded(state)
        nf = 6 if self._cfg.mhd else 3
        out = np.empty((nf, self.N[0], self.N[1], self.N[2]), np.float64)
        nat.check(self._lib.pcf_physical_fields(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def _backward_velocity(self, u, padded: bool = False):
        """``_backward_velocity`` (``channelflow_kmm.py:688-697``): the physical
        fluctuation velocity on the unpadded grid (real float64 (n0, ny, nz)
        arrays; nodes ascending in x).  ``u`` is the velocity of a state this
        solver holds (the last one loaded, stepped or returned); a foreign
        velocity is loaded with g = i ky u2 - i kz u1 for the transform."""
        if padded:
            raise NotImplementedError("the padded physical velocity stays inside the step")
        loaded = self._loaded
        if loaded is None or not all(a is b for a, b in zip(u, loaded.u)):
            ky, kz = self.wavenumbers()
            g = self.mask_nyquist(1j * ky[None] * u[2] - 1j * kz[None] * u[1])
            self.load(self._with_zero_history(tuple(u), g))
            self._loaded = KMMState(u=tuple(u), g=g)
        up = self._physical()
        return up[0], up[1], up[2]

    def velocity_energies(self, state: KMMState | None = None):
        """(int |u_0|^2, int |u_1|^2, int |u_2|^2) dV of ``state``, device reductions."""
        if state is not None:
            self._ensure_loaded(state)
        out = np.zeros(3)
        nat.check(self._lib.pcf_velocity_energies(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return tuple(float(v) for v in out)

    def fluctuation_scalars(self, state: KMMState | None = None) -> dict:
        """The scalars of the reference runner's ``_pcf_fluctuation_scalars``
        (``production/oracles.py:4566-4597``), same keys: the diagnostics plus
        streak_rms = sqrt(int |u_y|^2 / V) and roll_rms =
        sqrt((int |u_x|^2 + int |u_z|^2) / V), all from device reductions."""
        d = self.diagnostics(state)
        e = self.velocity_energies()
        volume = 2.0 * self.Ly * self.Lz
        return {
            "kinetic_energy": d["Epert"],
            "total_kinetic_energy": d["Etot"],
            "divergence_l2": d["divL2"],
            "wall_shear_lower": d["wall_shear_bot"],
            "wall_shear_upper": d["wall_shear_top"],
            "wall_velocity_lower": d["u_bot"],
            "wall_velocity_upper": d["u_top"],
            "mean_shear": d["mean_shear"],
            "streak_rms": math.sqrt(e[1] / volume),
            "roll_rms": math.sqrt((e[0] + e[2]) / volume),
        }

    def field_maxima(self, state: KMMState | None = None):
        """Largest |u_i + U_i| (total velocity, ``total_velocity_physical``,
        ``pcf_fluctuations_jax.py:114-116``) and |b_i + B0_i| (MHD; zeros
        otherwise) on the unpadded physical grid, reduced on the device."""
        if state is not None:
            self._ensure_loaded(state)
        out = np.zeros(6)
        nat.check(self._lib.pcf_field_maxima(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return tuple(float(v) for v in out[:3]), tuple(float(v) for v in out[3:])

    def cfl_scalars(self, state: KMMState | None = None) -> dict:
        """The CFL numbers of the reference's health scalars (``_cfl_scalars``,
        ``production/health.py:120-153``, with the spacings and rates of
        ``curl_health_scalars``, ``:156-180``): per-axis advective and Alfven
        numbers, the diffusive number and the rotation / shear-source rates;
        ``cfl_total`` is what ``production/adaptive.py:115-175`` reads."""
        u_max, b_max = self.field_maxima(state)
        n0 = self.N[0]
        x = np.sort(np.cos((2 * np.arange(n0) + 1) * np.pi / (2 * n0)))
        spacings = (float(np.diff(x).min()), self.Ly / self.N[1], self.Lz / self.N[2])
        dt = self.dt
        out = {}
        worst = 0.0
        for axis, label in enumerate("xyz"):
            out["cfl_advective_" + label] = u_max[axis] * dt / spacings[axis]
            out["cfl_alfven_" + label] = b_max[axis] * dt / spacings[axis]
            worst = max(worst, out["cfl_advective_" + label] + out["cfl_alfven_" + label])
        out["cfl_diffusive"] = max(self.nu, getattr(self, "eta", 0.0)) * dt / min(spacings) ** 2
        out["cfl_rotation"] = 2.0 * abs(getattr(self, "omega", 0.0)) * dt
        out["cfl_shear_source"] = abs(getattr(self, "shear_rate", 0.0)) * dt
        out["cfl_total"] = max(worst, out["cfl_rotation"], out["cfl_shear_source"])
        return out

    # ------------------------------------------------------------ reverse mode
    def vjp(self, state: KMMState, steps: int, cotangent=None, checkpoint_every: int = 20):
        """Run ``steps`` steps from ``state`` and pull a cotangent back to it.

        ``cotangent`` is ``(u0, u1, u2, g)`` of the final state in jax's
        convention for real functions of complex leaves (directional
        derivative ``Re sum(c * d)``, ``examples/pcf_minimal_seed_jax.py:78-94``);
        ``None`` differentiates the final perturbation energy ``Epert``
        (``channelflow_kmm.py:1302-1308``).  The forward is recomputed from
        snapshots every ``checkpoint_every`` steps.  Returns
        ``(final_state, gradient)`` with the gradient a ``KMMState`` holding
        d/d(leaf) of every leaf of ``state`` (history scalars 0.0).
        """
        if self.time_integrator != "SBDF3":
            raise NotImplementedError("the reverse mode covers the SBDF3 step only")
        self._ensure_loaded(state)
        self._loaded = None
        seed = nat.SEED_ENERGY
        if cotangent is not None:
            for name, arr in zip(("U0", "U1", "U2", "G"), cotangent):
                a = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128))
                nat.check(self._lib.pcf_set_cotangent(self._h, nat.FIELD_ID[name],
                                                      a.ctypes.data_as(C.POINTER(C.c_double))))
            seed = nat.SEED_COTANGENT
        nat.check(self._lib.pcf_vjp(self._h, int(steps), int(checkpoint_every), seed))
        grad = self._gradient()
        return self.fetch(), grad

    def value_and_grad(self, state: KMMState, steps: int, checkpoint_every: int = 20):
        """Final ``Epert`` and its gradient with respect to the leaves of ``state``
        (the reference's ``jax.value_and_grad`` of the rolled-out energy,
        ``production/autograd_cost_benchmark.py:228-330``)."""
        final, grad = self.vjp(state, steps, None, checkpoint_every)
        return self.diagnostics()["Epert"], grad, final

    def _gradient(self) -> KMMState:
        nb, nd = self.nb, self.nd

        def get(name, rows, vector=False):
            out = np.empty(rows, np.float64) if vector else np.empty((rows, self.N[1], self.N[2]), np.complex128)
            nat.check(self._lib.pcf_get_gradient(self._h, nat.FIELD_ID[name],
                                                 out.ctypes.data_as(C.POINTER(C.c_double))))
            return out

        groups = []
        for tag in ("NL_OLD", "NL_OLDER", "SOL_OLD", "SOL_OLDER"):
            groups.append((get(f"{tag}_U", nb), get(f"{tag}_G", nd), get(f"{tag}_V", nd, True),
                           get(f"{tag}_W", nd, True)))
        return KMMState(u=(get("U0", nb), get("U1", nd), get("U2", nd)), g=get("G", nd),
                        nonlinear_old=groups[0], nonlinear_older=groups[1], solution_old=groups[2],
                        solution_older=groups[3])

    @staticmethod
    def benchmark_state_fields(state: KMMState):
        return (*state.u, state.g)


def state_leaves(state: KMMState) -> list:
    """Leaves in the reference pytree order (``channelflow_kmm.py:65-76``)."""
    out = list(state.u) + [state.g]
    for grp in (state.nonlinear_old, state.nonlinear_older, state.solution_old, state.solution_older):
        if grp is not None:  # None is an empty subtree, as in jax
            out.extend(grp)
    out.extend([np.asarray(state.history_steps), np.asarray(state.have_old), np.asarray(state.previous_dt)])
    return out


def state_checksum(state: KMMState) -> dict:
    """Full-state checksum as ``production/benchmark.py:234-260`` computes it."""
    real_sum = imag_sum = norm2 = 0.0
    count = 0
    leaves = state_leaves(state)
    for v in leaves:
        v = np.asarray(v)
        real_sum += float(np.sum(np.real(v), dtype=np.float64))
        imag_sum += float(np.sum(np.imag(v), dtype=np.float64))
        norm2 += float(np.sum(np.abs(v) ** 2, dtype=np.float64))
        count += int(v.size)
    return {"leaf_count": len(leaves), "element_count": count, "real_sum": real_sum,
            "imag_sum": imag_sum, "l2_norm": math.sqrt(norm2)}
