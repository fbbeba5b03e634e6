/*
 * pcf_dns.h — C ABI of the MI355X plane-Couette KMM SBDF3 step library
 * (libpcfdns.so).  Plain C types only; every entry point returns 0 on success
 * and a negative PCF_E_* code on failure, with a message from pcf_last_error().
 *
 * Layout convention at the boundary: every spectral field is exchanged in the
 * reference's full-complex layout, C order (n, Ny, Nz) of interleaved
 * complex128 (re, im) doubles — axis 0 is the wall-normal Shen/Chebyshev index,
 * axis 1 the streamwise Fourier index, axis 2 the spanwise Fourier index, both
 * in FFT order.  Mean-mode vectors (v00, w00 and their nonlinear rows) are real
 * float64 of length n.  Internally the library keeps the Hermitian half
 * spectrum (spanwise kz < Nz/2; the kz = Nz/2 Nyquist plane is identically
 * zero in every field) and the two halves are converted exactly at the
 * boundary.
 *
 * Threading: one handle per stream; calls on one handle are not thread-safe.
 */
#ifndef PCF_DNS_H
#define PCF_DNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCF_ABI_VERSION 2

/* error codes */
#define PCF_OK 0
#define PCF_E_INVALID (-1)   /* bad argument or configuration        */
#define PCF_E_HIP (-2)       /* HIP runtime failure                  */
#define PCF_E_UNSUPPORTED (-3) /* option not built                   */
#define PCF_E_FIXED_DT (-4)  /* SBDF3 refuses a changed dt           */

/* precision modes */
#define PCF_PREC_FP64 0   /* every stage in FP64                              */
#define PCF_PREC_MIXED 1  /* transforms + nonlinear FP32; solves, mean mode,
                             history and constraint rows FP64                */

/* time integrators (reference: examples/channelflow_kmm.py:149-159).
 * SBDF3: two IMEXRK3 start-up steps, then steady SBDF3 (channelflow_kmm.py:937-1064).
 * IMEXRK3: every step a low-storage IMEXRK3 step (:872-935); the history fields
 *   are not touched.
 * IMEXRK222: every step an ARS(2,2,2) step (:1123-1201 with
 *   src/jaxfun/integrators/imex_rk.py:175-181).  The device history fields are
 *   scratch under IMEXRK3 and IMEXRK222 (the caller keeps the state's history).
 * CNAB2: Crank-Nicolson for the linear terms, variable-step Adams-Bashforth-2
 *   for the nonlinear ones (_cnab2_flow_update, :1071-1121, with
 *   src/jaxfun/integrators/cnab2.py:211-239): reads NL_OLD when have_old != 0
 *   (extrapolation ratio dt / previous_dt), then NL_OLD holds the step's
 *   nonlinear rows, have_old = 1 and previous_dt = dt.
 * The MHD step runs SBDF3, IMEXRK222 (examples/pcf_mhd_jax.py:389-486, every
 * history field passed through) and CNAB2 (:359-387, NL_A_OLD beside NL_OLD);
 * the reverse mode and the steady / graph entry points are SBDF3 only. */
#define PCF_TI_IMEXRK3 1
#define PCF_TI_IMEXRK222 2
#define PCF_TI_SBDF3 3
#define PCF_TI_CNAB2 4

/* nonlinear forms (reference: examples/channelflow_kmm.py:140-142) */
#define PCF_NL_ROTATIONAL 1

/*
 * Problem specification.  Mirrors the constructor of the reference's
 * PlaneCouetteFluctuationJax (examples/pcf_fluctuations_jax.py:35-75) with the
 * benchmark knobs of examples/channelflow_kmm.py:111-193:
 *   N = (n0, ny, nz), domain = ((-1,1), (0,Ly), (0,Lz)), nu = U_wall/Re,
 *   padding = (1, 1.5, 1.5), time_integrator = SBDF3,
 *   nonlinear_form = rotational, coefficient_path = optimized.
 */
typedef struct pcf_config {
  int32_t abi_version;     /* = PCF_ABI_VERSION                              */
  int32_t n0, ny, nz;      /* Chebyshev points; Fourier modes (even)         */
  double lx[2];            /* wall-normal domain, must be (-1, 1)            */
  double ly, lz;           /* periodic lengths                               */
  double nu;               /* kinematic viscosity = U_wall / Re              */
  double u_wall;           /* laminar base flow U = u_wall * x               */
  double dt;               /* fixed time step                                */
  double padding_y, padding_z; /* dealiasing factors (1.5 = 3/2 rule)        */
  int32_t precision;       /* PCF_PREC_*                                     */
  int32_t time_integrator; /* PCF_TI_*                                       */
  int32_t nonlinear_form;  /* PCF_NL_ROTATIONAL                              */
  int32_t device;          /* HIP device ordinal                             */
  /* Magnetohydrodynamics: mhd = 1 evolves a magnetic vector potential A with
   * magnetic diffusivity eta, coupled to the flow by the Lorentz force J x B
   * and the induction EMF u_tot x B (examples/pcf_mhd_jax.py:69-114).
   *   walls = PCF_WALLS_CONDUCTING: A in TD^3 (PlaneCouetteMHDJax);
   *   walls = PCF_WALLS_INSULATING: A_x in TD, A_y, A_z in TC with per-mode
   *     vacuum-matching wall rows (PlaneCouetteMRIShearpyInsulatingJax,
   *     examples/pcf_mhd_mri_shearpy_jax.py:516-712).
   * b0 is a uniform background field added to B in physical space and omega
   * the rotation rate of the Coriolis force (-2 omega u1, 2 omega u0, 0)
   * (examples/pcf_mhd_mri_shearpy_jax.py:149-176); a shear rate S of that
   * class is u_wall = -S.  mhd = 0 ignores walls, eta, b0 and omega.       */
  int32_t mhd;
  int32_t walls;
  double eta;
  double b0[3];
  double omega;
} pcf_config;

#define PCF_WALLS_CONDUCTING 0
#define PCF_WALLS_INSULATING 1

typedef struct pcf_handle pcf_handle;

/* Field identifiers for pcf_set_field / pcf_get_field.  Names follow the
 * reference KMMState (examples/channelflow_kmm.py:50-101):
 *   u = (u0 in TB, u1 in TD, u2 in TD), g in TD,
 *   nonlinear_old/older = (Nu in TB, Ng in TD, Nv00, Nw00),
 *   solution_old/older  = (u0 in TB, g in TD, v00, w00).                    */
enum pcf_field {
  PCF_F_U0 = 0, PCF_F_U1, PCF_F_U2, PCF_F_G,
  PCF_F_NL_OLD_U, PCF_F_NL_OLD_G, PCF_F_NL_OLD_V, PCF_F_NL_OLD_W,
  PCF_F_NL_OLDER_U, PCF_F_NL_OLDER_G, PCF_F_NL_OLDER_V, PCF_F_NL_OLDER_W,
  PCF_F_SOL_OLD_U, PCF_F_SOL_OLD_G, PCF_F_SOL_OLD_V, PCF_F_SOL_OLD_W,
  PCF_F_SOL_OLDER_U, PCF_F_SOL_OLDER_G, PCF_F_SOL_OLDER_V, PCF_F_SOL_OLDER_W,
  /* MHD handles only (MHDState, examples/pcf_mhd_jax.py:34-66): the vector
   * potential A = (Ax, Ay, Az), its nonlinear rows M_A HA of the two previous
   * steps and A of the two previous steps.  Ax rows: n0-2 (TD); Ay, Az rows:
   * n0-2 (TD, conducting walls) or n0 (TC, insulating walls). */
  PCF_F_AX, PCF_F_AY, PCF_F_AZ,
  PCF_F_NL_A_OLD_X, PCF_F_NL_A_OLD_Y, PCF_F_NL_A_OLD_Z,
  PCF_F_NL_A_OLDER_X, PCF_F_NL_A_OLDER_Y, PCF_F_NL_A_OLDER_Z,
  PCF_F_A_OLD_X, PCF_F_A_OLD_Y, PCF_F_A_OLD_Z,
  PCF_F_A_OLDER_X, PCF_F_A_OLDER_Y, PCF_F_A_OLDER_Z,
  PCF_F_COUNT
};

/* Library identity / errors ------------------------------------------------ */
int32_t pcf_abi_version(void);
const char *pcf_last_error(void);

/* Size (elements) of a field in the boundary layout: complex elements for
 * spectral fields (n*ny*nz), real elements for mean-mode vectors (n).
 * Host-only; needs no GPU. */
int64_t pcf_field_size(const pcf_config *cfg, int32_t field);

/* Host-only operator export for verification (no GPU needed).  Writes the
 * closed-form setup matrix `which` (row-major, rows x cols doubles) to `out`
 * when out != NULL; always reports its shape.  Matrices:
 *   0 C0_B=(psi_i,psi_j)  1 C2_B=(psi_i,psi_j'')  2 C4_B=(psi_i,psi_j'''')
 *   3 C0_D=(phi_i,phi_j)  4 C2_D=(phi_i,phi_j'')   (Chebyshev-weighted)
 *   5 VB  = clamped-basis values at the Gauss nodes   (n0 x n0-4)
 *   6 VD  = Dirichlet-basis values                    (n0 x n0-2)
 *   7 VD1 = Dirichlet-basis x-derivative values        (n0 x n0-2)
 *   8 A1, 9 A0: Nu rows from nodal values (n0-4 x n0)
 *   10 AG: Ng rows from nodal values (n0-2 x n0)
 *   11 R: f = P_TD(d/dx u0) reconstruction (n0-2 x n0-4)
 *   12 VB1 = clamped-basis x-derivative values        (n0 x n0-4)
 *   13 FB, 14 FD: Galerkin projection of nodal values onto the clamped /
 *      Dirichlet basis (n0-4 | n0-2 x n0) — composite.py:361-365
 *   15 VQ1, 16 VQ2: nodal values VD P_D D^k S_D^T (k = 1, 2) of the Galerkin
 *      projection P_D of x-derivatives of Dirichlet fields (n0 x n0-2), which
 *      J = curl B needs — composite.py:293-299, pcf_mhd_jax.py:291-311
 *   17 VC, 18 VC1, 19 VC2: Chebyshev values T_j and derivatives T_j', T_j''
 *      at the nodes (n0 x n0); 20 WQ: Chebyshev scalar products of nodal
 *      values (T_k, f) (n0 x n0) — the TC transforms of insulating walls
 *   21 VDR = VD R, 22 VD1R = VD1 R: nodal values of the reconstructed
 *      f = R u0 and of its x-derivative (n0 x n0-4)
 * Replaces the SymPy inner() assembly of examples/channelflow_kmm.py:261-409. */
int32_t pcf_setup_matrix(const pcf_config *cfg, int32_t which, double *out,
                         int32_t *rows, int32_t *cols);

/* Lifecycle ----------------------------------------------------------------- */
/* Replaces PlaneCouetteFluctuationJax.__init__ (examples/pcf_fluctuations_jax.py:35-75,
 * examples/channelflow_kmm.py:111-246): builds operators, allocates device state. */
int32_t pcf_create(const pcf_config *cfg, pcf_handle **out);
int32_t pcf_destroy(pcf_handle *h);

/* Stream: all work of the handle is enqueued on this hipStream_t (NULL = the
 * library's own stream).  Insulating-wall MHD handles fork an internal side
 * stream from it for the vector-potential kernels and join it back within
 * every call.  Passed as void* to keep HIP types out of the ABI. */
int32_t pcf_set_stream(pcf_handle *h, void *hip_stream);
int32_t pcf_synchronize(pcf_handle *h);

/* State exchange (host memory, boundary layout).  Together these carry the
 * whole KMMState pytree; history scalars travel through pcf_set_history. */
int32_t pcf_set_field(pcf_handle *h, int32_t field, const double *host);
int32_t pcf_get_field(pcf_handle *h, int32_t field, double *host);
/* The same exchange with device pointers (memory on the handle's device, same
 * layout: complex fields (rows, ny, nz) complex128, vectors rows float64),
 * ordered on the handle's stream: the caller orders its own producers and
 * consumers against it (pcf_set_stream / pcf_synchronize).  The reference's
 * state leaves live in device memory (jax arrays in KMMState,
 * examples/channelflow_kmm.py:50-101); these keep a solve loop free of host
 * copies. */
int32_t pcf_set_field_device(pcf_handle *h, int32_t field, const double *dev);
int32_t pcf_get_field_device(pcf_handle *h, int32_t field, double *dev);
int32_t pcf_set_history(pcf_handle *h, double history_steps, double have_old,
                        double previous_dt);
int32_t pcf_get_history(pcf_handle *h, double *history_steps, double *have_old,
                        double *previous_dt);

/* Advance nsteps steps of the configured integrator (SBDF3: IMEXRK3 startup
 * while history_steps < 2, then the steady step).  Replaces KMM.solve /
 * KMM.step (examples/channelflow_kmm.py:937-1064,1123-1137,1231-1247).
 * Asynchronous. */
int32_t pcf_step(pcf_handle *h, int32_t nsteps);

/* A new step size for the following steps.  Replaces KMM.set_dt
 * (examples/channelflow_kmm.py:1207-1222): SBDF3 refuses a changed dt
 * (PCF_E_FIXED_DT); IMEXRK222 refactors its stage systems at the new dt;
 * IMEXRK3 and CNAB2 factor per step and take it as is.  Synchronous. */
int32_t pcf_set_dt(pcf_handle *h, double dt);

/* Steady SBDF3 steps only (requires history_steps >= 2); the unit the
 * benchmark times.  Replaces _step_sbdf3_steady (channelflow_kmm.py:1016-1064). */
int32_t pcf_step_steady(pcf_handle *h, int32_t nsteps);

/* nsteps * replays steady steps replayed from HIP graphs of nsteps steps owned
 * by the handle (pcf_step_steady semantics, history scalars included).  A
 * graph holds the ring slots of the phase it was captured at, so the handle
 * caches one per (ring phase, nsteps) and captures on a miss; whole multiples
 * of 12 steps return the rings to their phase and replay one graph.
 * replays = 0 only captures the graph for the current phase (nothing runs and
 * the state is unchanged), so a timed region can start with it cached. */
int32_t pcf_step_steady_graph(pcf_handle *h, int32_t nsteps, int32_t replays);

/* Counters of the handle's graph caches (the steady graphs above and those of
 * the reverse pass's forward sweeps), the counterpart of the reference's
 * rollout cache introspection (ScanRolloutCacheInfo,
 * src/jaxfun/integrators/cnab2.py:23-33; KMM.rollout_cache_info,
 * examples/channelflow_kmm.py:1255-1257).  out[0] generation (times the caches
 * were dropped: pcf_set_stream), out[1] live graphs, out[2] lookups served
 * from the cache, out[3] lookups that captured, out[4] graphs dropped, out[5..n)
 * the distinct step counts of the live graphs, ascending, zero-padded.  n >= 5;
 * returns the number of distinct step counts. */
int32_t pcf_graph_cache_info(pcf_handle *h, int64_t *out, int32_t n);

/* Diagnostics of the current state (synchronous).  out[8] =
 * {Epert, Etot, divL2, u_top, u_bot, wall_shear_top, wall_shear_bot,
 *  mean_shear}; replaces PlaneCouetteFluctuationJax.diagnostics
 * (examples/pcf_fluctuations_jax.py:131-161). */
int32_t pcf_diagnostics(pcf_handle *h, double *out8);

/* Magnetic diagnostics of an MHD handle (synchronous).  out[4] =
 * {Emag, divB_L2, Emag_total, insulating_bc_residual}: the energy
 * sum_i int |b_i|^2 of the fluctuation field b = curl A, the L2 norm of div b
 * on the unpadded grid, the energy of b + b0, and (insulating walls; 0
 * otherwise) the largest wall residual of the vacuum-matching conditions.
 * Replaces PlaneCouetteMHDJax.diagnostics (examples/pcf_mhd_jax.py:760-789)
 * and the corresponding entries of examples/pcf_mhd_mri_shearpy_jax.py:431-513,802-836. */
int32_t pcf_mhd_diagnostics(pcf_handle *h, double *out4);

/* The rotating-shear family's cadence diagnostics (synchronous), on the
 * unpadded grid of the current state, b = curl A the fluctuation field and
 * b + B0 the total field, u_tot = u + (0, u_wall x, 0), V = 2 Ly Lz.
 * out[PCF_XDIAG_COUNT] =
 *   0 reynolds_stress  <u_x u_y>          1 maxwell_stress  -<(b+B0)_x (b+B0)_y>
 *   2 Emag_total  int |b+B0|^2           3..5 mean_b{x,y,z}  <b + B0>
 *   6 mag_energy_fluct_total  int |b|^2
 *   7 dissipation_kinetic  nu int |grad u|^2   8 dissipation_magnetic  eta int |grad b|^2
 *   9, 10 channel_kz1_{velocity,magnetic}_rms  11 bmax  12 bmax_total
 *   13..15 electric_{ideal,resistive,total}_l2 of -(u_tot x (b+B0)), eta curl b, their sum
 *   16..18 divergence_e_l2, divergence_e_ideal_l2, divergence_e_resistive_l2
 *          (after mask_nyquist(TC.forward)), 19 divergence_a_l2,
 *   20 divergence_a_wall_linf, 21..23 electric_{ideal,resistive,}_wall_tangential_linf,
 *   24, 25 faraday_mean_b{y,z}_tendency, 26..28 mean_b{x,y,z}_trace,
 *   29..37 mean_e{x,y,z}_{ideal,resistive,total} (label-major).
 * The derived keys (total_stress, alpha_Sh, alpha, mag_energy_mean,
 * mag_energy_fluct, channel_kz1_total_rms, mean_b_trace_mismatch_linf) are
 * scalar arithmetic on these.  Replaces PlaneCouetteMRIShearpyJax.diagnostics
 * with _dissipation_parts, _kz1_rms, _a_divergence, magnetic_flux_from_a,
 * electric_field_parts, _electric_wall_parts and _electromagnetic_diagnostics
 * (examples/pcf_mhd_mri_shearpy_jax.py:178-513).  The first call allocates
 * the diagnostics' device buffers (kept until pcf_destroy).  MHD handles only. */
#define PCF_XDIAG_COUNT 38
int32_t pcf_mhd_extended_diagnostics(pcf_handle *h, double *out);

/* Largest magnitudes on the unpadded physical grid (synchronous): out[6] =
 * {max|u_x|, max|u_y + U(x)|, max|u_z|, max|b_x + B0_x|, max|b_y + B0_y|,
 *  max|b_z + B0_z|} of the current state, total velocity (base flow U = u_wall x)
 * and, for MHD handles, total magnetic field b = curl A + B0 (zeros for
 * hydrodynamic handles).  The device reductions behind the CFL numbers of
 * production/health.py:120-193 (u_max, b_max of _cfl_scalars from
 * total_velocity_physical and _total_B_physical(B, padded=False)).
 * PCF_E_UNSUPPORTED when ny or nz has no 2/3/4 FFT plan. */
int32_t pcf_field_maxima(pcf_handle *h, double *out6);

/* pcf_field_maxima for callers that cannot allocate (serving loops and
 * harnesses that forbid device or host allocation between steps): the same
 * out[6], bit for bit, with no hipMalloc, hipFree or host heap use.  No
 * counterpart in the reference.  The per-block maxima and their reduction go
 * to one fixed device array of the library (65536 doubles, 512 KiB, one per
 * GPU, held until the process ends); a grid that needs more (3 or 6 doubles
 * per block of the maxima kernel plus 6; a few thousand at 129x128x128) is
 * PCF_E_INVALID.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (two host threads, or two streams in flight) race.  Synchronous. */
int32_t pcf_field_maxima_noalloc(pcf_handle *h, double *out6);

/* Per-component perturbation energies of the current state (synchronous):
 * out[i] = int |u_i|^2 dV, i = x, y, z, by the quadrature of
 * pcf_diagnostics (their sum is Epert).  The integrals behind streak_rms =
 * sqrt(out[1] / V) and roll_rms = sqrt((out[0] + out[2]) / V) of the
 * reference runner's _pcf_fluctuation_scalars (production/oracles.py:4566-4597). */
int32_t pcf_velocity_energies(pcf_handle *h, double *out3);

/* Physical fluctuation fields of the current state on the unpadded grid
 * (synchronous): out holds nf x (n0, ny, nz) real doubles, C order, x the
 * Gauss-node index (ascending nodes), y and z the grid indices, with nf = 3
 * (velocity u) or, for MHD handles, 6 (u, then b = curl A without B0).
 * Replaces solver._backward_velocity (examples/channelflow_kmm.py:688-697)
 * and _backward_B(B, padded=False) (examples/pcf_mhd_jax.py:313-333).
 * PCF_E_UNSUPPORTED when ny or nz has no 2/3/4 FFT plan. */
int32_t pcf_physical_fields(pcf_handle *h, double *out);

/* Per-stage timing.  One steady step launches pcf_kernels_per_step() kernels,
 * one per stage, in the order pcf_stage_name(0..pcf_stage_count()-1):
 * nodal_backward, fft_y_inverse, rotational, fft_y_forward, forward_rhs, solve,
 * reconstruct.
 * pcf_profile_steady runs nsteps steady steps eagerly with HIP events around
 * every kernel on the handle's stream and writes the mean duration of each
 * stage (milliseconds) to stage_ms[pcf_stage_count()].  No reference
 * counterpart (the reference reports only whole-step wall time,
 * production/benchmark.py:282-354). */
int32_t pcf_kernels_per_step(pcf_handle *h);
int32_t pcf_stage_count(void);
const char *pcf_stage_name(int32_t i);
int32_t pcf_profile_steady(pcf_handle *h, int32_t nsteps, double *stage_ms);

/* Reverse-pass timing outside pcf_vjp (synchronous): nsteps (1..256) eager
 * steady steps keep their tape on the handle's stream, then the energy seed
 * and nsteps - 1 reverse steady steps run alone on the reverse stream (no
 * graph capture, no recompute beside them).  out[2] = {ms per reverse steady
 * step, ms per forward steady step with its tape}.  The state is unchanged
 * afterwards and no gradient is left.  Hydrodynamic SBDF3 handles with a
 * bootstrapped history.  No reference counterpart: the reverse pass's kernel
 * timing (production/autograd_cost_benchmark.py times only whole value_and_grad
 * calls). */
int32_t pcf_profile_reverse(pcf_handle *h, int32_t nsteps, double *out2);

/* Mean duration (milliseconds) of one launch of nonlinear stage `stage`
 * (0 nodal_backward .. 4 forward_rhs) for the current state: the stages before
 * it run once to fill its inputs, then `reps` back-to-back launches between two
 * HIP events on the handle's stream (the launch gaps hide behind the queue).
 * The stages only rewrite workspace and the current nonlinear slot, which the
 * next step recomputes.  No reference counterpart. */
int32_t pcf_profile_stage(pcf_handle *h, int32_t stage, int32_t reps, double *ms_per_launch);

/* Reverse mode.  Replaces jax.value_and_grad / jax.vjp over the scanned,
 * rematerialised step (production/autograd_cost_benchmark.py:228-330,
 * src/jaxfun/integrators/cnab2.py:55) for the velocity/vorticity state.
 *
 * pcf_vjp advances the current state nsteps steps (exactly as pcf_step) and
 * then computes the gradient, with respect to every leaf of the state it
 * started from, of either
 *   PCF_SEED_ENERGY:    the final perturbation energy Epert (diagnostics out[0],
 *                       channelflow_kmm.py:1302-1308), or
 *   PCF_SEED_COTANGENT: the linear functional Re sum(c * u_final) given by the
 *                       cotangents c of the final U0, U1, U2, G set with
 *                       pcf_set_cotangent (unset fields count as zero).
 * The forward is recomputed from state snapshots every checkpoint_every steps
 * (the reference's jax.checkpoint per step, cnab2.py:55, is checkpoint_every = 1).
 * On return the handle holds the final state.  Synchronous.
 *
 * Gradients and cotangents use jax's convention for real functions of complex
 * leaves (directional derivative Re sum(g d), examples/pcf_minimal_seed_jax.py:78-94)
 * in the boundary layout of pcf_get_field.  They are exact on the states this
 * library represents (real velocity fields: Hermitian spectra with the Nyquist
 * planes zero); components that no such state can move are returned as zero. */
#define PCF_SEED_ENERGY 0
#define PCF_SEED_COTANGENT 1
int32_t pcf_set_cotangent(pcf_handle *h, int32_t field, const double *host);
int32_t pcf_vjp(pcf_handle *h, int32_t nsteps, int32_t checkpoint_every, int32_t seed);
/* The gradient with respect to flow leaf `field` (PCF_F_U0 .. PCF_F_SOL_OLDER_W;
 * any other field is PCF_E_INVALID) after pcf_vjp. */
int32_t pcf_get_gradient(pcf_handle *h, int32_t field, double *host);

#ifdef __cplusplus
}
#endif

#endif /* PCF_DNS_H */
