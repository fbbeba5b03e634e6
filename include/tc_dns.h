/*
 * tc_dns.h — C ABI of the Taylor-Couette DNS step in
 * libpcfdns.so (the second geometry of the same library).  Plain C types only;
 * every entry point returns 0 on success and a negative TC_E_* code on
 * failure, with a message from tc_last_error().
 *
 * The step is AxisymmetricTCDNSJax of the reference
 * (examples/taylor_couette_dns_jax.py:217-1078; ntheta = 1) or its full-3D
 * TaylorCouetteDNSJax (:1081-1356; ntheta > 1): perturbations
 * (u_r, u_theta, u_z, p) about the circular-Couette flow V = a r + b / r on
 * [Fourier(theta) x] Fourier(z) x Chebyshev(r), Dirichlet velocity modes and
 * the P_N / P_{N-2} pressure pair, per-(m, kz) coupled saddle-point blocks
 * with the (0, 0) pressure constant pinned.
 *
 * Layout at the boundary: every field is exchanged in the reference's
 * full-complex layout, C order ([Ntheta,] Nz, Nr - 2) of interleaved
 * complex128 (re, im) doubles, the Fourier axes in FFT order.  Internally the
 * library keeps the Hermitian half (kz index < Nz / 2, every m); the axial
 * Nyquist mode is not stored (the reference masks it in every nonlinear
 * row).
 */
#ifndef TC_DNS_H
#define TC_DNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TC_ABI_VERSION 3

#define TC_OK 0
#define TC_E_INVALID (-1)
#define TC_E_HIP (-2)
#define TC_E_UNSUPPORTED (-3)
#define TC_E_FIXED_DT (-4)

/* time integrators (examples/taylor_couette_dns_jax.py:286-292):
 * CNAB2 (:758-795): reads NL_OLD when have_old != 0, then NL_OLD holds the
 *   step's nonlinear rows and have_old = 1.
 * SBDF3 (:885-946): ARS(4,4,3) start-up steps while history_steps < 2, then
 *   steady SBDF3 steps; the history fields shift every step.
 * IMEXRK3 (:797-850): every step an ARS(4,4,3) step; history untouched. */
#define TC_TI_CNAB2 0
#define TC_TI_SBDF3 1
#define TC_TI_IMEXRK3 2

/* AxisymmetricTCDNSJax(base, nu, Nr, Nz, Lz, dt, family="C", dealias,
 * time_integrator) with base = CircularCouette(R1, R2, Omega1, Omega2)
 * (examples/taylor_couette_linear_jax.py:37-48). */
typedef struct tc_config {
  int32_t abi_version;   /* = TC_ABI_VERSION                           */
  int32_t nr, nz;        /* Chebyshev points (>= 6); Fourier modes (even) */
  double r1, r2;         /* cylinder radii, 0 < r1 < r2                */
  double omega1, omega2; /* angular velocities of the cylinders        */
  double nu;             /* kinematic viscosity                        */
  double lz;             /* axial period                               */
  double dt;             /* time step                                  */
  double dealias;        /* padding factor of the product grid (1 or 1.5) */
  int32_t time_integrator; /* TC_TI_*                                  */
  int32_t device;        /* HIP device ordinal                         */
  int32_t ntheta;        /* azimuthal Fourier modes: 1 (axisymmetric) or even */
  /* conducting-wall MHD, AxisymmetricMRIDNSJax (:1359-1903): mhd = 1 evolves
   * (u_r, u_theta, u_z, b_r, b_theta, b_z) with the total pressure about the
   * imposed axial field b0 (Alfven units), magnetic diffusivity eta_mag; b_r in
   * the Dirichlet basis, b_theta in the Robin basis of d(r b_theta)/dr = 0,
   * b_z in the Neumann basis.  Axisymmetric only (ntheta = 1). */
  int32_t mhd;
  double b0, eta_mag;
} tc_config;

typedef struct tc_handle tc_handle;

/* field ids for tc_set_field / tc_get_field: ([Ntheta,] Nz, Nr - 2) complex each */
#define TC_F_UR 0
#define TC_F_UT 1
#define TC_F_UZ 2
#define TC_F_P 3
#define TC_F_NL_OLD 4      /* 3 components: 4, 5, 6    */
#define TC_F_NL_OLDER 7    /* 7, 8, 9                  */
#define TC_F_SOL_OLD 10    /* 10, 11, 12               */
#define TC_F_SOL_OLDER 13  /* 13, 14, 15               */
#define TC_F_COUNT 16

/* diagnostics (examples/taylor_couette_dns_jax.py:1061-1069) */
#define TC_DIAG_E 0
#define TC_DIAG_DIV_LINF 1
#define TC_DIAG_CONTINUITY_L2 2
#define TC_DIAG_WALL 3
#define TC_DIAG_ETH 4
#define TC_DIAG_COUNT 5

/* groups for tc_set_component / tc_get_component (component c < 3, or < 6 with
 * mhd; the pressure group has the one component 0) */
#define TC_G_X 0          /* the evolved fields (u, or u then b)  */
#define TC_G_P 1          /* pressure (total pressure with mhd)   */
#define TC_G_NL_OLD 2
#define TC_G_NL_OLDER 3
#define TC_G_SOL_OLD 4
#define TC_G_SOL_OLDER 5

/* diagnostics with mhd (AxisymmetricMRIDNSJax.diagnostics, :1876-1886) */
#define TC_MHD_DIAG_EKIN 0
#define TC_MHD_DIAG_EMAG 1
#define TC_MHD_DIAG_E 2
#define TC_MHD_DIAG_DIVU 3
#define TC_MHD_DIAG_DIVB 4
#define TC_MHD_DIAG_WALL_U 5
#define TC_MHD_DIAG_COUNT 6

const char *tc_last_error(void);

/* AxisymmetricTCDNSJax.__init__ (:229-284): builds the per-mode operators and
 * their inverses on the host in long double, uploads them. */
int32_t tc_create(const tc_config *cfg, tc_handle **out);
int32_t tc_destroy(tc_handle *h);

/* State exchange in the full-complex layout; scalars: have_old, history_steps
 * (AxisymmetricTCState, :114-165). */
int32_t tc_set_field(tc_handle *h, int32_t field, const double *host);
int32_t tc_get_field(tc_handle *h, int32_t field, double *host);
int32_t tc_set_scalars(tc_handle *h, double have_old, double history_steps);
/* The same exchange by group and component (AxisymmetricMRIState, :169-191,
 * with mhd; any configuration). */
int32_t tc_set_component(tc_handle *h, int32_t group, int32_t comp, const double *host);
int32_t tc_get_component(tc_handle *h, int32_t group, int32_t comp, double *host);
int32_t tc_get_scalars(tc_handle *h, double *have_old, double *history_steps);

/* AxisymmetricTCDNSJax.step / solve (:948-990): nsteps steps of the
 * configured integrator on the device state. */
int32_t tc_step(tc_handle *h, int32_t nsteps);

/* tc_step with blocks of 12 steady SBDF3 steps, or of CNAB2 steps after the
 * first, replayed from captured HIP graphs (no reference counterpart: a
 * launch-bound small step); start-up steps and remainders run eagerly.  Same
 * results as tc_step. */
int32_t tc_step_graph(tc_handle *h, int32_t nsteps);

/* E, div_linf, continuity_l2, wall, Eth of the device state (TC_DIAG_*); with
 * mhd the TC_MHD_DIAG_COUNT values TC_MHD_DIAG_*. */
int32_t tc_diagnostics(tc_handle *h, double *out);

/* The linear eigenproblem pencil L q = lambda M q of
 * TaylorCouetteLinearJax.assemble (examples/taylor_couette_linear_jax.py:209-237)
 * on the configuration's radial space (N = cfg->nr, base flow and nu of cfg):
 * out_L, out_M (4n x 4n) complex interleaved, n = Nr - 2.  Host only, no GPU. */
int32_t tc_linear_pencil(const tc_config *cfg, int32_t m, double kz, double *out_L, double *out_M);

/* The conducting-wall MRI pencil of TaylorCouetteMRIJax.assemble
 * (examples/taylor_couette_mri_jax.py:214-321,393-395): L = L0 + nu Lnu +
 * eta Leta and M, (7n x 7n) complex interleaved, unknowns (u_r, u_theta, u_z,
 * p, b_r, b_theta, b_z); cfg supplies the base flow, nu, b0 and eta_mag.
 * Host only. */
int32_t tc_mri_pencil(const tc_config *cfg, int32_t m, double kz, double *out_L, double *out_M);

/* dt change (:971-985): SBDF3 refuses a changed dt (TC_E_FIXED_DT); CNAB2 and
 * IMEXRK3 rebuild their operators. */
int32_t tc_set_dt(tc_handle *h, double dt);

#ifdef __cplusplus
}
#endif

#endif /* TC_DNS_H */
