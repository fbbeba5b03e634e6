// Kernels of the Taylor-Couette step (tc_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace tc {

// Sizes and device tables of one handle.
struct TcGeo {
  int nf = 3;         // evolved fields: u (3) or u, b (6, MHD)
  int n = 0;          // radial dofs per mode (Nr - 2)
  int kh = 0;         // stored axial modes 0..kh-1
  int nth = 1;        // azimuthal modes (1: axisymmetric)
  int nm = 0;         // stored (m, kz) modes, q = im * kh + k
  int mr = 0, mz = 0; // product grid
  int mth = 1;        // azimuthal points of the product grid
  int nr = 0;         // standard radial points
  int nz = 0;         // standard axial points
  double lz = 0.0;
  double fz = 0.0;    // Fourier factor of the forms: Lz, or 2 pi Lz with an azimuthal axis
  const double *mq = nullptr, *mq_odd = nullptr;  // (nm) azimuthal wavenumbers (odd derivatives: Nyquist 0)
  const double *Dtc = nullptr;                    // (n x n) continuity block (T_i, phi_j / r) times fz
  const double2 *twt = nullptr;                   // (mth) e^{2 pi i q / mth}
  const double2 *twn_t = nullptr;                 // (nth) e^{2 pi i q / nth}
  double2 *dv = nullptr;                          // (nm x nr) diagnostics scratch
  const double *Ev = nullptr, *Ed = nullptr;  // [basis](mr x n): Dirichlet, then Robin b_theta and Neumann b_z (MHD)
  const double *QbrT = nullptr, *QbtT = nullptr, *QbtdT = nullptr, *QbzT = nullptr;  // (mr x n) MHD EMF rows
  const double *QbrrT = nullptr, *QbzrT = nullptr;  // (mr x n) their (1/r) d/dtheta parts (3D MHD)
  const double *PrT = nullptr;                // (mr x n): the projection, transposed
  const double *invr = nullptr;               // (mr)
  const double *kz = nullptr;                 // (nm) axial wavenumbers of the modes
  const double2 *twz = nullptr;               // (mz) e^{2 pi i q / mz}
  const double2 *twn = nullptr;               // (nz) e^{2 pi i q / nz}
  const double *MassT = nullptr;              // [nf](n x n), transposed
  const double2 *LinT = nullptr, *LexpT = nullptr;  // [nm][nf n x nf n] complex, transposed
  const double *Es = nullptr, *Esd = nullptr;    // [basis](nr x n)
  const double *rstd = nullptr, *fejer = nullptr;  // (nr)
  const double *Drc = nullptr, *Dzc = nullptr;     // (n x n)
};

// Operators of a right-hand-side term.
enum TcOp { TC_OP_ID = 0, TC_OP_MASS = 1, TC_OP_LIN = 2, TC_OP_LEXP = 3 };

constexpr int TC_TERMS = 8;

// rhs = sum_t c_t Op_t(X_t) + cN N, N = Pr G (the projected nonlinear rows of
// the G written by k_tc_nonlinear; stored to nout when set); then
// (u, p) = Ainv rhs.  Fields are [component][q][i] with components nm n apart.
// The terms of one call use at most one of TC_OP_LIN and TC_OP_LEXP (the
// kernel applies each operator once to the combined operand).
struct TcRhs {
  int nt = 0;
  int op[TC_TERMS];
  const double2 *x[TC_TERMS];
  double c[TC_TERMS];
  const double2 *G = nullptr;
  double cN = 0.0;
  double2 *nout = nullptr;
  const double2 *AinvT = nullptr;  // [nm][nf n x (nf + 1) n]: the inverse's field columns, transposed
  double2 *uout = nullptr;
  double2 *pout = nullptr;
};

hipError_t launch_nonlinear(const TcGeo &g, const double2 *u, double2 *G, hipStream_t s);
hipError_t launch_rhs_solve(const TcGeo &g, const TcRhs &a, hipStream_t s);
// out[0..4] = E, div_linf, continuity_l2, wall, Eth of u; with nf = 6,
// out[0..5] = Ekin, Emag, E, divu, divb, wall_u.
hipError_t launch_diag(const TcGeo &g, const double2 *u, double *out, hipStream_t s);
// Dynamic LDS bytes of the three kernels (the handle checks them at creation).
size_t nonlinear_lds(const TcGeo &g);  // the axisymmetric or the 3D product kernel
size_t rhs_lds(const TcGeo &g);
size_t diag_lds(const TcGeo &g);
// Raises the dynamic LDS limit of every kernel to bytes (at most the device's
// per-block LDS).
hipError_t allow_lds(size_t bytes);

}  // namespace tc
