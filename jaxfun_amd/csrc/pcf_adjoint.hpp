// Reverse-mode (adjoint) kernels of the plane-Couette KMM step.
//
// Cotangents mu are gradients for the real inner product Re sum conj(a) b,
// held in the same Hermitian half-spectrum layout as the state; a linear stage
// A propagates them by its conjugate transpose A^H.  With that convention the
// adjoint of the zero-padded inverse FFT is the truncating forward FFT without
// the 1/n factor and vice versa, so the adjoint nonlinear pipeline reuses the
// forward FFT kernels.
#pragma once

#include "pcf_kernels.hpp"

namespace pcf {

// Contraction matrices of the adjoint in precision T, in their natural
// (row-major, rows = Shen index i, columns = node x) orientation so that a
// chunk of outputs is contiguous for the reversed contractions.
template <typename T>
struct AdjMats {
  const T *A1;   // [nb][ldx]  Nu rows (A1 = G1 FD)
  const T *A0;   // [nb][ldx]
  const T *AG;   // [nd][ldx]
  int ldx;       // >= n0
  const T *VB;   // [n0][ldi]  basis values at the nodes (VB[x][i])
  const T *VD;   // [n0][ldi]
  const T *VD1;  // [n0][ldi]
  const T *VDR;   // [n0][ldi]  VD R (columns: clamped index < nb), R the reconstruction f = R u0
  const T *VD1R;  // [n0][ldi]  VD1 R
  int ldi;       // >= nd, multiple of the row chunk
};

// One cotangent set of the primary variables (u0, g, v00, w00).
struct Prim {
  double2 *u0;
  double2 *g;
  double *v;
  double *w;
};
struct PrimC {
  const double2 *u0;
  const double2 *g;
  const double *v;
  const double *w;
};

// Up to four weighted primary sets.
struct PrimSum {
  PrimC t[4];
  double wgt[4];
  int n;
};

// mu_out = S^{-T} mu_in per mode, S = M - gamma L (startup and steady systems);
// smt holds the transposed, index-reversed structure of every parity block.
// With nu.u0 non-null it also writes nu = e0 out + e1 r1 + e2 r2.
hipError_t launch_adj_solve(const Geo &g, const StructMats &smt, double gamma, PrimC in, Prim out,
                            double *scratch, hipStream_t s, Prim nu = Prim{}, PrimC r1 = PrimC{},
                            PrimC r2 = PrimC{}, double e0 = 0, double e1 = 0, double e2 = 0);
// out = S^{-T} in on the Dirichlet blocks for ncomp fields cstride apart (g
// slots only), one launch; scratch holds adj_solve_dir_scratch_doubles.
hipError_t launch_adj_solve_dir(const Geo &g, const StructMats &smt, double gamma, const double2 *in,
                                double2 *out, double *scratch, int ncomp, size_t cstride, hipStream_t s);
size_t adj_solve_dir_scratch_doubles(const Geo &g, int ncomp);
// Steady transposed solve with the factors of the transposed blocks (smt);
// also nu = e0 out + e1 r1 + e2 r2 and, with sig.u0 set, sig = c0 out + c1 r1
// + c2 r2.
hipError_t launch_adj_solve_factored(const Geo &g, const StructMats &smt, const FactorSet &fs, PrimC in,
                                     Prim out, Prim nu, PrimC r1, PrimC r2, double e0, double e1, double e2,
                                     hipStream_t s, Prim sig = Prim{}, double c0 = 0, double c1 = 0,
                                     double c2 = 0);
// The same; ncomp > 0: the Dirichlet blocks only, for ncomp fields whose g
// pointers (in, out, nu, r1, r2, sig) advance by cstride per field.
hipError_t launch_adj_solve_factored_dir(const Geo &g, const StructMats &smt, const FactorSet &fs, PrimC in,
                                         Prim out, Prim nu, PrimC r1, PrimC r2, double e0, double e1, double e2,
                                         hipStream_t s, Prim sig, double c0, double c1, double c2, int ncomp,
                                         size_t cstride);
// out += (M + gamma L)^T r_stage (if r_stage.u0) + M^T (sum of gather terms).
hipError_t launch_adj_mass(const Geo &g, const StructMats &sm, double gamma, PrimC r_stage,
                           const PrimSum &gather, Prim out, hipStream_t s);
// out = reconstruction adjoint of (mu_u0, mu_u1, mu_u2): (u0, g, v00, w00);
// g_extra (optional) is added to out.g.
hipError_t launch_recon_adj(const Geo &g, const double *RcT, int ldk, const double2 *mu0,
                            const double2 *mu1, const double2 *mu2, const double2 *g_extra,
                            Prim out, hipStream_t s);
// Cross-product adjoint on the padded grid: forward fields Yf (6 padded y-line
// fields), cotangent of H in ws.Hz (3 fields) -> cotangents of (u, omega) in ws.Y.
template <typename T>
hipError_t launch_rot_adj(const Geo &g, const NodalMats<T> &nm, const void *Yf, Workspace &ws,
                          const FFTPlan &pz, hipStream_t s);
// MHD: forward fields Yf (12 padded y-line fields u, omega, b, j), cotangents
// of H and E in ws.Hz (6 fields) -> cotangents of (u, omega, b, j) in ws.Y.
template <typename T>
hipError_t launch_rot_adj_mhd(const Geo &g, const NodalMats<T> &nm, const void *Yf, Workspace &ws,
                              const FFTPlan &pz, hipStream_t s);
// Streamwise FFT of nf fields between arbitrary buffers (layouts as k_fft_y).
template <typename T>
hipError_t launch_fft_y_fields(const Geo &g, const NodalMats<T> &nm, const FFTPlan &py, bool inverse,
                               int nf, const void *src, void *dst, hipStream_t s);
// Cotangent of Epert = Ly Lz sum_k c^H Q c: mu = 2 Ly Lz Q c for u0 (QB) and u1, u2 (QD).
hipError_t launch_energy_seed(const Geo &g, const double *QB, const double *QD, double scale,
                              const double2 *u0, const double2 *u1, const double2 *u2, double2 *mu0,
                              double2 *mu1, double2 *mu2, hipStream_t s);
// out = sum_t w_t in_t (every component; overwrites out).
hipError_t launch_prim_sum(const Geo &g, const PrimSum &in, Prim out, hipStream_t s);
// Matrix-core (MFMA) wall-normal contractions: forward nodal transform and
// nonlinear rows, and their adjoints (same outputs as the launchers above).
template <typename T>
hipError_t launch_nodal_mfma(const Geo &g, const NodalMats<T> &nm, const double2 *u0, const double2 *u1,
                             const double2 *u2, Workspace &ws, hipStream_t s);
// The nodal fields from (u0, g, v00, w00) with the velocity reconstruction
// folded in (needs NodalMats::VDRt, VD1Rt).
template <typename T>
hipError_t launch_nodal_prim_mfma(const Geo &g, const NodalMats<T> &nm, const double2 *u0, const double2 *gf,
                                  const double *v00, const double *w00, Workspace &ws, hipStream_t s);
template <typename T>
hipError_t launch_rhs_mfma(const Geo &g, const NodalMats<T> &nm, Workspace &ws, double2 *Nu, double2 *Ng,
                           double *Nv, double *Nw, hipStream_t s);
template <typename T>
hipError_t launch_nodal_adj_mfma(const Geo &g, const AdjMats<T> &am, Workspace &ws, double2 *mu0,
                                 double2 *mu1, double2 *mu2, hipStream_t s);
template <typename T>
hipError_t launch_rhs_adj_mfma(const Geo &g, const AdjMats<T> &am, Prim nu, Workspace &ws, hipStream_t s);
// mu_x of a steady reverse step in one pass: the transform adjoint above, the
// reconstruction adjoint (launch_recon_adj) and M^T sig, sig = sum_j C_j rho_{k+j}
// (the banded steady SBDF3 mass gather).
template <typename T>
hipError_t launch_nodal_adj_x_mfma(const Geo &g, const AdjMats<T> &am, const StructMats &sm, PrimC sig,
                                   Workspace &ws, Prim out, hipStream_t s);
// Per-configuration attributes of the matrix-core kernels (dynamic LDS above 64 KB).
hipError_t mfma_init(const Geo &g, bool fp64);
// LDS bytes of k_rot_adj; false when the padded spanwise grid does not fit.
bool rot_adj_fits(const Geo &g, const FFTPlan &pz, int elem_bytes);

}  // namespace pcf
