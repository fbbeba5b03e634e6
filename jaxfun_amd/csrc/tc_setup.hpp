// Host setup of the Taylor-Couette step (tc_setup.cpp): radial
// Galerkin blocks, per-axial-mode saddle-point inverses and the transform
// matrices of the nonlinear term, all formed in long double.
#pragma once

#include <complex>
#include <string>
#include <vector>

namespace tc {

struct Params {
  int nr = 0, nz = 0;          // radial Chebyshev points, axial Fourier modes
  int ntheta = 1;              // azimuthal Fourier modes (1: the axisymmetric class)
  double r1 = 1.0, r2 = 2.0;   // cylinder radii
  double a = 0.0, b = 0.0;     // base flow V = a r + b / r
  double nu = 0.0, lz = 0.0, dt = 0.0;
  double dealias = 1.0;
  int integrator = 0;          // TC_TI_* (tc_dns.h)
  int mhd = 0;                 // 1: conducting-wall MHD (fields u, b; axisymmetric only)
  double b0 = 0.0, eta = 0.0;  // imposed axial field, magnetic diffusivity
};

// Dense row-major matrices; complex ones as interleaved (re, im) doubles.
struct Ops {
  int nf = 3;       // evolved fields: u (3) or u, b (6)
  int nb = 1;       // radial bases of the fields: Dirichlet (and Robin b_theta, Neumann b_z)
  int n = 0;        // velocity (and pressure) dofs per mode: nr - 2
  int kh = 0;       // stored axial modes 0..kh-1 (the Nyquist mode is not stored)
  int nth = 1;      // azimuthal modes (all stored, FFT order)
  int nm = 0;       // stored (m, kz) modes, index q = im * kh + k
  int mth = 1;      // azimuthal points of the product grid
  double fz = 0.0;  // Fourier factor of every form: Lz (axisymmetric) or 2 pi Lz
  int mr = 0, mz = 0;  // padded radial points and axial points of the product grid
  double df = 0.0;  // 2 / (r2 - r1)
  std::vector<double> r_std, r_pad;       // radial nodes (ascending)
  std::vector<double> fejer;              // Fejer-I weights / df at the standard nodes
  std::vector<double> Ev, Ed;             // [nb](mr x n): psi_i and d psi_i / dr at the padded nodes
  std::vector<double> Es, Esd;            // [nb](nr x n): the same at the standard nodes
  std::vector<double> PrT;                // (mr x n): dealiased scalar products of padded values, transposed
  // MHD: the magnetic rows of minus the curl of the EMF from the axial analyses
  // of its padded values (truncated Chebyshev coefficients, derivatives on the
  // standard grid, scalar products with the row's test basis), transposed (mr x n)
  std::vector<double> QbrT, QbtT, QbtdT, QbzT;
  std::vector<double> QbrrT, QbzrT;       // 3D: the (1/r) d/dtheta parts of the b_r and b_z rows
  std::vector<double> MassT;              // [nf](n x n): fz (psi_i, psi_j) / dt of each field's basis, transposed
  std::vector<double> Drc, Dzc, Dtc;      // (n x n): continuity blocks (T_i, d phi_j/dr + phi_j/r), (T_i, phi_j), (T_i, phi_j/r), times fz
  // per-mode tables, each stored transposed (entry (i, j) at [j][i]) so that a
  // thread per output row reads consecutive addresses
  std::vector<double> LinT;               // [nm][nf n x nf n] complex: the linear operator L (ARS stages)
  std::vector<double> LexpT;              // [nm][nf n x nf n] complex: mass + L / 2 (CNAB2)
  // [nm][nf n x (nf + 1) n] complex: the field columns of the pinned implicit
  // block's inverse, unknowns ordered fields first, pressure last
  std::vector<double> AinvT;
  std::vector<double> Ainv0T;             // the same at the start-up scale 1/2 (SBDF3)
  std::vector<double> kz, mq, mq_odd;     // (nm): kz, the azimuthal wavenumber m, m with the Nyquist one zero
};

// Returns an empty string on success, else an error message.
std::string build_ops(const Params &p, Ops &o);

// The linear eigenproblem pencil of examples/taylor_couette_linear_jax.py:209-237
// for azimuthal mode m and axial wavenumber kz: L = L0 + nu Lv and M, (4n x 4n)
// complex, interleaved.
std::string linear_pencil(const Params &p, int m, double kz, std::vector<double> &L, std::vector<double> &M);

// The conducting-wall MRI pencil of examples/taylor_couette_mri_jax.py:214-321
// (L = L0 + nu Lnu + eta Leta and M, (7n x 7n) complex, interleaved, unknowns
// (u_r, u_theta, u_z, p, b_r, b_theta, b_z)).
std::string mri_pencil(const Params &p, int m, double kz, std::vector<double> &L, std::vector<double> &M);

}  // namespace tc
