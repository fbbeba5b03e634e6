// Host setup of the axisymmetric Taylor-Couette step (tc_setup.cpp): radial
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
};

// Dense row-major matrices; complex ones as interleaved (re, im) doubles.
struct Ops {
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
  std::vector<double> Ev, Ed;             // (mr x n): phi_i and d phi_i / dr at the padded nodes
  std::vector<double> Es, Esd;            // (nr x n): the same at the standard nodes
  std::vector<double> PrT;                // (mr x n): dealiased scalar products of padded values, transposed
  std::vector<double> MassT;              // (n x n): Lz (phi_i, phi_j) / dt, transposed
  std::vector<double> Drc, Dzc, Dtc;      // (n x n): continuity blocks (T_i, d phi_j/dr + phi_j/r), (T_i, phi_j), (T_i, phi_j/r), times fz
  // per-mode tables, each stored transposed (entry (i, j) at [j][i]) so that a
  // thread per output row reads consecutive addresses
  std::vector<double> LinT;               // [nm][3n x 3n] complex: the linear operator L (ARS stages)
  std::vector<double> LexpT;              // [nm][3n x 3n] complex: mass + L / 2 (CNAB2)
  std::vector<double> AinvT;              // [nm][3n x 4n] complex: velocity columns of the pinned implicit block's inverse
  std::vector<double> Ainv0T;             // the same at the start-up scale 1/2 (SBDF3)
  std::vector<double> kz, mq, mq_odd;     // (nm): kz, the azimuthal wavenumber m, m with the Nyquist one zero
};

// Returns an empty string on success, else an error message.
std::string build_ops(const Params &p, Ops &o);

// The linear eigenproblem pencil of examples/taylor_couette_linear_jax.py:209-237
// for azimuthal mode m and axial wavenumber kz: L = L0 + nu Lv and M, (4n x 4n)
// complex, interleaved.
std::string linear_pencil(const Params &p, int m, double kz, std::vector<double> &L, std::vector<double> &M);

}  // namespace tc
