// Host-side operator construction for the plane-Couette KMM step.
//
// Everything the kernels need is a dense 1-D matrix along the wall-normal
// Chebyshev axis, built once here in long double from closed forms and then
// rounded to double (and float for the mixed-precision stages):
//
//  * Shen stencils (reference src/jaxfun/galerkin/composite.py:816-826)
//      Dirichlet  phi_i = T_i - T_{i+2}
//      clamped    psi_i = T_i + 2(-i-2)/(i+3) T_{i+2} + (i+1)/(i+3) T_{i+4}
//  * Chebyshev derivative matrix D (integer valued: T_b' = sum_a D[a][b] T_a),
//    so D^2 and D^4 are exact in double.
//  * Galerkin blocks C_m = S diag(h) D^m S^T with h = (pi, pi/2, ...) — the
//    Chebyshev-weighted (psi_i, psi_j^{(m)}) the reference assembles with
//    SymPy inner() (examples/channelflow_kmm.py:261-409).
//  * Transform matrices: basis values / x-derivatives at the Gauss nodes
//    x_j = cos(pi + (2j+1) pi / (2 n0)) (Chebyshev.py:159-181), and the
//    Galerkin projection of nodal values onto the Dirichlet basis,
//    F_D = M_D^{-1} S_D Wq with Wq[n][j] = (pi/n0) T_n(x_j)
//    (Chebyshev.py:303-318, composite.py:361-365).
//  * For MHD, the nodal values VD P_D D^k S_D^T (k = 1, 2) of the Galerkin
//    projection P_D = M_D^{-1} S_D diag(h) of x-derivatives of Dirichlet
//    fields (composite.py:293-299), which J = curl curl A needs.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pcf {

struct Mat {
  int rows = 0, cols = 0;
  std::vector<double> a;  // row-major
  double &operator()(int i, int j) { return a[(size_t)i * cols + j]; }
  double operator()(int i, int j) const { return a[(size_t)i * cols + j]; }
};

struct Operators {
  int n0 = 0, nb = 0, nd = 0;
  // Galerkin blocks
  Mat C0B, C2B, C4B, C0D, C2D;
  // nodal transforms
  Mat VB, VD, VD1, VB1;       // (n0 x nb|nd): basis values / x-derivatives at nodes
  Mat A1, A0, AG;             // (nb|nd x n0): nonlinear rows from nodal values
  Mat R;                      // (nd x nb): f = P_D(d/dx u0)
  Mat FB, FD;                 // (nb|nd x n0): Galerkin projection of nodal values
  Mat VQ1, VQ2;               // (n0 x nd): VD P_D D^k S_D^T, k = 1, 2 (MHD current)
  Mat VC, VC1, VC2;           // (n0 x n0): T_j, T_j', T_j'' at the nodes (TC fields)
  Mat WQ;                     // (n0 x n0): Chebyshev scalar products (T_k, f) of nodal values
  Mat VDR, VD1R;              // (n0 x nb): VD R, VD1 R (nodal f = R u0 and its x-derivative)
  std::vector<double> x;      // Gauss nodes (ascending)
  std::vector<double> fejer;  // Fejer-I weights on the nodes
  std::vector<double> dphi_top, dphi_bot;  // phi_i'(+1), phi_i'(-1)
  // Band + far low-rank form of the parity blocks (see build_operators):
  // [basis 0 clamped / 1 Dirichlet][parity]
  struct Struct {
    int n = 0, P = 0, Q = 0, R = 0;
    std::vector<double> b0, b2, b4;  // [n][P+Q+1], offsets -P..Q
    std::vector<double> U, V;        // [n][R]
  } st[2][2];
};

// Evaluation matrices of the cadence diagnostics on P = n0 + 2 points: the
// n0 Gauss nodes (ascending) followed by the walls x = -1 and x = +1, so that
// wall traces (examples/pcf_fluctuations_jax.py:118-129) come out of the same
// contraction as the nodal values.  Ek<d> holds the d-th x-derivative of the
// basis k at the points: EB (clamped, P x nb), ED (Dirichlet, P x nd), EC
// (Chebyshev, P x n0).  Dn = T D T^{-1} (n0 x n0) differentiates nodal values
// of a degree < n0 interpolant, the TC forward followed by backward_primitive
// in x; I (n0 x n0) is the identity for terms without an x-derivative.
struct DiagOps {
  int P = 0;
  std::vector<double> xp;  // the P points
  Mat EB0, EB1, ED0, ED1, ED2, EC0, EC1, EC2, Dn, I;
};

// Returns an empty string on success, else an error message.
std::string build_operators(int n0, Operators &ops);
std::string build_diag_operators(int n0, DiagOps &d);

// Export one matrix by index (see pcf_setup_matrix in include/pcf_dns.h).
const Mat *operator_by_index(const Operators &ops, int which);

}  // namespace pcf
