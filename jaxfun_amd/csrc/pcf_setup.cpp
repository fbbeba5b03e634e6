// Closed-form operator construction (see pcf_setup.hpp for the formulas).
#include "pcf_setup.hpp"

#include <cmath>

namespace pcf {

namespace {

using ld = long double;
const ld PI_L = 3.141592653589793238462643383279502884L;

struct LMat {
  int rows = 0, cols = 0;
  std::vector<ld> a;
  LMat() = default;
  LMat(int r, int c) : rows(r), cols(c), a((size_t)r * c, 0.0L) {}
  ld &operator()(int i, int j) { return a[(size_t)i * cols + j]; }
  ld operator()(int i, int j) const { return a[(size_t)i * cols + j]; }
};

LMat matmul(const LMat &A, const LMat &B) {
  LMat C(A.rows, B.cols);
  for (int i = 0; i < A.rows; ++i)
    for (int k = 0; k < A.cols; ++k) {
      ld aik = A(i, k);
      if (aik == 0) continue;
      for (int j = 0; j < B.cols; ++j) C(i, j) += aik * B(k, j);
    }
  return C;
}

LMat transpose(const LMat &A) {
  LMat T(A.cols, A.rows);
  for (int i = 0; i < A.rows; ++i)
    for (int j = 0; j < A.cols; ++j) T(j, i) = A(i, j);
  return T;
}

// Solve A X = B (A square) by Gaussian elimination with partial pivoting.
bool solve(LMat A, LMat B, LMat &X) {
  const int n = A.rows;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (fabsl(A(i, k)) > fabsl(A(piv, k))) piv = i;
    if (A(piv, k) == 0) return false;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(A(k, j), A(piv, j));
      for (int j = 0; j < B.cols; ++j) std::swap(B(k, j), B(piv, j));
    }
    for (int i = k + 1; i < n; ++i) {
      ld l = A(i, k) / A(k, k);
      if (l == 0) continue;
      for (int j = k; j < n; ++j) A(i, j) -= l * A(k, j);
      for (int j = 0; j < B.cols; ++j) B(i, j) -= l * B(k, j);
    }
  }
  X = LMat(n, B.cols);
  for (int j = 0; j < B.cols; ++j)
    for (int i = n - 1; i >= 0; --i) {
      ld s = B(i, j);
      for (int k = i + 1; k < n; ++k) s -= A(i, k) * X(k, j);
      X(i, j) = s / A(i, i);
    }
  return true;
}

Mat to_double(const LMat &A) {
  Mat M;
  M.rows = A.rows;
  M.cols = A.cols;
  M.a.resize(A.a.size());
  for (size_t i = 0; i < A.a.size(); ++i) M.a[i] = (double)A.a[i];
  return M;
}

}  // namespace

std::string build_operators(int n0, Operators &ops) {
  if (n0 < 9) return "n0 must be >= 9";
  const int N = n0, NB = n0 - 4, ND = n0 - 2;
  ops.n0 = N;
  ops.nb = NB;
  ops.nd = ND;

  // Chebyshev values at the Gauss nodes: T[j][n] = cos(n theta_j).
  LMat T(N, N);
  std::vector<ld> theta(N);
  ops.x.resize(N);
  for (int j = 0; j < N; ++j) {
    theta[j] = PI_L + (2 * j + 1) * PI_L / (2 * N);
    ops.x[j] = (double)cosl(theta[j]);
    for (int n = 0; n < N; ++n) T(j, n) = cosl(n * theta[j]);
  }
  // Derivative matrix: T_b' = sum_{a<b, b-a odd} (2b / c_a) T_a, c_0 = 2.
  LMat D(N, N);
  for (int a = 0; a < N; ++a)
    for (int b = a + 1; b < N; b += 2) D(a, b) = (a == 0) ? (ld)b : (ld)(2 * b);
  LMat D2 = matmul(D, D), D4 = matmul(D2, D2);
  std::vector<ld> h(N, PI_L / 2);
  h[0] = PI_L;

  LMat SB(NB, N), SD(ND, N);
  for (int i = 0; i < NB; ++i) {
    ld n = i;
    SB(i, i) = 1;
    SB(i, i + 2) = 2 * (-n - 2) / (n + 3);
    SB(i, i + 4) = (n + 1) / (n + 3);
  }
  for (int i = 0; i < ND; ++i) {
    SD(i, i) = 1;
    SD(i, i + 2) = -1;
  }
  LMat SBt = transpose(SB), SDt = transpose(SD);
  LMat HSB(NB, N), HSD(ND, N);  // S diag(h)
  for (int i = 0; i < NB; ++i)
    for (int n = 0; n < N; ++n) HSB(i, n) = SB(i, n) * h[n];
  for (int i = 0; i < ND; ++i)
    for (int n = 0; n < N; ++n) HSD(i, n) = SD(i, n) * h[n];

  LMat C0B = matmul(HSB, SBt);
  LMat C2B = matmul(HSB, matmul(D2, SBt));
  LMat C4B = matmul(HSB, matmul(D4, SBt));
  LMat C0D = matmul(HSD, SDt);
  LMat C2D = matmul(HSD, matmul(D2, SDt));

  // Nodal transforms.
  LMat VB = matmul(T, SBt);              // (N x NB)
  LMat VD = matmul(T, SDt);              // (N x ND)
  LMat VD1 = matmul(T, matmul(D, SDt));  // (N x ND)
  LMat VB1 = matmul(T, matmul(D, SBt));  // (N x NB)

  // Chebyshev scalar-product matrix Wq[n][j] = (pi/N) T_n(x_j).
  LMat Wq(N, N);
  for (int n = 0; n < N; ++n)
    for (int j = 0; j < N; ++j) Wq(n, j) = PI_L / N * T(j, n);
  LMat SDWq = matmul(SD, Wq);  // (ND x N) = S_D Wq
  LMat FD;                     // M_D^{-1} S_D Wq
  if (!solve(C0D, SDWq, FD)) return "singular Dirichlet mass matrix";
  LMat G1 = matmul(HSB, matmul(D, SDt));  // S_B diag(h) D S_D^T
  LMat G0 = matmul(HSB, SDt);             // S_B diag(h) S_D^T
  LMat A1 = matmul(G1, FD), A0 = matmul(G0, FD);
  // Ng = W * S_D diag(h) S_D^T F_D c = W * M_D M_D^{-1} S_D Wq c = W * S_D Wq c.
  LMat AG = SDWq;
  // Reconstruction f = M_D^{-1} S_D diag(h) D S_B^T u0.
  LMat rhsR = matmul(HSD, matmul(D, SBt));
  LMat R;
  if (!solve(C0D, rhsR, R)) return "singular Dirichlet mass matrix";
  // Projection onto the clamped basis (state_from_physical of the initial data).
  LMat FB;
  if (!solve(C0B, matmul(SB, Wq), FB)) return "singular clamped mass matrix";

  ops.C0B = to_double(C0B);
  ops.C2B = to_double(C2B);
  ops.C4B = to_double(C4B);
  ops.C0D = to_double(C0D);
  ops.C2D = to_double(C2D);
  ops.VB = to_double(VB);
  ops.VD = to_double(VD);
  ops.VD1 = to_double(VD1);
  ops.VB1 = to_double(VB1);
  ops.A1 = to_double(A1);
  ops.A0 = to_double(A0);
  ops.AG = to_double(AG);
  ops.R = to_double(R);
  ops.FB = to_double(FB);
  ops.FD = to_double(FD);

  // Fejer-I weights (Chebyshev.py:183-209).
  ops.fejer.resize(N);
  for (int j = 0; j < N; ++j) {
    ld corr = 0;
    for (int m = 1; m <= N / 2; ++m) corr += 2 * cosl(2 * theta[j] * m) / (4.0L * m * m - 1);
    ops.fejer[j] = (double)((2.0L / N) * (1 - corr));
  }
  // Wall derivatives of phi_i = T_i - T_{i+2}: T_n'(1) = n^2, T_n'(-1) = (-1)^{n+1} n^2.
  ops.dphi_top.resize(ND);
  ops.dphi_bot.resize(ND);
  for (int i = 0; i < ND; ++i) {
    ld a = (ld)i * i, b = (ld)(i + 2) * (i + 2);
    ops.dphi_top[i] = (double)(a - b);
    ld s = (i % 2 == 0) ? -1.0L : 1.0L;  // (-1)^{i+1}
    ops.dphi_bot[i] = (double)(s * a - s * b);
  }
  return "";
}

const Mat *operator_by_index(const Operators &ops, int which) {
  switch (which) {
    case 0: return &ops.C0B;
    case 1: return &ops.C2B;
    case 2: return &ops.C4B;
    case 3: return &ops.C0D;
    case 4: return &ops.C2D;
    case 5: return &ops.VB;
    case 6: return &ops.VD;
    case 7: return &ops.VD1;
    case 8: return &ops.A1;
    case 9: return &ops.A0;
    case 10: return &ops.AG;
    case 11: return &ops.R;
    case 12: return &ops.VB1;
    case 13: return &ops.FB;
    case 14: return &ops.FD;
    default: return nullptr;
  }
}

}  // namespace pcf
