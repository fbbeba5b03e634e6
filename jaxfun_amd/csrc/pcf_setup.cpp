// Closed-form operator construction (see pcf_setup.hpp for the formulas).
#include "pcf_setup.hpp"

#include <algorithm>
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
  // MHD: Q_k = P_D D^k S_D^T, nodal values VD Q_k.
  LMat Q1, Q2;
  if (!solve(C0D, matmul(HSD, matmul(D, SDt)), Q1)) return "singular Dirichlet mass matrix";
  if (!solve(C0D, matmul(HSD, matmul(D2, SDt)), Q2)) return "singular Dirichlet mass matrix";
  LMat VQ1 = matmul(VD, Q1), VQ2 = matmul(VD, Q2);
  LMat VC1 = matmul(T, D), VC2 = matmul(T, D2);
  // Nodal values of f = R u0 and of its x-derivative: the reconstruction
  // folded into the nodal contraction of the next nonlinear evaluation.
  LMat VDR = matmul(VD, R), VD1R = matmul(VD1, R);

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
  ops.VQ1 = to_double(VQ1);
  ops.VQ2 = to_double(VQ2);
  ops.VC = to_double(T);
  ops.VC1 = to_double(VC1);
  ops.VC2 = to_double(VC2);
  ops.WQ = to_double(Wq);
  ops.VDR = to_double(VDR);
  ops.VD1R = to_double(VD1R);

  // Structured form of the parity blocks.  Far-upper factors from
  //   (T_a, T_b^(2)) = (pi/2)  b (b^2 - a^2)
  //   (T_a, T_b^(4)) = (pi/48) [ (b^7 - 8b^5 + 16b^3) - (3b^5 + 16b) a^2
  //                              + (3b^3 + 8b) a^4 - b a^6 ]      (b >= a)
  // and the stencil boundary sums: for the clamped basis sum S = sum S a^2 = 0,
  // so C2_B has no far part and C4_B's far part is rank 2 (the a^4, a^6 terms);
  // for the Dirichlet basis sum S = 0 and sum S a^2 = -4(I+1), so C2_D's far
  // part is rank 1: -4 pi (I+1).
  {
    const LMat *full[2][3] = {{&C0B, &C2B, &C4B}, {&C0D, &C2D, nullptr}};
    const LMat *stc[2] = {&SB, &SD};
    for (int b = 0; b < 2; ++b)
      for (int p = 0; p < 2; ++p) {
        Operators::Struct &st = ops.st[b][p];
        const int nfull = full[b][0]->rows;
        const int n = (nfull - p + 1) / 2;
        st.n = n;
        st.P = (b == 0) ? 2 : 1;
        st.Q = (b == 0) ? 2 : 1;
        st.R = (b == 0) ? 2 : 1;
        const int BW = st.P + st.Q + 1;
        st.b0.assign((size_t)n * BW, 0.0);
        st.b2.assign((size_t)n * BW, 0.0);
        st.b4.assign((size_t)n * BW, 0.0);
        st.U.assign((size_t)n * st.R, 0.0);
        st.V.assign((size_t)n * st.R, 0.0);
        std::vector<double> *bands[3] = {&st.b0, &st.b2, &st.b4};
        for (int i = 0; i < n; ++i)
          for (int o = 0; o < BW; ++o) {
            const int j = i + o - st.P;
            if (j < 0 || j >= n) continue;
            for (int t = 0; t < 3; ++t)
              if (full[b][t]) (*bands[t])[(size_t)i * BW + o] = (double)(*full[b][t])(2 * i + p, 2 * j + p);
          }
        const LMat &S = *stc[b];
        for (int i = 0; i < n; ++i) {
          const int I = 2 * i + p;
          ld m2 = 0, m4 = 0, m6 = 0, b1 = 0, b3 = 0;
          for (int a = 0; a < N; ++a) {
            const ld sv = S(I, a);
            if (sv == 0) continue;
            const ld aa = (ld)a * a;
            m2 += sv * aa;
            m4 += sv * aa * aa;
            m6 += sv * aa * aa * aa;
            b1 += sv * a;
            b3 += sv * a * aa;
          }
          if (b == 0) {
            st.U[(size_t)i * 2 + 0] = (double)(PI_L / 48 * m4);
            st.U[(size_t)i * 2 + 1] = (double)(PI_L / 48 * m6);
            st.V[(size_t)i * 2 + 0] = (double)(3 * b3 + 8 * b1);
            st.V[(size_t)i * 2 + 1] = (double)(-b1);
          } else {
            // (pi/2) (sum_a S a^2) (-sum_b S b), and -sum_b S b = 2 for every row
            st.U[i] = (double)(PI_L / 2 * m2 * 2);
            st.V[i] = 1.0;
          }
        }
        // verify: band + far reproduces the block that has a far part (C4_B,
        // C2_D); every other block vanishes off the band
        const LMat *farm = full[b][b == 0 ? 2 : 1];
        ld worst = 0, scale = 0, bscale[3] = {0, 0, 0};
        for (int t = 0; t < 3; ++t)
          if (full[b][t])
            for (int i = 0; i < n; ++i)
              for (int j = 0; j < n; ++j)
                bscale[t] = std::max(bscale[t], fabsl((*full[b][t])(2 * i + p, 2 * j + p)));
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) {
            const ld ref = (*farm)(2 * i + p, 2 * j + p);
            scale = std::max(scale, fabsl(ref));
            ld rec = 0;
            if (j - i > st.Q) {
              for (int r = 0; r < st.R; ++r)
                rec += (ld)st.U[(size_t)i * st.R + r] * st.V[(size_t)j * st.R + r];
            } else if (j - i >= -st.P) {
              rec = (b == 0 ? st.b4 : st.b2)[(size_t)i * BW + (j - i + st.P)];
            }
            worst = std::max(worst, fabsl(rec - ref));
            for (int t = 0; t < 3; ++t) {
              if (!full[b][t] || full[b][t] == farm) continue;
              // exact zeros up to the long-double residue of the rational stencils
              if ((j - i > st.Q || j - i < -st.P) &&
                  fabsl((*full[b][t])(2 * i + p, 2 * j + p)) > 1e-14L * bscale[t])
                return "structured form: unexpected entry outside the band";
            }
          }
        if (worst > 1e-11L * scale) return "structured form does not reproduce the Galerkin block";
      }
  }

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

std::string build_diag_operators(int n0, DiagOps &d) {
  if (n0 < 9) return "n0 must be >= 9";
  const int N = n0, NB = n0 - 4, ND = n0 - 2, P = n0 + 2;
  d.P = P;
  d.xp.resize(P);
  // T^(k)_n at the points: T[k](p, n), k = 0, 1, 2, from the three-term
  // recurrence and its derivatives (exact at x = +-1, where the nodal
  // cos(n theta) form has no derivative).
  LMat T[3] = {LMat(P, N), LMat(P, N), LMat(P, N)};
  for (int p = 0; p < P; ++p) {
    ld x = (p < N) ? cosl(PI_L + (2 * p + 1) * PI_L / (2 * N)) : (p == N ? -1.0L : 1.0L);
    d.xp[p] = (double)x;
    T[0](p, 0) = 1;
    if (N > 1) {
      T[0](p, 1) = x;
      T[1](p, 1) = 1;
    }
    for (int n = 2; n < N; ++n) {
      T[0](p, n) = 2 * x * T[0](p, n - 1) - T[0](p, n - 2);
      T[1](p, n) = 2 * T[0](p, n - 1) + 2 * x * T[1](p, n - 1) - T[1](p, n - 2);
      T[2](p, n) = 4 * T[1](p, n - 1) + 2 * x * T[2](p, n - 1) - T[2](p, n - 2);
    }
  }
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
  d.EB0 = to_double(matmul(T[0], SBt));
  d.EB1 = to_double(matmul(T[1], SBt));
  d.ED0 = to_double(matmul(T[0], SDt));
  d.ED1 = to_double(matmul(T[1], SDt));
  d.ED2 = to_double(matmul(T[2], SDt));
  d.EC0 = to_double(T[0]);
  d.EC1 = to_double(T[1]);
  d.EC2 = to_double(T[2]);
  // T^{-1} = diag(1/h) Wq on the Gauss nodes (discrete orthogonality), so the
  // nodal derivative of the interpolant is Dn = T' diag(1/h) Wq.
  LMat Tn(N, N), T1n(N, N), Winv(N, N);
  for (int j = 0; j < N; ++j) {
    const ld th = PI_L + (2 * j + 1) * PI_L / (2 * N);
    for (int n = 0; n < N; ++n) {
      Tn(j, n) = T[0](j, n);
      T1n(j, n) = T[1](j, n);
      Winv(n, j) = (PI_L / N) * cosl(n * th) / (n == 0 ? PI_L : PI_L / 2);
    }
  }
  d.Dn = to_double(matmul(T1n, Winv));
  LMat I(N, N);
  for (int j = 0; j < N; ++j) I(j, j) = 1;
  d.I = to_double(I);
  // check: Dn reproduces the derivative of every T_n, n < N, at the nodes
  ld worst = 0, scale = 0;
  LMat chk = matmul(matmul(T1n, Winv), Tn);
  for (int j = 0; j < N; ++j)
    for (int n = 0; n < N; ++n) {
      worst = std::max(worst, fabsl(chk(j, n) - T1n(j, n)));
      scale = std::max(scale, fabsl(T1n(j, n)));
    }
  if (worst > 1e-12L * scale) return "nodal differentiation matrix does not reproduce T'";
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
    case 15: return &ops.VQ1;
    case 16: return &ops.VQ2;
    case 17: return &ops.VC;
    case 18: return &ops.VC1;
    case 19: return &ops.VC2;
    case 20: return &ops.WQ;
    case 21: return &ops.VDR;
    case 22: return &ops.VD1R;
    default: return nullptr;
  }
}

}  // namespace pcf
