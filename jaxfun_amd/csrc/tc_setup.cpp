// Host setup of the axisymmetric Taylor-Couette step, in long double.
//
// Weak forms (examples/taylor_couette_dns_jax.py:523-606): every bilinear form
// (v, c(r) d^m u / dr^m) is the N-point Gauss-Chebyshev quadrature of the
// mapped integral, (1/df) sum_q (pi/N) v(x_q) c(r_q) df^m u^(m)(x_q) with
// df = 2 / (r2 - r1), times Lz (i kz)^m from the Fourier part; velocity modes
// phi_i = T_i - T_{i+2}, pressure modes T_i, i < N - 2.
#include "tc_setup.hpp"

#include <cmath>

#include "../../include/tc_dns.h"

namespace tc {
namespace {

using LD = long double;
using CL = std::complex<LD>;
constexpr LD PI = 3.141592653589793238462643383279502884L;

// Gauss-Chebyshev nodes in ascending order: -cos((2q + 1) pi / (2N)).
std::vector<LD> nodes(int N) {
  std::vector<LD> x(N);
  for (int q = 0; q < N; ++q) x[q] = -cosl((2 * q + 1) * PI / (2 * N));
  return x;
}

// T_k(x), T_k'(x), T_k''(x) for k < n by the three-term recurrences.
void cheb_eval(LD x, int n, std::vector<LD> &t, std::vector<LD> &t1, std::vector<LD> &t2) {
  t.assign(n, 0);
  t1.assign(n, 0);
  t2.assign(n, 0);
  t[0] = 1;
  if (n > 1) {
    t[1] = x;
    t1[1] = 1;
  }
  for (int k = 1; k + 1 < n; ++k) {
    t[k + 1] = 2 * x * t[k] - t[k - 1];
    t1[k + 1] = 2 * t[k] + 2 * x * t1[k] - t1[k - 1];
    t2[k + 1] = 4 * t1[k] + 2 * x * t2[k] - t2[k - 1];
  }
}

// Radial bases: psi_i = sum_k S[i][k] T_k, i < N - 2.  Dirichlet
// T_i - T_{i+2}; Robin psi + alpha psi_x = 0 at both walls (the conducting
// wall's d(r b_theta)/dr = 0 with alpha = R df); Neumann psi_x = 0,
// T_i - i^2/(i+2)^2 T_{i+2} (src/jaxfun/galerkin/composite.py:797-871); the
// pressure modes T_i.
enum { BD = 0, BR = 1, BN = 2, BP = 3, NBASIS = 4 };

std::vector<LD> stencil(int N, int kind, LD aL, LD aR) {
  const int n = N - 2;
  std::vector<LD> S((size_t)n * N, 0);
  for (int i = 0; i < n; ++i) {
    S[(size_t)i * N + i] = 1;
    if (kind == BD) {
      S[(size_t)i * N + i + 2] = -1;
    } else if (kind == BN) {
      S[(size_t)i * N + i + 2] = -(LD)i * i / ((LD)(i + 2) * (i + 2));
    } else if (kind == BR) {
      // T_j(-1) + aL T_j'(-1) = (-1)^j (1 - aL j^2), T_j(1) + aR T_j'(1) = 1 + aR j^2
      auto L = [&](int j) { return ((j % 2) ? -1.0L : 1.0L) * (1 - aL * (LD)j * j); };
      auto R = [&](int j) { return 1 + aR * (LD)j * j; };
      const LD a11 = L(i + 1), a12 = L(i + 2), a21 = R(i + 1), a22 = R(i + 2), b1 = -L(i), b2 = -R(i);
      const LD det = a11 * a22 - a12 * a21;
      S[(size_t)i * N + i + 1] = (b1 * a22 - a12 * b2) / det;
      S[(size_t)i * N + i + 2] = (a11 * b2 - b1 * a21) / det;
    }
  }
  return S;
}

// Basis values at nodes x: V[kind][d][q * n + i], derivative order d in the
// reference coordinate.
struct Basis {
  int N = 0, n = 0;
  std::vector<LD> V[NBASIS][3];
};

Basis basis_at(const std::vector<LD> &x, int N, const std::vector<LD> *S) {
  Basis b;
  b.N = N;
  b.n = N - 2;
  const int Q = (int)x.size(), n = b.n;
  for (int k = 0; k < NBASIS; ++k)
    for (int d = 0; d < 3; ++d) b.V[k][d].assign((size_t)Q * n, 0);
  std::vector<LD> t[3];
  for (int q = 0; q < Q; ++q) {
    cheb_eval(x[q], N, t[0], t[1], t[2]);
    for (int d = 0; d < 3; ++d)
      for (int i = 0; i < n; ++i) {
        b.V[BD][d][(size_t)q * n + i] = t[d][i] - t[d][i + 2];
        b.V[BP][d][(size_t)q * n + i] = t[d][i];
        for (int k : {BR, BN}) {
          const LD *s = &S[k][(size_t)i * N];
          b.V[k][d][(size_t)q * n + i] = t[d][i] + s[i + 1] * t[d][i + 1] + s[i + 2] * t[d][i + 2];
        }
      }
  }
  return b;
}

using Mat = std::vector<LD>;  // row-major n x n

struct Forms {
  int N = 0, n = 0;
  LD df = 1, r1 = 1;
  std::vector<LD> x, r;
  std::vector<LD> S[NBASIS];
  Basis B;
  Forms(int N_, LD r1_, LD r2_) : N(N_), n(N_ - 2), df(2 / (r2_ - r1_)), r1(r1_) {
    x = nodes(N);
    r.resize(N);
    for (int q = 0; q < N; ++q) r[q] = r1 + (x[q] + 1) / df;
    S[BR] = stencil(N, BR, r1_ * df, r2_ * df);
    S[BN] = stencil(N, BN, 0, 0);
    B = basis_at(x, N, S);
  }
  // (test_i, c(r) d^m trial_j / dr^m); test/trial: 'D', 'R', 'N' or 'P'
  static int kind(char c) { return c == 'D' ? BD : c == 'R' ? BR : c == 'N' ? BN : BP; }
  template <class Coef>
  Mat form(char test, char trial, int m, Coef c) const {
    const std::vector<LD> &vt = B.V[kind(test)][0];
    const std::vector<LD> &vu = B.V[kind(trial)][m];
    const LD scale = (PI / N) / df * powl(df, (LD)m);
    Mat out((size_t)n * n, 0);
    for (int q = 0; q < N; ++q) {
      const LD cq = c(r[q]) * scale;
      for (int i = 0; i < n; ++i) {
        const LD a = vt[(size_t)q * n + i] * cq;
        if (a == 0) continue;
        for (int j = 0; j < n; ++j) out[(size_t)i * n + j] += a * vu[(size_t)q * n + j];
      }
    }
    return out;
  }
};

LD one(LD) { return 1; }

// Gauss-Jordan inverse with partial pivoting, complex long double.
bool invert(std::vector<CL> &A, int n) {
  std::vector<CL> I((size_t)n * n, CL(0));
  for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    LD best = std::abs(A[(size_t)c * n + c]);
    for (int r = c + 1; r < n; ++r) {
      const LD v = std::abs(A[(size_t)r * n + c]);
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    if (best == 0) return false;
    if (piv != c)
      for (int j = 0; j < n; ++j) {
        std::swap(A[(size_t)c * n + j], A[(size_t)piv * n + j]);
        std::swap(I[(size_t)c * n + j], I[(size_t)piv * n + j]);
      }
    const CL inv = CL(1) / A[(size_t)c * n + c];
    for (int j = 0; j < n; ++j) {
      A[(size_t)c * n + j] *= inv;
      I[(size_t)c * n + j] *= inv;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const CL f = A[(size_t)r * n + c];
      if (f == CL(0)) continue;
      for (int j = 0; j < n; ++j) {
        A[(size_t)r * n + j] -= f * A[(size_t)c * n + j];
        I[(size_t)r * n + j] -= f * I[(size_t)c * n + j];
      }
    }
  }
  A.swap(I);
  return true;
}

struct Blocks {
  Mat MD, D2, D1r, R2D, TwoOm, Om, Gr, Gz0, Gt0, Dr, Dz0, Dt0;
};

Blocks blocks(const Forms &F, LD a, LD b) {
  Blocks k;
  k.MD = F.form('D', 'D', 0, one);
  k.D2 = F.form('D', 'D', 2, one);
  k.D1r = F.form('D', 'D', 1, [](LD r) { return 1 / r; });
  k.R2D = F.form('D', 'D', 0, [](LD r) { return 1 / (r * r); });
  k.TwoOm = F.form('D', 'D', 0, [&](LD r) { return 2 * (a + b / (r * r)); });
  k.Om = F.form('D', 'D', 0, [&](LD r) { return a + b / (r * r); });
  k.Gt0 = F.form('D', 'P', 0, [](LD r) { return 1 / r; });
  k.Dt0 = F.form('P', 'D', 0, [](LD r) { return 1 / r; });
  k.Gr = F.form('D', 'P', 1, one);
  k.Gz0 = F.form('D', 'P', 0, one);
  k.Dr = F.form('P', 'D', 1, one);
  const Mat Dr1 = F.form('P', 'D', 0, [](LD r) { return 1 / r; });
  for (size_t i = 0; i < k.Dr.size(); ++i) k.Dr[i] += Dr1[i];
  k.Dz0 = F.form('P', 'D', 0, one);
  return k;
}

// The operator part of one (m, kz) mode times sgn (3n x 3n, complex): the
// blocks of _add_avv_terms (examples/taylor_couette_dns_jax.py:1157-1190;
// :523-569 for m = 0): nu (Laplacian - 1/r^2), the base-flow advection
// -Omega d/dtheta, the 2 nu / r^2 azimuthal couplings, 2 Omega u_theta and
// -2a u_r.
std::vector<CL> operator_part(const Blocks &k, int n, LD m, LD kz, LD sgn, const Params &p) {
  const int N3 = 3 * n;
  std::vector<CL> X((size_t)N3 * N3, CL(0));
  const LD nu = p.nu;
  const CL I(0, 1);
  auto at = [&](int bi, int bj, int i, int j) -> CL & { return X[(size_t)(bi * n + i) * N3 + bj * n + j]; };
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const size_t o = (size_t)i * n + j;
      const LD lap = k.D2[o] + k.D1r[o] - kz * kz * k.MD[o] - m * m * k.R2D[o];
      const CL adv = -I * m * k.Om[o];
      at(0, 0, i, j) = sgn * (nu * lap - nu * k.R2D[o] + adv);
      at(1, 1, i, j) = sgn * (nu * lap - nu * k.R2D[o] + adv);
      at(2, 2, i, j) = sgn * (nu * lap + adv);
      at(0, 1, i, j) = sgn * (-nu * 2 * I * m * k.R2D[o] + k.TwoOm[o]);
      at(1, 0, i, j) = sgn * (nu * 2 * I * m * k.R2D[o] - 2 * (LD)p.a * k.MD[o]);
    }
  return X;
}

// Implicit block of one mode at scale ls (:539-569, :1198-1214), times fz,
// with the (m, kz) = (0, 0) pressure constant pinned (:608-612).
std::vector<CL> implicit_block(const Blocks &k, int n, LD m, LD kz, LD ls, LD fz, const Params &p, bool pin) {
  const int N4 = 4 * n, N3 = 3 * n;
  std::vector<CL> A((size_t)N4 * N4, CL(0));
  const std::vector<CL> X = operator_part(k, n, m, kz, -ls, p);
  const CL I(0, 1);
  for (int i = 0; i < N3; ++i)
    for (int j = 0; j < N3; ++j) A[(size_t)i * N4 + j] = fz * X[(size_t)i * N3 + j];
  auto at = [&](int bi, int bj, int i, int j) -> CL & { return A[(size_t)(bi * n + i) * N4 + bj * n + j]; };
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const size_t o = (size_t)i * n + j;
      for (int c = 0; c < 3; ++c) at(c, c, i, j) += fz * k.MD[o] / (LD)p.dt;
      at(0, 3, i, j) = fz * k.Gr[o];
      at(1, 3, i, j) = fz * I * m * k.Gt0[o];
      at(2, 3, i, j) = fz * I * kz * k.Gz0[o];
      at(3, 0, i, j) = fz * k.Dr[o];
      at(3, 1, i, j) = fz * I * m * k.Dt0[o];
      at(3, 2, i, j) = fz * I * kz * k.Dz0[o];
    }
  if (pin) {
    const int row = 3 * n;
    for (int j = 0; j < N4; ++j) A[(size_t)row * N4 + j] = 0;
    A[(size_t)row * N4 + row] = 1;
  }
  return A;
}

// Explicit operator mass + ls L of one mode, times fz (:571-605, :1216-1231);
// with_mass = false gives L alone.  Stored transposed into dst (complex).
void explicit_block_T(const Blocks &k, int n, LD m, LD kz, LD ls, LD fz, const Params &p, bool with_mass,
                      double *dst) {
  const int N3 = 3 * n;
  const std::vector<CL> X = operator_part(k, n, m, kz, ls, p);
  for (int i = 0; i < N3; ++i)
    for (int j = 0; j < N3; ++j) {
      CL v = fz * X[(size_t)i * N3 + j];
      if (with_mass && i / n == j / n) v += fz * k.MD[(size_t)(i % n) * n + (j % n)] / (LD)p.dt;
      dst[2 * ((size_t)j * N3 + i)] = (double)v.real();
      dst[2 * ((size_t)j * N3 + i) + 1] = (double)v.imag();
    }
}

// The velocity columns of the inverse, transposed: dst[j][i] = inv(A)[i][j], j < 3n.
bool inverse_columns(const Blocks &k, int n, LD m, LD kz, LD ls, LD fz, const Params &p, bool pin, double *dst) {
  std::vector<CL> A = implicit_block(k, n, m, kz, ls, fz, p, pin);
  const int N4 = 4 * n, N3 = 3 * n;
  if (!invert(A, N4)) return false;
  for (int i = 0; i < N4; ++i)  // transposed: [j][i]
    for (int j = 0; j < N3; ++j) {
      const CL v = A[(size_t)i * N4 + j];
      dst[2 * ((size_t)j * N4 + i)] = (double)v.real();
      dst[2 * ((size_t)j * N4 + i) + 1] = (double)v.imag();
    }
  return true;
}

// ---- conducting-wall MHD (examples/taylor_couette_dns_jax.py:1359-1571) ----
// Fields u_r, u_theta, u_z, b_r (Dirichlet), b_theta (Robin), b_z (Neumann).
int field_basis(int c) { return c < 4 ? BD : (c == 4 ? BR : BN); }
char basis_char(int k) { return k == BD ? 'D' : k == BR ? 'R' : 'N'; }

struct MhdBlocks {
  Mat M[3], D2[3], D1r[3], R2[3], Om[3];  // per basis: mass, Laplacian pieces, Omega
  Mat TwoOm, M_DR, M_DN, M_RD, M_ND, RomD, R2_DR, R2_RD, Gr, Gz0, Gt0, Dr, Dz0, Dt0;
};

MhdBlocks mhd_blocks(const Forms &F, LD a, LD b) {
  MhdBlocks k;
  for (int t = 0; t < 3; ++t) {
    const char c = basis_char(t);
    k.M[t] = F.form(c, c, 0, one);
    k.D2[t] = F.form(c, c, 2, one);
    k.D1r[t] = F.form(c, c, 1, [](LD r) { return 1 / r; });
    k.R2[t] = F.form(c, c, 0, [](LD r) { return 1 / (r * r); });
    k.Om[t] = F.form(c, c, 0, [&](LD r) { return a + b / (r * r); });
  }
  k.R2_DR = F.form('D', 'R', 0, [](LD r) { return 1 / (r * r); });
  k.R2_RD = F.form('R', 'D', 0, [](LD r) { return 1 / (r * r); });
  k.Gt0 = F.form('D', 'P', 0, [](LD r) { return 1 / r; });
  k.Dt0 = F.form('P', 'D', 0, [](LD r) { return 1 / r; });
  k.TwoOm = F.form('D', 'D', 0, [&](LD r) { return 2 * (a + b / (r * r)); });
  k.M_DR = F.form('D', 'R', 0, one);
  k.M_DN = F.form('D', 'N', 0, one);
  k.M_RD = F.form('R', 'D', 0, one);
  k.M_ND = F.form('N', 'D', 0, one);
  k.RomD = F.form('R', 'D', 0, [&](LD r) { return -2 * b / (r * r); });  // r dOmega/dr
  k.Gr = F.form('D', 'P', 1, one);
  k.Gz0 = F.form('D', 'P', 0, one);
  k.Dr = F.form('P', 'D', 1, one);
  const Mat Dr1 = F.form('P', 'D', 0, [](LD r) { return 1 / r; });
  for (size_t i = 0; i < k.Dr.size(); ++i) k.Dr[i] += Dr1[i];
  k.Dz0 = F.form('P', 'D', 0, one);
  return k;
}

// _add_mhd_terms (:1460-1499; 3D :2005-2059) of one (m, kz) mode times sgn
// (6n x 6n): viscous and resistive Laplacians (with -1/r^2 on the r and
// theta components), the base-flow advection -i m Omega of every field, the
// +-2 i m / r^2 azimuthal couplings of u and of b, the base-flow couplings
// 2 Omega u_theta, -2a u_r and r Omega' b_r, and the i kz B0 line-tension and
// induction couplings.
std::vector<CL> mhd_operator_part(const MhdBlocks &k, int n, LD m, LD kz, LD sgn, const Params &p) {
  const int N6 = 6 * n;
  std::vector<CL> X((size_t)N6 * N6, CL(0));
  const LD nu = p.nu, eta = p.eta;
  const CL ikB(0, kz * (LD)p.b0);
  auto at = [&](int bi, int bj, int i, int j) -> CL & { return X[(size_t)(bi * n + i) * N6 + bj * n + j]; };
  const CL I(0, 1);
  auto lap = [&](int t, size_t o, bool hoop) {
    return k.D2[t][o] + k.D1r[t][o] - kz * kz * k.M[t][o] - m * m * k.R2[t][o] - (hoop ? k.R2[t][o] : 0);
  };
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const size_t o = (size_t)i * n + j;
      at(0, 0, i, j) = sgn * nu * lap(BD, o, true);
      at(0, 1, i, j) = sgn * k.TwoOm[o];
      at(0, 3, i, j) = sgn * ikB * k.M[BD][o];
      at(1, 1, i, j) = sgn * nu * lap(BD, o, true);
      at(1, 0, i, j) = sgn * (-2 * (LD)p.a) * k.M[BD][o];
      at(1, 4, i, j) = sgn * ikB * k.M_DR[o];
      at(2, 2, i, j) = sgn * nu * lap(BD, o, false);
      at(2, 5, i, j) = sgn * ikB * k.M_DN[o];
      at(3, 3, i, j) = sgn * eta * lap(BD, o, true);
      at(3, 0, i, j) = sgn * ikB * k.M[BD][o];
      at(4, 4, i, j) = sgn * eta * lap(BR, o, true);
      at(4, 1, i, j) = sgn * ikB * k.M_RD[o];
      at(4, 3, i, j) = sgn * k.RomD[o];
      at(5, 5, i, j) = sgn * eta * lap(BN, o, false);
      at(5, 2, i, j) = sgn * ikB * k.M_ND[o];
      if (m != 0) {
        for (int f = 0; f < 6; ++f) at(f, f, i, j) += sgn * (-I * m) * k.Om[field_basis(f)][o];
        at(0, 1, i, j) += sgn * (-nu * 2 * I * m) * k.R2[BD][o];
        at(1, 0, i, j) += sgn * (nu * 2 * I * m) * k.R2[BD][o];
        at(3, 4, i, j) += sgn * (-eta * 2 * I * m) * k.R2_DR[o];
        at(4, 3, i, j) += sgn * (eta * 2 * I * m) * k.R2_RD[o];
      }
    }
  return X;
}

// mass + ls L (with_mass) or L alone, times fz, stored transposed (complex).
void mhd_explicit_T(const MhdBlocks &k, int n, LD m, LD kz, LD ls, LD fz, const Params &p, bool with_mass,
                    double *dst) {
  const int N6 = 6 * n;
  const std::vector<CL> X = mhd_operator_part(k, n, m, kz, ls, p);
  for (int i = 0; i < N6; ++i)
    for (int j = 0; j < N6; ++j) {
      CL v = fz * X[(size_t)i * N6 + j];
      if (with_mass && i / n == j / n) v += fz * k.M[field_basis(i / n)][(size_t)(i % n) * n + (j % n)] / (LD)p.dt;
      dst[2 * ((size_t)j * N6 + i)] = (double)v.real();
      dst[2 * ((size_t)j * N6 + i) + 1] = (double)v.imag();
    }
}

// The implicit 7-field block (:1501-1549) in the order (u_r, u_theta, u_z,
// b_r, b_theta, b_z, Pi), times fz, kz = 0 pinned; the field columns of its
// inverse, transposed.
bool mhd_inverse_columns(const MhdBlocks &k, int n, LD m, LD kz, LD ls, LD fz, const Params &p, bool pin,
                         double *dst) {
  const int N6 = 6 * n, N7 = 7 * n;
  std::vector<CL> A((size_t)N7 * N7, CL(0));
  const std::vector<CL> X = mhd_operator_part(k, n, m, kz, -ls, p);
  for (int i = 0; i < N6; ++i)
    for (int j = 0; j < N6; ++j) {
      CL v = fz * X[(size_t)i * N6 + j];
      if (i / n == j / n) v += fz * k.M[field_basis(i / n)][(size_t)(i % n) * n + (j % n)] / (LD)p.dt;
      A[(size_t)i * N7 + j] = v;
    }
  const CL I(0, 1);
  auto at = [&](int bi, int bj, int i, int j) -> CL & { return A[(size_t)(bi * n + i) * N7 + bj * n + j]; };
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const size_t o = (size_t)i * n + j;
      at(0, 6, i, j) = fz * k.Gr[o];
      at(1, 6, i, j) = fz * I * m * k.Gt0[o];
      at(2, 6, i, j) = fz * I * kz * k.Gz0[o];
      at(6, 0, i, j) = fz * k.Dr[o];
      at(6, 1, i, j) = fz * I * m * k.Dt0[o];
      at(6, 2, i, j) = fz * I * kz * k.Dz0[o];
    }
  if (pin) {
    const int row = N6;
    for (int j = 0; j < N7; ++j) A[(size_t)row * N7 + j] = 0;
    A[(size_t)row * N7 + row] = 1;
  }
  if (!invert(A, N7)) return false;
  for (int i = 0; i < N7; ++i)
    for (int j = 0; j < N6; ++j) {
      const CL v = A[(size_t)i * N7 + j];
      dst[2 * ((size_t)j * N7 + i)] = (double)v.real();
      dst[2 * ((size_t)j * N7 + i) + 1] = (double)v.imag();
    }
  return true;
}

}  // namespace

std::string build_ops(const Params &p, Ops &o) {
  if (p.nr < 6) return "Nr must be at least 6";
  if (p.nz < 2 || p.nz % 2) return "Nz must be even and at least 2";
  if (p.ntheta < 1 || (p.ntheta > 1 && p.ntheta % 2)) return "Ntheta must be 1 or even";
  if (!(p.r2 > p.r1 && p.r1 > 0)) return "require 0 < R1 < R2";
  if (!(p.lz > 0) || !(p.dt > 0) || !(p.nu > 0)) return "Lz, dt and nu must be positive";
  if (p.mhd && !(p.eta > 0)) return "eta_mag must be positive";
  const int N = p.nr, n = N - 2;
  o.n = n;
  o.nf = p.mhd ? 6 : 3;
  o.nb = p.mhd ? 3 : 1;
  o.kh = p.nz / 2;
  o.mr = p.dealias > 1.0 ? (int)(p.nr * p.dealias) : p.nr;
  o.mz = p.dealias > 1.0 ? (int)(p.nz * p.dealias) : p.nz;
  o.nth = p.ntheta;
  o.nm = o.nth * o.kh;
  o.mth = (p.ntheta > 1 && p.dealias > 1.0) ? (int)(p.ntheta * p.dealias) : p.ntheta;
  o.fz = p.ntheta > 1 ? (double)(2 * PI * (LD)p.lz) : p.lz;
  if (o.mr < N || o.mz < p.nz || o.mth < p.ntheta) return "invalid dealias factor";
  const Forms F(N, p.r1, p.r2);
  o.df = (double)F.df;
  const Blocks K = blocks(F, p.a, p.b);
  // nonlinear transforms
  const std::vector<LD> xp = nodes(o.mr);
  const Basis Bp = basis_at(xp, N, F.S);
  const size_t pe = (size_t)o.mr * n, se = (size_t)N * n;
  o.r_pad.resize(o.mr);
  o.Ev.assign(o.nb * pe, 0);
  o.Ed.assign(o.nb * pe, 0);
  o.PrT.assign(pe, 0);
  for (int j = 0; j < o.mr; ++j) {
    o.r_pad[j] = (double)(p.r1 + (xp[j] + 1) / F.df);
    for (int i = 0; i < n; ++i) {
      for (int t = 0; t < o.nb; ++t) {
        o.Ev[t * pe + (size_t)j * n + i] = (double)Bp.V[t][0][(size_t)j * n + i];
        o.Ed[t * pe + (size_t)j * n + i] = (double)(F.df * Bp.V[t][1][(size_t)j * n + i]);
      }
      // the padded forward to N coefficients followed by the N-point scalar
      // product is the mr-point Gauss quadrature of phi_i f (phi_i has modes
      // i and i + 2 < N only)
      o.PrT[(size_t)j * n + i] = (double)((PI / o.mr) / F.df * Bp.V[BD][0][(size_t)j * n + i]);
    }
  }
  o.r_std.resize(N);
  o.Es.assign(o.nb * se, 0);
  o.Esd.assign(o.nb * se, 0);
  for (int q = 0; q < N; ++q) {
    o.r_std[q] = (double)F.r[q];
    for (int t = 0; t < o.nb; ++t)
      for (int i = 0; i < n; ++i) {
        o.Es[t * se + (size_t)q * n + i] = (double)F.B.V[t][0][(size_t)q * n + i];
        o.Esd[t * se + (size_t)q * n + i] = (double)(F.df * F.B.V[t][1][(size_t)q * n + i]);
      }
  }
  if (p.mhd) {
    // EMF rows (:1626-1649): the padded forward keeps the first N Chebyshev
    // coefficients, c_k = (2 - [k = 0]) / mr sum_j f(x_j) T_k(x_j); the curl
    // is taken on the standard grid and projected with (pi / N) / df psi_i(x_q)
    std::vector<LD> Wv((size_t)N * o.mr, 0), Wd((size_t)N * o.mr, 0), tp[3], ts[3];
    std::vector<LD> Cf((size_t)N * o.mr);
    for (int j = 0; j < o.mr; ++j) {
      cheb_eval(xp[j], N, tp[0], tp[1], tp[2]);
      for (int k = 0; k < N; ++k) Cf[(size_t)k * o.mr + j] = (k == 0 ? 1 : 2) * tp[0][k] / (LD)o.mr;
    }
    for (int q = 0; q < N; ++q) {
      cheb_eval(F.x[q], N, ts[0], ts[1], ts[2]);
      for (int j = 0; j < o.mr; ++j) {
        LD v = 0, d = 0;
        for (int k = 0; k < N; ++k) {
          v += ts[0][k] * Cf[(size_t)k * o.mr + j];
          d += F.df * ts[1][k] * Cf[(size_t)k * o.mr + j];
        }
        Wv[(size_t)q * o.mr + j] = v;
        Wd[(size_t)q * o.mr + j] = d;
      }
    }
    o.QbrT.assign(pe, 0);
    o.QbtT.assign(pe, 0);
    o.QbtdT.assign(pe, 0);
    o.QbzT.assign(pe, 0);
    o.QbrrT.assign(pe, 0);
    o.QbzrT.assign(pe, 0);
    const LD w = (PI / N) / F.df;
    for (int j = 0; j < o.mr; ++j)
      for (int i = 0; i < n; ++i) {
        LD br = 0, bt = 0, btd = 0, bz = 0, brr = 0, bzr = 0;
        for (int q = 0; q < N; ++q) {
          const LD v = Wv[(size_t)q * o.mr + j], d = Wd[(size_t)q * o.mr + j];
          br += F.B.V[BD][0][(size_t)q * n + i] * v;
          bt += F.B.V[BR][0][(size_t)q * n + i] * v;
          btd += F.B.V[BR][0][(size_t)q * n + i] * d;
          bz -= F.B.V[BN][0][(size_t)q * n + i] * (d + v / F.r[q]);
          brr += F.B.V[BD][0][(size_t)q * n + i] * v / F.r[q];
          bzr += F.B.V[BN][0][(size_t)q * n + i] * v / F.r[q];
        }
        o.QbrrT[(size_t)j * n + i] = (double)(w * brr);
        o.QbzrT[(size_t)j * n + i] = (double)(w * bzr);
        o.QbrT[(size_t)j * n + i] = (double)(w * br);
        o.QbtT[(size_t)j * n + i] = (double)(w * bt);
        o.QbtdT[(size_t)j * n + i] = (double)(w * btd);
        o.QbzT[(size_t)j * n + i] = (double)(w * bz);
      }
  }
  // Fejer-I weights for the unweighted measure (src/jaxfun/galerkin/Chebyshev.py:183-209)
  o.fejer.resize(N);
  for (int q = 0; q < N; ++q) {
    const LD th = PI + (2 * q + 1) * PI / (2 * N);
    LD corr = 0;
    for (int m = 1; m <= N / 2; ++m) corr += 2 * cosl(2 * th * m) / (4 * (LD)m * m - 1);
    o.fejer[q] = (double)((2 / (LD)N) * (1 - corr) / F.df);
  }
  const LD fz = p.ntheta > 1 ? 2 * PI * (LD)p.lz : (LD)p.lz;
  MhdBlocks KM;
  if (p.mhd) KM = mhd_blocks(F, p.a, p.b);
  o.MassT.resize((size_t)o.nf * n * n);
  o.Drc.resize((size_t)n * n);
  o.Dzc.resize((size_t)n * n);
  o.Dtc.resize((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const size_t ij = (size_t)i * n + j;
      for (int c = 0; c < o.nf; ++c)
        o.MassT[(size_t)c * n * n + (size_t)j * n + i] =
            (double)(fz * (p.mhd ? KM.M[field_basis(c)][ij] : K.MD[ij]) / p.dt);
      o.Drc[ij] = (double)(fz * K.Dr[ij]);
      o.Dzc[ij] = (double)(fz * K.Dz0[ij]);
      o.Dtc[ij] = (double)(fz * K.Dt0[ij]);
    }
  // per-mode operators, mode q = im * kh + k (im in FFT order)
  const int kh = o.kh, nm = o.nm, N3 = o.nf * n, N4 = N3 + n;
  o.kz.resize(nm);
  o.mq.resize(nm);
  o.mq_odd.resize(nm);
  o.LinT.assign((size_t)nm * N3 * N3 * 2, 0);
  o.LexpT.assign((size_t)nm * N3 * N3 * 2, 0);
  o.AinvT.assign((size_t)nm * N4 * N3 * 2, 0);
  o.Ainv0T.assign((size_t)nm * N4 * N3 * 2, 0);
  const LD ls = (p.integrator == TC_TI_SBDF3) ? 6.0L / 11.0L : 0.5L;
  for (int im = 0; im < o.nth; ++im)
    for (int k = 0; k < kh; ++k) {
      const int q = im * kh + k;
      const int mi = (im < (o.nth + 1) / 2) ? im : im - o.nth;
      const LD m = mi, kz = 2 * PI * k / p.lz;
      o.kz[q] = (double)kz;
      o.mq[q] = (double)mi;
      o.mq_odd[q] = (o.nth % 2 == 0 && im == o.nth / 2) ? 0.0 : (double)mi;
      const bool pin = (mi == 0 && k == 0);
      if (p.mhd) {
        mhd_explicit_T(KM, n, m, kz, 1, fz, p, false, &o.LinT[(size_t)q * N3 * N3 * 2]);
        mhd_explicit_T(KM, n, m, kz, 0.5L, fz, p, true, &o.LexpT[(size_t)q * N3 * N3 * 2]);
        if (!mhd_inverse_columns(KM, n, m, kz, ls, fz, p, pin, &o.AinvT[(size_t)q * N4 * N3 * 2]))
          return "singular implicit block";
        if (!mhd_inverse_columns(KM, n, m, kz, 0.5L, fz, p, pin, &o.Ainv0T[(size_t)q * N4 * N3 * 2]))
          return "singular start-up block";
        continue;
      }
      explicit_block_T(K, n, m, kz, 1, fz, p, false, &o.LinT[(size_t)q * N3 * N3 * 2]);
      explicit_block_T(K, n, m, kz, 0.5L, fz, p, true, &o.LexpT[(size_t)q * N3 * N3 * 2]);
      if (!inverse_columns(K, n, m, kz, ls, fz, p, pin, &o.AinvT[(size_t)q * N4 * N3 * 2]))
        return "singular implicit block";
      if (!inverse_columns(K, n, m, kz, 0.5L, fz, p, pin, &o.Ainv0T[(size_t)q * N4 * N3 * 2]))
        return "singular start-up block";
    }
  return "";
}

std::string linear_pencil(const Params &p, int m, double kz_, std::vector<double> &Lout, std::vector<double> &Mout) {
  if (p.nr < 6) return "Nr must be at least 6";
  if (!(p.r2 > p.r1 && p.r1 > 0)) return "require 0 < R1 < R2";
  const Forms F(p.nr, p.r1, p.r2);
  const int n = F.n, N4 = 4 * n;
  const LD a = p.a, b = p.b, mm = m, kz = kz_, nu = p.nu;
  const Mat MD = F.form('D', 'D', 0, one);
  const Mat D2 = F.form('D', 'D', 2, one);
  const Mat D1r = F.form('D', 'D', 1, [](LD r) { return 1 / r; });
  const Mat Kc = F.form('D', 'D', 0, [&](LD r) { return mm * mm / (r * r) + kz * kz; });
  const Mat R2 = F.form('D', 'D', 0, [](LD r) { return 1 / (r * r); });
  const Mat Adv = F.form('D', 'D', 0, [&](LD r) { return mm * (a + b / (r * r)); });
  const Mat Two = F.form('D', 'D', 0, [&](LD r) { return 2 * a + 2 * b / (r * r); });
  const Mat Sh = F.form('D', 'D', 0, [&](LD) { return 2 * a; });
  const Mat Cp = F.form('D', 'D', 0, [&](LD r) { return 2 * mm / (r * r); });
  const Mat Gr = F.form('D', 'P', 1, one);
  const Mat Gt = F.form('D', 'P', 0, [&](LD r) { return mm / r; });
  const Mat G0 = F.form('D', 'P', 0, one);
  Mat Dr = F.form('P', 'D', 1, one);
  const Mat Dr1 = F.form('P', 'D', 0, [](LD r) { return 1 / r; });
  for (size_t i = 0; i < Dr.size(); ++i) Dr[i] += Dr1[i];
  const Mat Dt = F.form('P', 'D', 0, [&](LD r) { return mm / r; });
  const Mat D0 = F.form('P', 'D', 0, one);
  std::vector<CL> L((size_t)N4 * N4, CL(0)), M((size_t)N4 * N4, CL(0));
  auto put = [&](std::vector<CL> &X, int bi, int bj, int i, int j) -> CL & {
    return X[(size_t)(bi * n + i) * N4 + bj * n + j];
  };
  const CL I(0, 1);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const size_t o = (size_t)i * n + j;
      const LD lap = D2[o] + D1r[o] - Kc[o];
      const CL adv = -I * Adv[o], couple = I * Cp[o];
      put(L, 0, 0, i, j) = adv + nu * (lap - R2[o]);
      put(L, 0, 1, i, j) = Two[o] - nu * couple;
      put(L, 0, 3, i, j) = -Gr[o];
      put(L, 1, 0, i, j) = -Sh[o] + nu * couple;
      put(L, 1, 1, i, j) = adv + nu * (lap - R2[o]);
      put(L, 1, 3, i, j) = -I * Gt[o];
      put(L, 2, 2, i, j) = adv + nu * lap;
      put(L, 2, 3, i, j) = -I * kz * G0[o];
      put(L, 3, 0, i, j) = Dr[o];
      put(L, 3, 1, i, j) = I * Dt[o];
      put(L, 3, 2, i, j) = I * kz * D0[o];
      for (int c = 0; c < 3; ++c) put(M, c, c, i, j) = MD[o];
    }
  Lout.resize((size_t)N4 * N4 * 2);
  Mout.resize((size_t)N4 * N4 * 2);
  for (size_t i = 0; i < L.size(); ++i) {
    Lout[2 * i] = (double)L[i].real();
    Lout[2 * i + 1] = (double)L[i].imag();
    Mout[2 * i] = (double)M[i].real();
    Mout[2 * i + 1] = (double)M[i].imag();
  }
  return "";
}

std::string mri_pencil(const Params &p, int m, double kz_, std::vector<double> &Lout, std::vector<double> &Mout) {
  if (p.nr < 6) return "Nr must be at least 6";
  if (!(p.r2 > p.r1 && p.r1 > 0)) return "require 0 < R1 < R2";
  const Forms F(p.nr, p.r1, p.r2);
  const int n = F.n, N7 = 7 * n;
  const LD a = p.a, b = p.b, mm = m, kz = kz_, nu = p.nu, eta = p.eta;
  const CL I(0, 1), ikB(0, kz * (LD)p.b0);
  std::vector<CL> L((size_t)N7 * N7, CL(0)), M((size_t)N7 * N7, CL(0));
  enum { UR, UT, UZ, PP, BRR, BTT, BZZ };
  const char bas[7] = {'D', 'D', 'D', 'P', 'D', 'R', 'N'};
  auto add = [&](std::vector<CL> &X, int row, int col, CL c, const Mat &f) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) X[(size_t)(row * n + i) * N7 + col * n + j] += c * f[(size_t)i * n + j];
  };
  auto frm = [&](int row, int col, int d, auto coef) { return F.form(bas[row], bas[col], d, coef); };
  auto lapv = [&](int row, int col, bool hoop) {
    Mat f = frm(row, col, 2, one);
    const Mat f1 = frm(row, col, 1, [](LD r) { return 1 / r; });
    const Mat f0 = frm(row, col, 0, [&](LD r) { return -(mm * mm / (r * r) + kz * kz) - (hoop ? 1 / (r * r) : 0); });
    for (size_t i = 0; i < f.size(); ++i) f[i] += f1[i] + f0[i];
    return f;
  };
  const auto om = [&](LD r) { return a + b / (r * r); };
  // (examples/taylor_couette_mri_jax.py:244-320)
  for (int c : {UR, UT, UZ, BRR, BTT, BZZ}) {
    add(L, c, c, -I * mm, frm(c, c, 0, om));
    add(M, c, c, 1, frm(c, c, 0, one));
  }
  add(L, UR, UR, nu, lapv(UR, UR, true));
  add(L, UR, UT, 1, frm(UR, UT, 0, [&](LD r) { return 2 * om(r); }));
  add(L, UR, UT, -nu * I, frm(UR, UT, 0, [&](LD r) { return 2 * mm / (r * r); }));
  add(L, UR, PP, -1, frm(UR, PP, 1, one));
  add(L, UR, BRR, ikB, frm(UR, BRR, 0, one));
  add(L, UT, UR, -2 * a, frm(UT, UR, 0, one));
  add(L, UT, UR, nu * I, frm(UT, UR, 0, [&](LD r) { return 2 * mm / (r * r); }));
  add(L, UT, UT, nu, lapv(UT, UT, true));
  add(L, UT, PP, -I, frm(UT, PP, 0, [&](LD r) { return mm / r; }));
  add(L, UT, BTT, ikB, frm(UT, BTT, 0, one));
  add(L, UZ, UZ, nu, lapv(UZ, UZ, false));
  add(L, UZ, PP, -I * kz, frm(UZ, PP, 0, one));
  add(L, UZ, BZZ, ikB, frm(UZ, BZZ, 0, one));
  add(L, PP, UR, 1, frm(PP, UR, 1, one));
  add(L, PP, UR, 1, frm(PP, UR, 0, [](LD r) { return 1 / r; }));
  add(L, PP, UT, I, frm(PP, UT, 0, [&](LD r) { return mm / r; }));
  add(L, PP, UZ, I * kz, frm(PP, UZ, 0, one));
  add(L, BRR, UR, ikB, frm(BRR, UR, 0, one));
  add(L, BRR, BRR, eta, lapv(BRR, BRR, true));
  add(L, BRR, BTT, -eta * I, frm(BRR, BTT, 0, [&](LD r) { return 2 * mm / (r * r); }));
  add(L, BTT, UT, ikB, frm(BTT, UT, 0, one));
  add(L, BTT, BRR, 1, frm(BTT, BRR, 0, [&](LD r) { return -2 * b / (r * r); }));
  add(L, BTT, BRR, eta * I, frm(BTT, BRR, 0, [&](LD r) { return 2 * mm / (r * r); }));
  add(L, BTT, BTT, eta, lapv(BTT, BTT, true));
  add(L, BZZ, UZ, ikB, frm(BZZ, UZ, 0, one));
  add(L, BZZ, BZZ, eta, lapv(BZZ, BZZ, false));
  Lout.resize((size_t)N7 * N7 * 2);
  Mout.resize((size_t)N7 * N7 * 2);
  for (size_t i = 0; i < L.size(); ++i) {
    Lout[2 * i] = (double)L[i].real();
    Lout[2 * i + 1] = (double)L[i].imag();
    Mout[2 * i] = (double)M[i].real();
    Mout[2 * i + 1] = (double)M[i].imag();
  }
  return "";
}

}  // namespace tc
