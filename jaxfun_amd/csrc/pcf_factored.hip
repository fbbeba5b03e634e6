// Steady SBDF3 solves with per-mode LU factors computed once per handle.
//
// The steady systems S = M - gamma L (gamma = 6 dt / 11, sbdf3.py:11-12) depend
// on the mode only through k^2 and do not change between steps.  k_factor runs
// the banded + far low-rank elimination of every (basis, parity, mode) block
// once and keeps, per row, the pivot inverse, the near-upper entries, the
// far-part coefficients and the multipliers (7 values per row for the clamped
// blocks, 4 for the Dirichlet ones: 11 MB at 65x64x64).  Each step then only
// forms the right-hand side (banded, in parallel over rows) and runs the
// forward and back substitutions, whose factor loads do not depend on the
// substitution chain.  The reference likewise factors once and substitutes per
// step (src/jaxfun/la/tpmatrix.py:1034-1150, 686-799).
#include <type_traits>

#include "pcf_adjoint.hpp"
#include "pcf_device.hpp"

namespace pcf {
namespace {

// factor value layout of one row: [0] 1/pivot, [1..Q] near-upper entries,
// [Q+1..Q+R] far coefficients, [Q+R+1..Q+R+P] multipliers of rows i-1..i-P
template <int P, int Q, int R, bool HAS4>
__device__ void factor_block(const StructSys &ss, int n, double s0, double s2, double s4, double sf,
                             double *F, size_t rs, size_t vs) {
  constexpr int BW = P + Q + 1;
  double pd[P + 1], pe[P + 1][Q + 1], pw[P + 1][R];
#pragma unroll
  for (int t = 0; t <= P; ++t) {
    pd[t] = 0;
#pragma unroll
    for (int q = 0; q <= Q; ++q) pe[t][q] = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) pw[t][r] = 0;
  }
  for (int i = 0; i < n; ++i) {
    double a[BW];
#pragma unroll
    for (int o = 0; o < BW; ++o) {
      double v = s0 * ss.b0[(size_t)i * BW + o] + s2 * ss.b2[(size_t)i * BW + o];
      if (HAS4) v += s4 * ss.b4[(size_t)i * BW + o];
      a[o] = v;
    }
    double wi[R], lt[P + 1];
#pragma unroll
    for (int r = 0; r < R; ++r) wi[r] = sf * ss.U[(size_t)i * R + r];
#pragma unroll
    for (int t = 0; t <= P; ++t) lt[t] = 0.0;
#pragma unroll
    for (int t = P; t >= 1; --t) {
      if (i - t < 0) continue;
      const double l = a[P - t] * pd[t];
      lt[t] = l;
#pragma unroll
      for (int o = -t + 1; o <= Q; ++o) {
        if (i + o >= n) continue;
        const int rel = o + t;
        double u;
        if (rel <= Q) {
          u = pe[t][rel];
        } else {
          u = 0.0;
#pragma unroll
          for (int r = 0; r < R; ++r) u += pw[t][r] * ss.V[(size_t)(i + o) * R + r];
        }
        a[P + o] -= l * u;
      }
#pragma unroll
      for (int r = 0; r < R; ++r) wi[r] -= l * pw[t][r];
    }
    const double dinv = 1.0 / a[P];
    double *f = F + (size_t)i * rs;
    f[0] = dinv;
#pragma unroll
    for (int q = 1; q <= Q; ++q) f[q * vs] = a[P + q];
#pragma unroll
    for (int r = 0; r < R; ++r) f[(1 + Q + r) * vs] = wi[r];
#pragma unroll
    for (int t = 1; t <= P; ++t) f[(Q + R + t) * vs] = lt[t];
#pragma unroll
    for (int t = P; t >= 2; --t) {
      pd[t] = pd[t - 1];
#pragma unroll
      for (int q = 0; q <= Q; ++q) pe[t][q] = pe[t - 1][q];
#pragma unroll
      for (int r = 0; r < R; ++r) pw[t][r] = pw[t - 1][r];
    }
    pd[1] = dinv;
#pragma unroll
    for (int q = 1; q <= Q; ++q) pe[1][q] = a[P + q];
#pragma unroll
    for (int r = 0; r < R; ++r) pw[1][r] = wi[r];
  }
}

// Forward then back substitution with stored factors.  load(i) gives the
// right-hand side of row i, keep(i, y)/kept(i) park the forward result, and
// store(i, x) receives the solution (rows visited n-1 .. 0).  Rows go in
// chunks of SC: the chunk's right-hand sides, factors and far coefficients are
// read into registers first (independent loads, one wait), then the serial
// chain runs on registers only, so a row costs its few dependent FMAs rather
// than a load latency.
// The value type V is double2 (one thread per complex right-hand side) or
// double (one thread per real or imaginary part: the factors are real, so the
// two parts never mix and half the instructions per thread carry the chain).
__device__ __forceinline__ void fms(double &acc, double a, double b) { acc -= a * b; }
__device__ __forceinline__ void fms(double2 &acc, double a, double2 b) {
  acc.x -= a * b.x;
  acc.y -= a * b.y;
}
__device__ __forceinline__ void fma_to(double &acc, double a, double b) { acc += a * b; }
__device__ __forceinline__ void fma_to(double2 &acc, double a, double2 b) {
  acc.x += a * b.x;
  acc.y += a * b.y;
}
__device__ __forceinline__ double scaled(double v, double f) { return v * f; }
__device__ __forceinline__ double2 scaled(double2 v, double f) { return double2{v.x * f, v.y * f}; }

template <int P, int Q, int R, class V, class Load, class Keep, class Kept, class Store>
__device__ __forceinline__ void substitute(const StructSys &ss, int n, const double *F, size_t rs, size_t vs,
                                           Load load, Keep keep, Kept kept, Store store) {
  constexpr int F1 = Q + 1, SC = 8;
  {
    V yw[P + 1];
#pragma unroll
    for (int t = 0; t <= P; ++t) yw[t] = V{};
    for (int i0 = 0; i0 < n; i0 += SC) {
      V r[SC];
      double l[SC][P + 1];
#pragma unroll
      for (int c = 0; c < SC; ++c) {
        const int i = i0 + c;
        r[c] = V{};
#pragma unroll
        for (int t = 1; t <= P; ++t) l[c][t] = 0.0;
        if (i < n) {
          r[c] = load(i);
#pragma unroll
          for (int t = 1; t <= P; ++t) l[c][t] = F[(size_t)i * rs + (Q + R + t) * vs];
        }
      }
#pragma unroll
      for (int c = 0; c < SC; ++c) {
        const int i = i0 + c;
        if (i < n) {
          // the newest row (t = 1) last: one dependent FMA per row on the chain
          V y = r[c];
#pragma unroll
          for (int t = P; t >= 1; --t) fms(y, l[c][t], yw[t]);
          keep(i, y);
#pragma unroll
          for (int t = P; t >= 2; --t) yw[t] = yw[t - 1];
          yw[1] = y;
        }
      }
    }
  }
  {
    V xw[F1 + 1];
#pragma unroll
    for (int k = 0; k <= F1; ++k) xw[k] = V{};
    V tau[R];
#pragma unroll
    for (int r = 0; r < R; ++r) tau[r] = V{};
    for (int i1 = n - 1; i1 >= 0; i1 -= SC) {
      V y[SC];
      double f[SC][1 + Q + R], v[SC][R];
#pragma unroll
      for (int c = 0; c < SC; ++c) {
        const int i = i1 - c;
        y[c] = V{};
#pragma unroll
        for (int w = 0; w < 1 + Q + R; ++w) f[c][w] = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) v[c][r] = 0.0;
        if (i >= 0) {
          y[c] = kept(i);
#pragma unroll
          for (int w = 0; w < 1 + Q + R; ++w) f[c][w] = F[(size_t)i * rs + w * vs];
          if (i + F1 < n) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[c][r] = ss.V[(size_t)(i + F1) * R + r];
          }
        }
      }
#pragma unroll
      for (int c = 0; c < SC; ++c) {
        const int i = i1 - c;
        if (i >= 0) {
          // tau_r = sum_{j > i + Q} V[j][r] x_j; x_{i+F1} joins it at row i
#pragma unroll
          for (int r = 0; r < R; ++r) fma_to(tau[r], v[c][r], xw[F1]);
          // far part and older rows first, the newest row (q = 1) last: the
          // chain through x_{i+1} is one FMA and the pivot product per row
          V acc = y[c];
#pragma unroll
          for (int r = 0; r < R; ++r) fms(acc, f[c][1 + Q + r], tau[r]);
#pragma unroll
          for (int q = Q; q >= 1; --q) fms(acc, f[c][q], xw[q]);
          const V xi = scaled(acc, f[c][0]);
          store(i, xi);
#pragma unroll
          for (int k = F1; k >= 2; --k) xw[k] = xw[k - 1];
          xw[1] = xi;
        }
      }
    }
  }
}

constexpr int NF_B = 7, NF_D = 4;  // factor values per row: clamped (2,2,2), Dirichlet (1,1,1)

struct Lane2 {
  int sys, par, m;
  bool mean, valid;
};
__device__ __forceinline__ Lane2 lane2(const Geo &g, int t) {
  const int M = g.m, Mp = (M + 63) / 64 * 64;
  Lane2 s{};
  const int sp = t / Mp, m = t - sp * Mp;
  s.valid = true;
  if (sp >= 4) {
    s.valid = (sp == 4 && m < 2);
    s.mean = true;
    s.sys = 1;
    s.par = m;
    s.m = 0;
    return s;
  }
  s.valid = m < M;
  s.mean = false;
  s.sys = sp >> 1;
  s.par = sp & 1;
  s.m = m;
  return s;
}

// dir_only: the Dirichlet blocks alone (the vector-potential systems), no mean.
__global__ void __launch_bounds__(256) k_factor(Geo g, StructMats sm, double gamma, FactorSet fs, bool dir_only) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (dir_only) {
    const int Mp = (g.m + 63) / 64 * 64;
    t += 2 * Mp;
    if (t >= 4 * Mp) return;
  }
  const Lane2 L = lane2(g, t);
  if (!L.valid) return;
  const StructSys &ss = sm.sys[L.sys][L.par];
  double kyd, kzd, kyo;
  mode_k(g, L.m, kyd, kzd, kyo);
  const double k2 = L.mean ? 0.0 : kyd * kyd + kzd * kzd;
  const double W = L.mean ? 1.0 : g.W, nu = g.nu, gm = gamma;
  const size_t M = g.m;
  if (L.sys == 0) {
    const double s2 = W * (1 + 2 * gm * nu * k2), s4 = -W * gm * nu, s0 = -W * (k2 + gm * nu * k2 * k2);
    factor_block<2, 2, 2, true>(ss, ss.n, s0, s2, s4, s4, fs.blk[0][L.par] + L.m, NF_B * M, M);
  } else {
    const double s0 = W * (1 + gm * nu * k2), s2 = -W * gm * nu;
    if (L.mean)
      factor_block<1, 1, 1, false>(ss, ss.n, s0, s2, 0.0, s2, fs.mean[L.par], NF_D, 1);
    else
      factor_block<1, 1, 1, false>(ss, ss.n, s0, s2, 0.0, s2, fs.blk[1][L.par] + L.m, NF_D * M, M);
  }
}

// Steady update: rhs = M z + N, z = sum c_l x_l, N = sum e_l N_l over the three
// levels, then the factored solve.  Block: SL modes of one (basis, parity)
// block; the Dirichlet blocks of mode 0 also carry the mean system in lane SL.
// spb0 = 2: the Dirichlet blocks only, without the mean system, for
// gridDim.y fields whose g-side pointers advance by cstride per field (the
// three vector-potential components).
template <int SL>
__global__ void __launch_bounds__(256) k_solve_factored(Geo g, StructMats sm, FactorSet fs, SolveArgs a,
                                                        int spb0, size_t cstride) {
  static_assert(2 * SL <= 64, "the substitution threads of the modes fill at most the first wave");
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int SLP = SL + 1;
  const int M = g.m;
  const int bps = (M + SL - 1) / SL;
  const int spb = blockIdx.x / bps + spb0;
  const int m0 = (blockIdx.x - (spb - spb0) * bps) * SL;
  const bool has_mean = (spb0 == 0) && (spb >= 2) && (m0 == 0);
  if (blockIdx.y) {
    const size_t off = blockIdx.y * cstride;
    for (int l = 0; l < 3; ++l) {
      if (a.g[l]) a.g[l] += off;
      if (a.Ng[l]) a.Ng[l] += off;
    }
    a.g_out += off;
  }
  double *zn = reinterpret_cast<double *>(smem);  // [row][4][SLP]: z.x, z.y, rhs.x, rhs.y
  double *fl = zn + (size_t)g.npmax * 4 * SLP;    // [row][NF][SLP]: the lanes' factors
  const int sys = spb >> 1, par = spb & 1;
  const StructSys &ss = sm.sys[sys][par];
  const int n = ss.n;
  auto ZN = [&](int j, int v, int lane) -> double & { return zn[((size_t)j * 4 + v) * SLP + lane]; };
  // factors of every (row, value, lane), consecutive lanes on consecutive modes;
  // the value count as a compile-time constant keeps the index split to
  // multiplications
  auto stage_factors = [&](auto nfc) {
    constexpr int NFC = decltype(nfc)::value;
    gather_batched<8>(
        n * NFC * SLP,
        [&](int idx) {
          const int j = idx / (NFC * SLP), rem = idx - j * NFC * SLP, v = rem / SLP, lane = rem - v * SLP;
          const bool mean = (lane == SL);
          const int m = m0 + lane;
          if (mean ? !has_mean : m >= M) return 0.0;
          return mean ? fs.mean[par][(size_t)j * NF_D + v] : fs.blk[sys][par][((size_t)j * NFC + v) * M + m];
        },
        [&](int idx, double v) { fl[idx] = v; });
  };
  if (sys == 0)
    stage_factors(std::integral_constant<int, NF_B>{});
  else
    stage_factors(std::integral_constant<int, NF_D>{});

  // phase 0: z and N of every (row, lane)
  struct ZNv {
    double2 z, n;
  };
  gather_batched<4>(
      n * SLP,
      [&](int idx) {
        const int j = idx / SLP, lane = idx - j * SLP;
        const bool mean = (lane == SL);
        const int m = m0 + lane;
        ZNv r{{0, 0}, {0, 0}};
        if (mean ? !has_mean : m >= M) return r;
        const int J = 2 * j + par;
        if (!mean) {
          const size_t o = (size_t)J * M + m;
          // all three levels of each kind are valid (the launcher points unset
          // ones at a set level with a zero coefficient): branch-free loads
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            const double2 xv = (sys == 0) ? a.u0[l][o] : a.g[l][o];
            const double2 nv = (sys == 0) ? a.Nu[l][o] : a.Ng[l][o];
            r.z.x += a.c[l] * xv.x;
            r.z.y += a.c[l] * xv.y;
            r.n.x += a.e[l] * nv.x;
            r.n.y += a.e[l] * nv.y;
          }
        } else {
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            r.z.x += a.c[l] * a.v[l][J];
            r.z.y += a.c[l] * a.w[l][J];
            r.n.x += a.e[l] * a.Nv[l][J];
            r.n.y += a.e[l] * a.Nw[l][J];
          }
        }
        return r;
      },
      [&](int idx, const ZNv &r) {
        const int j = idx / SLP, lane = idx - j * SLP;
        ZN(j, 0, lane) = r.z.x;
        ZN(j, 1, lane) = r.z.y;
        ZN(j, 2, lane) = r.n.x;
        ZN(j, 3, lane) = r.n.y;
      });
  __syncthreads();
  // phase 1: rhs = N + M z with the banded mass operator (no far part)
  const int P = (sys == 0) ? 2 : 1, Q = P, BW = P + Q + 1;
  for (int idx = threadIdx.x; idx < n * SLP; idx += blockDim.x) {
    const int j = idx / SLP, lane = idx - j * SLP;
    const bool mean = (lane == SL);
    const int m = m0 + lane;
    if (mean ? !has_mean : m >= M) continue;
    double kyd, kzd, kyo;
    mode_k(g, mean ? 0 : m, kyd, kzd, kyo);
    const double k2 = mean ? 0.0 : kyd * kyd + kzd * kzd;
    const double W = mean ? 1.0 : g.W;
    const double mb0 = (sys == 0) ? -W * k2 : W, mb2 = (sys == 0) ? W : 0.0;
    double rx = ZN(j, 2, lane), ry = ZN(j, 3, lane);
    for (int d = -P; d <= Q; ++d) {
      const int jj = j + d;
      if (jj < 0 || jj >= n) continue;
      const size_t bo = (size_t)j * BW + (d + P);
      const double c = mb0 * ss.b0[bo] + mb2 * ss.b2[bo];
      rx += c * ZN(jj, 0, lane);
      ry += c * ZN(jj, 1, lane);
    }
    ZN(j, 2, lane) = rx;
    ZN(j, 3, lane) = ry;
  }
  __syncthreads();
  // phase 2: substitutions, one thread per (lane, real/imaginary part):
  // threads [0, 2 SL) of the first wave; threads 64, 65 (a second wave) take
  // the mean system (parts v, w)
  int lane, part;
  bool mean = false;
  if (threadIdx.x < 2 * SL) {
    lane = threadIdx.x % SL;
    part = threadIdx.x / SL;
  } else if (threadIdx.x >= 64 && threadIdx.x < 66 && has_mean) {
    lane = SL;
    part = threadIdx.x - 64;
    mean = true;
  } else {
    return;
  }
  const int m = mean ? 0 : m0 + lane;
  if (!mean && m >= M) return;
  auto load = [&](int i) { return ZN(i, 2 + part, lane); };
  auto keep = [&](int i, double y) { ZN(i, 2 + part, lane) = y; };
  const double *F = fl + lane;
  const double sub = a.sub;  // output x - sub x_0 (level 0 of the state)
  if (sys == 0) {
    substitute<2, 2, 2, double>(ss, n, F, (size_t)NF_B * SLP, (size_t)SLP, load, keep, load, [&](int i, double x) {
      const size_t o = 2 * ((size_t)(2 * i + par) * M + m) + part;
      if (sub != 0.0) x -= sub * reinterpret_cast<const double *>(a.u0[0])[o];
      reinterpret_cast<double *>(a.u0_out)[o] = x;
    });
  } else if (mean) {
    const double *x0 = part ? a.w[0] : a.v[0];
    double *out = part ? a.w_out : a.v_out;
    substitute<1, 1, 1, double>(ss, n, F, (size_t)NF_D * SLP, (size_t)SLP, load, keep, load, [&](int i, double x) {
      const int J = 2 * i + par;
      if (sub != 0.0) x -= sub * x0[J];
      out[J] = x;
    });
  } else {
    substitute<1, 1, 1, double>(ss, n, F, (size_t)NF_D * SLP, (size_t)SLP, load, keep, load, [&](int i, double x) {
      const size_t o = 2 * ((size_t)(2 * i + par) * M + m) + part;
      if (sub != 0.0) x -= sub * reinterpret_cast<const double *>(a.g[0])[o];
      reinterpret_cast<double *>(a.g_out)[o] = x;
    });
  }
}

// Transposed steady solve of the reverse pass on the index-reversed transposed
// blocks: rho = S^{-T} mu, and nu = e0 rho + e1 r1 + e2 r2.  Block: SL modes of
// one (basis, parity) block (the Dirichlet blocks of mode 0 also carry the
// mean system in lane SL), as k_solve_factored: phase 0 stages the factors and
// the right-hand side rows (reversed) into LDS, phase 1 runs one substitution
// per lane out of LDS, phase 2 writes rho and nu with coalesced stores.
// spb0 = 2: the Dirichlet blocks only, without the mean system, for
// gridDim.y fields whose g pointers advance by cstride per field (the three
// vector-potential components).
template <int SL>
__global__ void __launch_bounds__(256) k_adj_solve_factored(Geo g, StructMats smt, FactorSet fs, PrimC in,
                                                            Prim out, Prim nuo, PrimC r1, PrimC r2,
                                                            double e0, double e1, double e2, Prim sg,
                                                            double c0, double c1, double c2, int spb0,
                                                            size_t cstride) {
  static_assert(2 * SL <= 64, "the substitution threads of the modes fill at most the first wave");
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int SLP = SL + 1;
  const int M = g.m;
  const int bps = (M + SL - 1) / SL;
  const int spb = blockIdx.x / bps + spb0;
  const int m0 = (blockIdx.x - (spb - spb0) * bps) * SL;
  const bool has_mean = (spb0 == 0) && (spb >= 2) && (m0 == 0);
  if (blockIdx.y) {
    const size_t off = blockIdx.y * cstride;
    in.g += off;
    out.g += off;
    nuo.g += off;
    r1.g += off;
    r2.g += off;
    if (sg.g) sg.g += off;
  }
  const int sys = spb >> 1, par = spb & 1;
  const StructSys &ss = smt.sys[sys][par];
  const int n = ss.n;
  double *zn = reinterpret_cast<double *>(smem);  // [ir][2][SLP]: rhs, then y, then x
  double *fl = zn + (size_t)g.npmax * 2 * SLP;    // [ir][NF][SLP]
  auto row = [&](int ir) { return 2 * (n - 1 - ir) + par; };
  auto ZN = [&](int ir, int v, int lane) -> double & { return zn[((size_t)ir * 2 + v) * SLP + lane]; };
  auto active = [&](int lane) { return (lane == SL) ? has_mean : (m0 + lane < M); };
  auto stage_factors = [&](auto nfc) {
    constexpr int NFC = decltype(nfc)::value;
    gather_batched<8>(
        n * NFC * SLP,
        [&](int idx) {
          const int j = idx / (NFC * SLP), rem = idx - j * NFC * SLP, v = rem / SLP, lane = rem - v * SLP;
          if (!active(lane)) return 0.0;
          return (lane == SL) ? fs.mean[par][(size_t)j * NF_D + v]
                              : fs.blk[sys][par][((size_t)j * NFC + v) * M + m0 + lane];
        },
        [&](int idx, double v) { fl[idx] = v; });
  };
  if (sys == 0)
    stage_factors(std::integral_constant<int, NF_B>{});
  else
    stage_factors(std::integral_constant<int, NF_D>{});
  gather_batched<4>(
      n * SLP,
      [&](int idx) {
        const int ir = idx / SLP, lane = idx - ir * SLP;
        if (!active(lane)) return double2{0, 0};
        const int J = row(ir);
        if (lane == SL) return double2{in.v[J], in.w[J]};
        return (sys == 0 ? in.u0 : in.g)[(size_t)J * M + m0 + lane];
      },
      [&](int idx, const double2 &v) {
        const int ir = idx / SLP, lane = idx - ir * SLP;
        ZN(ir, 0, lane) = v.x;
        ZN(ir, 1, lane) = v.y;
      });
  __syncthreads();
  // one substitution thread per (lane, real/imaginary part), as in the
  // forward solve: threads [0, 2 SL), and 64, 65 for the mean system
  int lane = -1, part = 0;
  if (threadIdx.x < 2 * SL) {
    lane = threadIdx.x % SL;
    part = threadIdx.x / SL;
  } else if (threadIdx.x >= 64 && threadIdx.x < 66 && has_mean) {
    lane = SL;
    part = threadIdx.x - 64;
  }
  if (lane >= 0 && active(lane)) {
    auto rw = [&](int i) { return ZN(i, part, lane); };
    auto wr = [&](int i, double v) { ZN(i, part, lane) = v; };
    const double *F = fl + lane;
    if (sys == 0)
      substitute<2, 2, 2, double>(ss, n, F, (size_t)NF_B * SLP, (size_t)SLP, rw, wr, rw, wr);
    else
      substitute<1, 1, 1, double>(ss, n, F, (size_t)NF_D * SLP, (size_t)SLP, rw, wr, rw, wr);
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < n * SLP; idx += blockDim.x) {
    const int ir = idx / SLP, ln = idx - ir * SLP;
    if (!active(ln)) continue;
    const int J = row(ir);
    double2 x{ZN(ir, 0, ln), ZN(ir, 1, ln)};
    if (ln == SL) {
      out.v[J] = x.x;
      out.w[J] = x.y;
      const double2 p{r1.v[J], r1.w[J]}, q{r2.v[J], r2.w[J]};
      nuo.v[J] = e0 * x.x + e1 * p.x + e2 * q.x;
      nuo.w[J] = e0 * x.y + e1 * p.y + e2 * q.y;
      if (sg.u0) {
        sg.v[J] = c0 * x.x + c1 * p.x + c2 * q.x;
        sg.w[J] = c0 * x.y + c1 * p.y + c2 * q.y;
      }
      continue;
    }
    const int m = m0 + ln;
    const size_t o = (size_t)J * M + m;
    double2 v{0, 0}, sv{0, 0};
    if (m / g.kzs != g.ny / 2) {  // the Nyquist row stays masked
      const double2 p = (sys == 0 ? r1.u0 : r1.g)[o], q = (sys == 0 ? r2.u0 : r2.g)[o];
      v = double2{e0 * x.x + e1 * p.x + e2 * q.x, e0 * x.y + e1 * p.y + e2 * q.y};
      sv = double2{c0 * x.x + c1 * p.x + c2 * q.x, c0 * x.y + c1 * p.y + c2 * q.y};
    } else {
      x = double2{0, 0};
    }
    (sys == 0 ? out.u0 : out.g)[o] = x;
    (sys == 0 ? nuo.u0 : nuo.g)[o] = v;
    if (sg.u0) (sys == 0 ? sg.u0 : sg.g)[o] = sv;
  }
}

// z/rhs (4) and factors (up to 7) per row and lane slot
size_t factored_lds_bytes(const Geo &g, int sl) { return (size_t)g.npmax * (4 + NF_B) * (sl + 1) * sizeof(double); }

}  // namespace

size_t factor_doubles(const Geo &g, const StructMats &sm, int basis, int par) {
  if (basis < 0) return (size_t)sm.sys[1][par].n * NF_D;  // mean system
  return (size_t)sm.sys[basis][par].n * (basis == 0 ? NF_B : NF_D) * g.m;
}

hipError_t launch_factor(const Geo &g, const StructMats &sm, double gamma, FactorSet fs, bool dir_only,
                         hipStream_t s) {
  const int Mp = (g.m + 63) / 64 * 64;
  const int threads = dir_only ? 2 * Mp : 4 * Mp + 64;
  hipLaunchKernelGGL(k_factor, dim3((threads + 255) / 256), dim3(256), 0, s, g, sm, gamma, fs, dir_only);
  return hipGetLastError();
}

hipError_t launch_solve_factored(const Geo &g, const StructMats &sm, const FactorSet &fs, const SolveArgs &args,
                                 int ncomp, size_t cstride, hipStream_t s) {
  SolveArgs a = args;
  // ncomp > 0: the Dirichlet blocks only, for ncomp fields cstride apart
  const int spb0 = ncomp > 0 ? 2 : 0, nf = ncomp > 0 ? ncomp : 1;
  // every level the kernel reads must be set
  if (a.nx < 1 || a.nx > 3 || a.nn < 0 || a.nn > 3 || !a.g_out) return hipErrorInvalidValue;
  for (int l = 0; l < a.nx; ++l)
    if (!a.g[l] || (!ncomp && (!a.u0[l] || !a.v[l] || !a.w[l]))) return hipErrorInvalidValue;
  for (int l = 0; l < a.nn; ++l)
    if (!a.Ng[l] || (!ncomp && (!a.Nu[l] || !a.Nv[l] || !a.Nw[l]))) return hipErrorInvalidValue;
  if (!ncomp && (!a.u0_out || !a.v_out || !a.w_out)) return hipErrorInvalidValue;
  // the kernel reads three levels of each kind: unset ones repeat level 0 of
  // the solutions with a zero coefficient
  for (int l = a.nx; l < 3; ++l) {
    a.u0[l] = a.u0[0];
    a.g[l] = a.g[0];
    a.v[l] = a.v[0];
    a.w[l] = a.w[0];
    a.c[l] = 0.0;
  }
  for (int l = a.nn; l < 3; ++l) {
    a.Nu[l] = a.u0[0];
    a.Ng[l] = a.g[0];
    a.Nv[l] = a.v[0];
    a.Nw[l] = a.w[0];
    a.e[l] = 0.0;
  }
  // the most lanes per block whose rows and factors fit the default 64 KB of LDS
#define PCF_FSOLVE(SLV)                                                                                  \
  hipLaunchKernelGGL(k_solve_factored<SLV>, dim3((4 - spb0) * ((g.m + SLV - 1) / SLV), nf), dim3(256),    \
                     factored_lds_bytes(g, SLV), s, g, sm, fs, a, spb0, cstride)
  if (factored_lds_bytes(g, 16) <= 65536) PCF_FSOLVE(16);
  else if (factored_lds_bytes(g, 8) <= 65536) PCF_FSOLVE(8);
  else return hipErrorInvalidValue;
#undef PCF_FSOLVE
  return hipGetLastError();
}

hipError_t launch_adj_solve_factored(const Geo &g, const StructMats &smt, const FactorSet &fs, PrimC in,
                                     Prim out, Prim nu, PrimC r1, PrimC r2, double e0, double e1, double e2,
                                     hipStream_t s, Prim sig, double c0, double c1, double c2) {
  return launch_adj_solve_factored_dir(g, smt, fs, in, out, nu, r1, r2, e0, e1, e2, s, sig, c0, c1, c2, 0, 0);
}

hipError_t launch_adj_solve_factored_dir(const Geo &g, const StructMats &smt, const FactorSet &fs, PrimC in,
                                         Prim out, Prim nu, PrimC r1, PrimC r2, double e0, double e1, double e2,
                                         hipStream_t s, Prim sig, double c0, double c1, double c2, int ncomp,
                                         size_t cstride) {
  const int spb0 = ncomp > 0 ? 2 : 0, nf = ncomp > 0 ? ncomp : 1;
  auto lds = [&](int sl) { return (size_t)g.npmax * (2 + NF_B) * (sl + 1) * sizeof(double); };
#define PCF_ASOLVE(SLV)                                                                                  \
  hipLaunchKernelGGL(k_adj_solve_factored<SLV>, dim3((4 - spb0) * ((g.m + SLV - 1) / SLV), nf), dim3(256), \
                     lds(SLV), s, g, smt, fs, in, out, nu, r1, r2, e0, e1, e2, sig, c0, c1, c2, spb0, cstride)
  if (lds(16) <= 65536) PCF_ASOLVE(16);
  else if (lds(8) <= 65536) PCF_ASOLVE(8);
  else return hipErrorInvalidValue;
#undef PCF_ASOLVE
  return hipGetLastError();
}

}  // namespace pcf
