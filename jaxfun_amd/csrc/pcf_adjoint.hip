// Reverse-mode kernels of the plane-Couette KMM SBDF3 step (gfx950).
//
// The reference differentiates its step with jax reverse mode
// (production/autograd_cost_benchmark.py:228-330); these kernels are the
// hand-written transpose of each forward kernel:
//   k_adj_solve      S^{-T} per mode: the structured O(n) solver run on the
//                    transposed, index-reversed parity blocks (a far-lower
//                    low-rank part becomes far-upper under index reversal)
//   k_adj_mass       (M + gamma L)^T and the SBDF3 history gather M^T sum c_l rho_l
//   k_recon_adj      transpose of the velocity reconstruction
//   k_rot_adj        cross-product adjoint between the spanwise FFTs
//   k_energy_seed    cotangent of the perturbation energy
// The transposed wall-normal contractions (nonlinear rows, transform of u and
// curl u) run on the matrix cores (pcf_mfma.hip).
#include "pcf_adjoint.hpp"
#include "pcf_device.hpp"

#include <algorithm>

namespace pcf {
namespace {


__device__ __forceinline__ bool nyquist_row(const Geo &g, int m) { return m / g.kzs == g.ny / 2; }

// Banded + far-upper low-rank system (the StructSys form) solved without
// pivoting for an explicit right-hand side: a top-down elimination that keeps
// each row's far coefficients, then a bottom-up back substitution with running
// suffix sums of the far part.  Row data between the sweeps: scratch laid out
// [row][value][lane].
template <int P, int Q, int R, bool HAS4, class Load, class Store>
__device__ void band_lowrank_solve(const StructSys &ss, int n, int lane, int M, double s0, double s2,
                                   double s4, double sf, double *scr, Load load, Store store) {
  constexpr int F = Q + 1;
  constexpr int BW = P + Q + 1;
  constexpr int NV = 1 + Q + R + 2;
  const size_t rowstride = (size_t)NV * M;
  auto at = [&](int i, int v) -> double & { return scr[(size_t)i * rowstride + (size_t)v * M + lane]; };
  {
    double pd[P + 1], pe[P + 1][Q + 1], pw[P + 1][R];
    double2 py[P + 1];
#pragma unroll
    for (int t = 0; t <= P; ++t) {
      pd[t] = 0;
      py[t] = double2{0, 0};
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
      double wi[R];
#pragma unroll
      for (int r = 0; r < R; ++r) wi[r] = sf * ss.U[(size_t)i * R + r];
      double2 yi = load(i);
#pragma unroll
      for (int t = P; t >= 1; --t) {
        if (i - t < 0) continue;
        const double l = a[P - t] * pd[t];
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
        yi.x -= l * py[t].x;
        yi.y -= l * py[t].y;
      }
      const double dinv = 1.0 / a[P];
      at(i, 0) = dinv;
#pragma unroll
      for (int q = 1; q <= Q; ++q) at(i, q) = a[P + q];
#pragma unroll
      for (int r = 0; r < R; ++r) at(i, 1 + Q + r) = wi[r];
      at(i, NV - 2) = yi.x;
      at(i, NV - 1) = yi.y;
#pragma unroll
      for (int t = P; t >= 2; --t) {
        pd[t] = pd[t - 1];
        py[t] = py[t - 1];
#pragma unroll
        for (int q = 0; q <= Q; ++q) pe[t][q] = pe[t - 1][q];
#pragma unroll
        for (int r = 0; r < R; ++r) pw[t][r] = pw[t - 1][r];
      }
      pd[1] = dinv;
      py[1] = yi;
#pragma unroll
      for (int q = 1; q <= Q; ++q) pe[1][q] = a[P + q];
#pragma unroll
      for (int r = 0; r < R; ++r) pw[1][r] = wi[r];
    }
  }
  {
    double2 xw[F + 1];
#pragma unroll
    for (int k = 0; k <= F; ++k) xw[k] = double2{0, 0};
    double2 tau[R];
#pragma unroll
    for (int r = 0; r < R; ++r) tau[r] = double2{0, 0};
    for (int i = n - 1; i >= 0; --i) {
      if (i + F < n) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double v = ss.V[(size_t)(i + F) * R + r];
          tau[r].x += v * xw[F].x;
          tau[r].y += v * xw[F].y;
        }
      }
      double2 acc{at(i, NV - 2), at(i, NV - 1)};
#pragma unroll
      for (int q = 1; q <= Q; ++q) {
        const double e = at(i, q);
        acc.x -= e * xw[q].x;
        acc.y -= e * xw[q].y;
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double w = at(i, 1 + Q + r);
        acc.x -= w * tau[r].x;
        acc.y -= w * tau[r].y;
      }
      const double dinv = at(i, 0);
      const double2 xi{acc.x * dinv, acc.y * dinv};
      store(i, xi);
#pragma unroll
      for (int k = F; k >= 2; --k) xw[k] = xw[k - 1];
      xw[1] = xi;
    }
  }
}

// Thread map shared by the per-mode kernels: threads [0, 4 Mp) are
// (system, parity) = t / Mp (system 0 clamped u0, 1 Dirichlet g) over modes;
// threads [4 Mp, 4 Mp + 2) are the two parities of the mean system S00.
struct SysLane {
  int sys, par, m;
  bool mean, valid;
};
__device__ __forceinline__ SysLane sys_lane(const Geo &g, int t) {
  const int M = g.m, Mp = (M + 63) / 64 * 64;
  SysLane s{};
  int sp = t / Mp, m = t - sp * Mp;
  s.valid = true;
  if (sp >= 4) {
    if (sp > 4 || m >= 2) s.valid = false;
    s.mean = true;
    s.sys = 1;
    s.par = m;
    s.m = 0;
    return s;
  }
  if (m >= M) s.valid = false;
  s.mean = false;
  s.sys = sp >> 1;
  s.par = sp & 1;
  s.m = m;
  return s;
}

// rho = S^{-T} mu, S = M - gamma L, on the transposed index-reversed blocks.
// With nu.u0 set, also nu = e0 rho + e1 r1 + e2 r2 (the cotangent of the
// step's nonlinear rows, r1/r2 the two later rho).  ncomp > 0: the Dirichlet
// blocks only, for gridDim.y fields whose g pointers advance by cstride per
// field, each with its own two parity slices of scratch.
__global__ void __launch_bounds__(256) k_adj_solve(Geo g, StructMats smt, double gamma, PrimC in,
                                                   Prim out, double *scratch, Prim nuo, PrimC r1,
                                                   PrimC r2, double e0, double e1, double e2, int ncomp,
                                                   size_t cstride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  SysLane L{};
  if (ncomp > 0) {
    const int Mp = (g.m + 63) / 64 * 64;
    L.par = t / Mp;
    L.m = t - L.par * Mp;
    L.sys = 1;
    L.mean = false;
    L.valid = L.par < 2 && L.m < g.m;
    const size_t off = blockIdx.y * cstride;
    in.g += off;
    out.g += off;
    if (nuo.g) nuo.g += off;
    if (r1.g) r1.g += off;
    if (r2.g) r2.g += off;
  } else {
    L = sys_lane(g, t);
  }
  if (!L.valid) return;
  const int M = g.m, m = L.m, par = L.par;
  const StructSys &ss = smt.sys[L.sys][par];
  const int n = ss.n;
  const int sp = L.mean ? 2 + par : 2 * L.sys + par;
  double *scr = L.mean      ? scratch + (size_t)4 * g.npmax * 7 * M + (size_t)par * g.npmax * 7
                : ncomp > 0 ? scratch + ((size_t)blockIdx.y * 2 + par) * g.npmax * 7 * M
                            : scratch + (size_t)sp * g.npmax * 7 * M;
  const int Ms = L.mean ? 1 : M;
  const int lane = L.mean ? 0 : m;
  const bool wnu = nuo.u0 != nullptr;
  if (!L.mean && nyquist_row(g, m)) {
    double2 *dst = (L.sys == 0) ? out.u0 : out.g;
    double2 *dn = (L.sys == 0) ? nuo.u0 : nuo.g;
    for (int i = 0; i < n; ++i) {
      dst[(size_t)(2 * i + par) * M + m] = double2{0, 0};
      if (wnu) dn[(size_t)(2 * i + par) * M + m] = double2{0, 0};
    }
    return;
  }
  auto comb = [&](double2 x, const double2 *a, const double2 *b, size_t o) {
    const double2 p = a[o], q = b[o];
    return double2{e0 * x.x + e1 * p.x + e2 * q.x, e0 * x.y + e1 * p.y + e2 * q.y};
  };
  double kyd, kzd, kyo;
  mode_k(g, m, kyd, kzd, kyo);
  const double k2 = L.mean ? 0.0 : kyd * kyd + kzd * kzd;
  const double W = L.mean ? 1.0 : g.W, nu = g.nu, gm = gamma;
  auto row = [&](int ir) { return 2 * (n - 1 - ir) + par; };
  if (L.sys == 0) {
    const double s2 = W * (1 + 2 * gm * nu * k2), s4 = -W * gm * nu, s0 = -W * (k2 + gm * nu * k2 * k2);
    band_lowrank_solve<2, 2, 2, true>(
        ss, n, lane, Ms, s0, s2, s4, s4, scr,
        [&](int ir) { return in.u0[(size_t)row(ir) * M + m]; },
        [&](int ir, double2 x) {
          const size_t o = (size_t)row(ir) * M + m;
          out.u0[o] = x;
          if (wnu) nuo.u0[o] = comb(x, r1.u0, r2.u0, o);
        });
  } else {
    const double s0 = W * (1 + gm * nu * k2), s2 = -W * gm * nu;
    if (L.mean) {
      band_lowrank_solve<1, 1, 1, false>(
          ss, n, lane, Ms, s0, s2, 0.0, s2, scr,
          [&](int ir) { return double2{in.v[row(ir)], in.w[row(ir)]}; },
          [&](int ir, double2 x) {
            const int j = row(ir);
            out.v[j] = x.x;
            out.w[j] = x.y;
            if (wnu) {
              nuo.v[j] = e0 * x.x + e1 * r1.v[j] + e2 * r2.v[j];
              nuo.w[j] = e0 * x.y + e1 * r1.w[j] + e2 * r2.w[j];
            }
          });
    } else {
      band_lowrank_solve<1, 1, 1, false>(
          ss, n, lane, Ms, s0, s2, 0.0, s2, scr,
          [&](int ir) { return in.g[(size_t)row(ir) * M + m]; },
          [&](int ir, double2 x) {
            const size_t o = (size_t)row(ir) * M + m;
            out.g[o] = x;
            if (wnu) nuo.g[o] = comb(x, r1.g, r2.g, o);
          });
    }
  }
}

// out[j] += sum_i A[i][j] rs[i] + sum_i Mb[i][j] rg[i] over one parity block:
// A = t0 C0 + t2 C2 + t4 C4 with far part tf U V^T, Mb = m0 C0 + m2 C2 (banded).
template <int P, int Q, int R, bool HAS4, class LoadS, class LoadG, class Acc>
__device__ void band_transpose_apply(const StructSys &ss, int n, bool has_stage, double t0, double t2,
                                     double t4, double tf, double m0, double m2, LoadS ls, LoadG lg,
                                     Acc acc) {
  constexpr int BW = P + Q + 1;
  double2 tau[R];
#pragma unroll
  for (int r = 0; r < R; ++r) tau[r] = double2{0, 0};
  for (int j = 0; j < n; ++j) {
    double2 s{0, 0};
    if (has_stage) {
      const int ip = j - Q - 1;
      if (ip >= 0) {
        const double2 v = ls(ip);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double u = ss.U[(size_t)ip * R + r];
          tau[r].x += u * v.x;
          tau[r].y += u * v.y;
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double vj = tf * ss.V[(size_t)j * R + r];
        s.x += vj * tau[r].x;
        s.y += vj * tau[r].y;
      }
    }
#pragma unroll
    for (int d = -P; d <= Q; ++d) {
      const int i = j - d;  // A[i][j] sits at offset d = j - i of row i
      if (i < 0 || i >= n) continue;
      const size_t o = (size_t)i * BW + (d + P);
      const double c0 = ss.b0[o], c2 = ss.b2[o], c4 = HAS4 ? ss.b4[o] : 0.0;
      if (has_stage) {
        const double a = t0 * c0 + t2 * c2 + t4 * c4;
        const double2 v = ls(i);
        s.x += a * v.x;
        s.y += a * v.y;
      }
      const double b = m0 * c0 + m2 * c2;
      if (b != 0.0) {
        const double2 v = lg(i);
        s.x += b * v.x;
        s.y += b * v.y;
      }
    }
    acc(j, s);
  }
}

__global__ void __launch_bounds__(256) k_adj_mass(Geo g, StructMats sm, double gamma, PrimC rs,
                                                  PrimSum gat, Prim out) {
  const SysLane L = sys_lane(g, blockIdx.x * blockDim.x + threadIdx.x);
  if (!L.valid) return;
  const int M = g.m, m = L.m, par = L.par;
  const StructSys &ss = sm.sys[L.sys][par];
  const int n = ss.n;
  const bool has_stage = rs.u0 != nullptr;
  double kyd, kzd, kyo;
  mode_k(g, m, kyd, kzd, kyo);
  const double k2 = L.mean ? 0.0 : kyd * kyd + kzd * kzd;
  const double W = L.mean ? 1.0 : g.W, nu = g.nu, gm = gamma;
  auto J = [&](int i) { return 2 * i + par; };
  if (L.sys == 0) {
    // M = W (C2 - k2 C0);  M + gamma L = W[(1 - 2 gm nu k2) C2 + gm nu C4 + (-k2 + gm nu k2^2) C0]
    const double t4 = W * gm * nu, t2 = W * (1 - 2 * gm * nu * k2), t0 = W * (-k2 + gm * nu * k2 * k2);
    band_transpose_apply<2, 2, 2, true>(
        ss, n, has_stage, t0, t2, t4, t4, -W * k2, W,
        [&](int i) { return rs.u0[(size_t)J(i) * M + m]; },
        [&](int i) {
          double2 s{0, 0};
          for (int t = 0; t < gat.n; ++t) {
            const double2 v = gat.t[t].u0[(size_t)J(i) * M + m];
            s.x += gat.wgt[t] * v.x;
            s.y += gat.wgt[t] * v.y;
          }
          return s;
        },
        [&](int j, double2 s) {
          double2 &o = out.u0[(size_t)J(j) * M + m];
          o.x += s.x;
          o.y += s.y;
        });
  } else {
    // M = W C0;  M + gamma L = W[(1 - gm nu k2) C0 + gm nu C2]
    const double t0 = W * (1 - gm * nu * k2), t2 = W * gm * nu;
    if (L.mean) {
      band_transpose_apply<1, 1, 1, false>(
          ss, n, has_stage, t0, t2, 0.0, t2, W, 0.0,
          [&](int i) { return double2{rs.v[J(i)], rs.w[J(i)]}; },
          [&](int i) {
            double2 s{0, 0};
            for (int t = 0; t < gat.n; ++t) {
              s.x += gat.wgt[t] * gat.t[t].v[J(i)];
              s.y += gat.wgt[t] * gat.t[t].w[J(i)];
            }
            return s;
          },
          [&](int j, double2 s) {
            out.v[J(j)] += s.x;
            out.w[J(j)] += s.y;
          });
    } else {
      band_transpose_apply<1, 1, 1, false>(
          ss, n, has_stage, t0, t2, 0.0, t2, W, 0.0,
          [&](int i) { return rs.g[(size_t)J(i) * M + m]; },
          [&](int i) {
            double2 s{0, 0};
            for (int t = 0; t < gat.n; ++t) {
              const double2 v = gat.t[t].g[(size_t)J(i) * M + m];
              s.x += gat.wgt[t] * v.x;
              s.y += gat.wgt[t] * v.y;
            }
            return s;
          },
          [&](int j, double2 s) {
            double2 &o = out.g[(size_t)J(j) * M + m];
            o.x += s.x;
            o.y += s.y;
          });
    }
  }
}

// Transpose of f = R u0, u1 = i (a f + b g), u2 = i (b f - a g) (a = ky/k^2,
// b = kz/k^2), mean (0,0) <- (v00, w00), Nyquist row masked.  R couples
// opposite parities; RcT[j][k] = R[2k + 1 - (j & 1)][j].
__global__ void __launch_bounds__(256) k_recon_adj(Geo g, const double *__restrict__ RcT, int ldk,
                                                   const double2 *__restrict__ mu0,
                                                   const double2 *__restrict__ mu1,
                                                   const double2 *__restrict__ mu2,
                                                   const double2 *__restrict__ g_extra, Prim out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (m >= g.m) return;
  const size_t o = (size_t)j * g.m + m;
  const bool nyq = nyquist_row(g, m);
  double kyd, kzd, kyo;
  mode_k(g, m, kyd, kzd, kyo);
  const double k2 = kyd * kyd + kzd * kzd;
  const double den = (k2 == 0.0) ? 1.0 : k2;
  const double a = kyd / den, b = kzd / den;
  if (j < g.nb) {
    double2 acc = nyq ? double2{0, 0} : mu0[o];
    if (m != 0 && !nyq) {
      const int i0 = 1 - (j & 1);
      const int nk = (g.nd - i0 + 1) / 2;
      const double *r = RcT + (size_t)j * ldk;
      for (int k = 0; k < nk; ++k) {
        const size_t oi = (size_t)(i0 + 2 * k) * g.m + m;
        const double2 p = mu1[oi], q = mu2[oi];
        // mu_f = -i (a mu1 + b mu2)
        const double2 s{a * p.x + b * q.x, a * p.y + b * q.y};
        acc.x += r[k] * s.y;
        acc.y -= r[k] * s.x;
      }
    }
    out.u0[o] = acc;
  }
  if (j < g.nd) {
    double2 gv{0, 0};
    if (m != 0 && !nyq) {
      // mu_g = -i (b mu1 - a mu2)
      const double2 p = mu1[o], q = mu2[o];
      const double2 s{b * p.x - a * q.x, b * p.y - a * q.y};
      gv = double2{s.y, -s.x};
    }
    if (g_extra && !nyq) {
      gv.x += g_extra[o].x;
      gv.y += g_extra[o].y;
    }
    out.g[o] = gv;
    if (m == 0) {
      out.v[j] = mu1[o].x;
      out.w[j] = mu2[o].x;
    }
  }
}

// Cross-product adjoint.  LDS layout [field][unit][MZ] (field-major, so the
// six outputs of all units form one uniformly strided FFT batch).  Fields 0-5
// are the forward u, omega; 6-8 the cotangent of H.  With a = u + (0, U x, 0),
// b = omega + (0, 0, U) and c = -mu_H / (MY MZ) on the grid:
//   mu_u = b x c,  mu_omega = c x a.
// NZC > 0: padded spanwise length fixed at compile time (0: from g).
template <typename T, int MAXB, int NZC>
__global__ void __launch_bounds__(256) k_rot_adj(Geo g, NodalMats<T> nm, FFTPlan plan, int UPB,
                                                 const typename VT<T>::V *__restrict__ Yf,
                                                 const typename VT<T>::V *__restrict__ Hz,
                                                 typename VT<T>::V *__restrict__ Yout) {
  using V = typename VT<T>::V;
  extern __shared__ __align__(16) unsigned char smem[];
  V *buf = reinterpret_cast<V *>(smem);  // [9][UPB][MZ]
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int MZ = NZC ? NZC : g.mz, KZ = g.kzs, MY = g.my;
  const int nunits = g.n0 * (MY / 2);
  const int ubase = blockIdx.x * UPB;
  const size_t fstride = (size_t)g.n0 * MY * KZ;
  V *tw = buf + (size_t)9 * UPB * MZ;  // twiddle table in LDS
  for (int i = tid; i < MZ; i += nthr) tw[i] = reinterpret_cast<const V *>(nm.twz)[i];

  gather_batched<4>(
      9 * UPB * MZ,
      [&](int idx) {
        const int f = idx / (UPB * MZ), rem = idx - f * UPB * MZ;
        const int lu = rem / MZ, k = rem - lu * MZ;
        const int u = ubase + lu;
        V z{0, 0};
        if (u < nunits) {
          const int x = u / (MY / 2), y2 = u - x * (MY / 2);
          const V *src = (f < 6) ? Yf + f * fstride : Hz + (f - 6) * fstride;
          const V *ra = src + ((size_t)x * MY + 2 * y2) * KZ;
          const V *rb = ra + KZ;
          if (k == 0) {
            z = V{ra[0].x, rb[0].x};
          } else if (k < KZ) {
            const V a = ra[k], b = rb[k];
            z = V{a.x - b.y, a.y + b.x};
          } else if (k > MZ - KZ) {
            const V a = ra[MZ - k], b = rb[MZ - k];
            z = V{a.x + b.y, b.x - a.y};
          }
        }
        return z;
      },
      [&](int idx, const V &z) { buf[idx] = z; });
  __syncthreads();
  if constexpr (NZC > 0)
    fft_lds_c<V, T, NZC, true, false, MAXB>(buf, tw, 9 * UPB, MZ, 1);
  else
    fft_lds<V, T, true, false, MAXB>(buf, tw, plan, 9 * UPB, MZ, 1);

  const T U = (T)g.uwall;
  const T sc = (T)(-1.0 / ((double)MY * (double)MZ));
  const size_t fs = (size_t)UPB * MZ;
  for (int idx = tid; idx < UPB * MZ; idx += nthr) {
    const int lu = idx / MZ;
    const int u = ubase + lu;
    const int x = (u < nunits) ? u / (MY / 2) : 0;
    const T xv = nm.xnode[x];
    V *p = buf + idx;
    V r[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) r[f] = p[f * fs];
    V o[6];
    // both packed lines (real and imaginary parts) in turn
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      auto get = [&](const V &v) { return part ? v.y : v.x; };
      const T a0 = get(r[0]), a1 = get(r[1]) + U * xv, a2 = get(r[2]);
      const T b0 = get(r[3]), b1 = get(r[4]), b2 = get(r[5]) + U;
      const T c0 = sc * get(r[6]), c1 = sc * get(r[7]), c2 = sc * get(r[8]);
      const T out[6] = {b1 * c2 - b2 * c1, b2 * c0 - b0 * c2, b0 * c1 - b1 * c0,
                        c1 * a2 - c2 * a1, c2 * a0 - c0 * a2, c0 * a1 - c1 * a0};
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        if (part)
          o[f].y = out[f];
        else
          o[f].x = out[f];
      }
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) p[f * fs] = o[f];
  }
  __syncthreads();
  if constexpr (NZC > 0)
    fft_lds_c<V, T, NZC, false, false, MAXB>(buf, tw, 6 * UPB, MZ, 1);
  else
    fft_lds<V, T, false, false, MAXB>(buf, tw, plan, 6 * UPB, MZ, 1);

  for (int idx = tid; idx < 6 * UPB * KZ; idx += nthr) {
    const int f = idx / (UPB * KZ), rem = idx - f * UPB * KZ;
    const int lu = rem / KZ, k = rem - lu * KZ;
    const int u = ubase + lu;
    if (u >= nunits) continue;
    const int x = u / (MY / 2), y2 = u - x * (MY / 2);
    const V *P = buf + ((size_t)f * UPB + lu) * MZ;
    const V p = P[k], q = cconj(P[(MZ - k) % MZ]);
    const V ha{(p.x + q.x) * (T)0.5, (p.y + q.y) * (T)0.5};
    const V hb{(p.y - q.y) * (T)0.5, -(p.x - q.x) * (T)0.5};
    V *dst = Yout + f * fstride + ((size_t)x * MY + 2 * y2) * KZ;
    dst[k] = ha;
    dst[KZ + k] = hb;
  }
}

// Cross-product adjoint of the MHD forcing and EMF (k_rotational_mhd):
// fields 0-11 of the LDS are the forward u, omega, b, j of one padded y-line
// pair, 12-14 the cotangent of H and 15-17 that of the EMF E.  With
// a = u + (0, U x, 0), o = omega + (0, 0, U), bt = b + B0, c = -mu_H / (MY MZ)
// and d = mu_E / (MY MZ), the forward H = -(a x o) + (-2 Omega u1, 2 Omega u0, 0)
// - (j x bt), E = a x bt gives
//   mu_u = o x c + bt x d + 2 Omega (-c1, c0, 0),  mu_omega = c x a,
//   mu_b = c x j + d x a,                          mu_j = bt x c.
template <typename T, int MAXB, int NZC>
__global__ void __launch_bounds__(256) k_rot_adj_mhd(Geo g, NodalMats<T> nm, FFTPlan plan,
                                                     const typename VT<T>::V *__restrict__ Yf,
                                                     const typename VT<T>::V *__restrict__ Hz,
                                                     typename VT<T>::V *__restrict__ Yout) {
  using V = typename VT<T>::V;
  constexpr int NIN = 18, NOUT = 12;
  extern __shared__ __align__(16) unsigned char smem[];
  V *buf = reinterpret_cast<V *>(smem);  // [NIN][MZ]
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int MZ = NZC ? NZC : g.mz, KZ = g.kzs, MY = g.my;
  const int nunits = g.n0 * (MY / 2);
  const int u = blockIdx.x;
  const size_t fstride = (size_t)g.n0 * MY * KZ;
  V *tw = buf + (size_t)NIN * MZ;
  for (int i = tid; i < MZ; i += nthr) tw[i] = reinterpret_cast<const V *>(nm.twz)[i];
  const int x = u / (MY / 2), y2 = u - x * (MY / 2);
  gather_batched<4>(
      NIN * MZ,
      [&](int idx) {
        const int f = idx / MZ, k = idx - f * MZ;
        V z{0, 0};
        if (u < nunits) {
          const V *src = (f < 12) ? Yf + f * fstride : Hz + (f - 12) * fstride;
          const V *ra = src + ((size_t)x * MY + 2 * y2) * KZ;
          const V *rb = ra + KZ;
          if (k == 0) {
            z = V{ra[0].x, rb[0].x};
          } else if (k < KZ) {
            const V a = ra[k], b = rb[k];
            z = V{a.x - b.y, a.y + b.x};
          } else if (k > MZ - KZ) {
            const V a = ra[MZ - k], b = rb[MZ - k];
            z = V{a.x + b.y, b.x - a.y};
          }
        }
        return z;
      },
      [&](int idx, const V &z) { buf[idx] = z; });
  __syncthreads();
  if constexpr (NZC > 0)
    fft_lds_c<V, T, NZC, true, false, MAXB>(buf, tw, NIN, MZ, 1);
  else
    fft_lds<V, T, true, false, MAXB>(buf, tw, plan, NIN, MZ, 1);

  const T U = (T)g.uwall, O2 = (T)g.omega2;
  const T B0x = (T)g.b0[0], B0y = (T)g.b0[1], B0z = (T)g.b0[2];
  const T sc = (T)(1.0 / ((double)MY * (double)MZ));
  const T xv = nm.xnode[u < nunits ? x : 0];
  for (int n = tid; n < MZ; n += nthr) {
    V r[NIN];
#pragma unroll
    for (int f = 0; f < NIN; ++f) r[f] = buf[f * MZ + n];
    V o[NOUT];
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      auto get = [&](const V &v) { return part ? v.y : v.x; };
      const T a0 = get(r[0]), a1 = get(r[1]) + U * xv, a2 = get(r[2]);
      const T w0 = get(r[3]), w1 = get(r[4]), w2 = get(r[5]) + U;
      const T t0 = get(r[6]) + B0x, t1 = get(r[7]) + B0y, t2 = get(r[8]) + B0z;
      const T j0 = get(r[9]), j1 = get(r[10]), j2 = get(r[11]);
      const T c0 = -sc * get(r[12]), c1 = -sc * get(r[13]), c2 = -sc * get(r[14]);
      const T d0 = sc * get(r[15]), d1 = sc * get(r[16]), d2 = sc * get(r[17]);
      const T out[NOUT] = {w1 * c2 - w2 * c1 + t1 * d2 - t2 * d1 - O2 * c1,
                           w2 * c0 - w0 * c2 + t2 * d0 - t0 * d2 + O2 * c0,
                           w0 * c1 - w1 * c0 + t0 * d1 - t1 * d0,
                           c1 * a2 - c2 * a1, c2 * a0 - c0 * a2, c0 * a1 - c1 * a0,
                           c1 * j2 - c2 * j1 + d1 * a2 - d2 * a1,
                           c2 * j0 - c0 * j2 + d2 * a0 - d0 * a2,
                           c0 * j1 - c1 * j0 + d0 * a1 - d1 * a0,
                           t1 * c2 - t2 * c1, t2 * c0 - t0 * c2, t0 * c1 - t1 * c0};
#pragma unroll
      for (int f = 0; f < NOUT; ++f) {
        if (part)
          o[f].y = out[f];
        else
          o[f].x = out[f];
      }
    }
#pragma unroll
    for (int f = 0; f < NOUT; ++f) buf[f * MZ + n] = o[f];
  }
  __syncthreads();
  if constexpr (NZC > 0)
    fft_lds_c<V, T, NZC, false, false, MAXB>(buf, tw, NOUT, MZ, 1);
  else
    fft_lds<V, T, false, false, MAXB>(buf, tw, plan, NOUT, MZ, 1);
  if (u >= nunits) return;
  for (int idx = tid; idx < NOUT * KZ; idx += nthr) {
    const int f = idx / KZ, k = idx - f * KZ;
    const V *P = buf + (size_t)f * MZ;
    const V p = P[k], q = cconj(P[(MZ - k) % MZ]);
    const V ha{(p.x + q.x) * (T)0.5, (p.y + q.y) * (T)0.5};
    const V hb{(p.y - q.y) * (T)0.5, -(p.x - q.x) * (T)0.5};
    V *dst = Yout + f * fstride + ((size_t)x * MY + 2 * y2) * KZ;
    dst[k] = ha;
    dst[KZ + k] = hb;
  }
}

__global__ void __launch_bounds__(256) k_energy_seed(Geo g, const double *__restrict__ QB,
                                                     const double *__restrict__ QD, double scale,
                                                     const double2 *__restrict__ u0,
                                                     const double2 *__restrict__ u1,
                                                     const double2 *__restrict__ u2,
                                                     double2 *__restrict__ mu0,
                                                     double2 *__restrict__ mu1,
                                                     double2 *__restrict__ mu2) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (m >= g.m) return;
  const bool nyq = nyquist_row(g, m);
  const size_t o = (size_t)i * g.m + m;
  if (i < g.nb) {
    double2 s{0, 0};
    if (!nyq)
      for (int j = 0; j < g.nb; ++j) {
        const double q = QB[(size_t)i * g.nb + j];
        const double2 c = u0[(size_t)j * g.m + m];
        s.x += q * c.x;
        s.y += q * c.y;
      }
    mu0[o] = double2{scale * s.x, scale * s.y};
  }
  double2 s1{0, 0}, s2{0, 0};
  if (!nyq)
    for (int j = 0; j < g.nd; ++j) {
      const double q = QD[(size_t)i * g.nd + j];
      const double2 c = u1[(size_t)j * g.m + m], d = u2[(size_t)j * g.m + m];
      s1.x += q * c.x;
      s1.y += q * c.y;
      s2.x += q * d.x;
      s2.y += q * d.y;
    }
  mu1[o] = double2{scale * s1.x, scale * s1.y};
  mu2[o] = double2{scale * s2.x, scale * s2.y};
}

// out = sum_t w_t in_t over every component of a primary set.
__global__ void __launch_bounds__(256) k_prim_sum(Geo g, PrimSum in, Prim out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (m >= g.m) return;
  const size_t o = (size_t)i * g.m + m;
  double2 su{0, 0}, sg{0, 0};
  double sv = 0, sw = 0;
  for (int t = 0; t < in.n; ++t) {
    const double w = in.wgt[t];
    if (i < g.nb) {
      su.x += w * in.t[t].u0[o].x;
      su.y += w * in.t[t].u0[o].y;
    }
    sg.x += w * in.t[t].g[o].x;
    sg.y += w * in.t[t].g[o].y;
    if (m == 0) {
      sv += w * in.t[t].v[i];
      sw += w * in.t[t].w[i];
    }
  }
  if (i < g.nb) out.u0[o] = su;
  out.g[o] = sg;
  if (m == 0) {
    out.v[i] = sv;
    out.w[i] = sw;
  }
}

int sys_threads(const Geo &g) { return 4 * ((g.m + 63) / 64 * 64) + 64; }

int max_bf(const FFTPlan &plan, int nbatch) {
  int mb = 0;
  for (int s = 0; s < plan.nst; ++s) mb = std::max(mb, (plan.n / plan.radix[s] * nbatch + 255) / 256);
  return mb;
}

template <typename T>
int rot_adj_units(const Geo &g) {
  using V = typename VT<T>::V;
  int upb = 1;  // nine lines per unit already fill the block (fewer butterflies per thread)
  while (upb > 1 && ((size_t)upb * 9 * g.mz + g.mz) * sizeof(V) > 65536) upb >>= 1;
  return upb;
}

}  // namespace

hipError_t launch_adj_solve(const Geo &g, const StructMats &smt, double gamma, PrimC in, Prim out,
                            double *scratch, hipStream_t s, Prim nu, PrimC r1, PrimC r2, double e0,
                            double e1, double e2) {
  hipLaunchKernelGGL(k_adj_solve, dim3((sys_threads(g) + 255) / 256), dim3(256), 0, s, g, smt, gamma,
                     in, out, scratch, nu, r1, r2, e0, e1, e2, 0, (size_t)0);
  return hipGetLastError();
}

size_t adj_solve_dir_scratch_doubles(const Geo &g, int ncomp) {
  return std::max(solve_scratch_doubles(g), (size_t)ncomp * 2 * g.npmax * 7 * g.m);
}

hipError_t launch_adj_solve_dir(const Geo &g, const StructMats &smt, double gamma, const double2 *in,
                                double2 *out, double *scratch, int ncomp, size_t cstride, hipStream_t s) {
  if (ncomp < 1 || !in || !out || !scratch) return hipErrorInvalidValue;
  const int threads = 2 * ((g.m + 63) / 64 * 64);
  PrimC pin{};
  pin.g = in;
  Prim pout{};
  pout.g = out;
  hipLaunchKernelGGL(k_adj_solve, dim3((threads + 255) / 256, ncomp), dim3(256), 0, s, g, smt, gamma, pin, pout,
                     scratch, Prim{}, PrimC{}, PrimC{}, 0.0, 0.0, 0.0, ncomp, cstride);
  return hipGetLastError();
}

hipError_t launch_adj_mass(const Geo &g, const StructMats &sm, double gamma, PrimC r_stage,
                           const PrimSum &gather, Prim out, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_mass, dim3((sys_threads(g) + 255) / 256), dim3(256), 0, s, g, sm, gamma,
                     r_stage, gather, out);
  return hipGetLastError();
}

hipError_t launch_recon_adj(const Geo &g, const double *RcT, int ldk, const double2 *mu0,
                            const double2 *mu1, const double2 *mu2, const double2 *g_extra,
                            Prim out, hipStream_t s) {
  dim3 grid((g.m + 255) / 256, g.nd);
  hipLaunchKernelGGL(k_recon_adj, grid, dim3(256), 0, s, g, RcT, ldk, mu0, mu1, mu2, g_extra, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rot_adj(const Geo &g, const NodalMats<T> &nm, const void *Yf, Workspace &ws,
                          const FFTPlan &pz, hipStream_t s) {
  using V = typename VT<T>::V;
  const int upb = rot_adj_units<T>(g);
  const int nunits = g.n0 * (g.my / 2);
  const size_t lds = ((size_t)upb * 9 * g.mz + g.mz) * sizeof(V);
  const int mb = max_bf(pz, upb * 9);
  const dim3 grid((nunits + upb - 1) / upb);
  const V *yf = reinterpret_cast<const V *>(Yf);
  const V *hz = reinterpret_cast<const V *>(ws.Hz);
  V *yo = reinterpret_cast<V *>(ws.Y);
#define PCF_ROTADJN(MB, NZ) \
  hipLaunchKernelGGL((k_rot_adj<T, MB, NZ>), grid, dim3(256), lds, s, g, nm, pz, upb, yf, hz, yo)
#define PCF_ROTADJ(MB)                            \
  switch (g.mz) {                                 \
    case 24: PCF_ROTADJN(MB, 24); break;          \
    case 48: PCF_ROTADJN(MB, 48); break;          \
    case 96: PCF_ROTADJN(MB, 96); break;          \
    case 144: PCF_ROTADJN(MB, 144); break;        \
    case 192: PCF_ROTADJN(MB, 192); break;        \
    default: PCF_ROTADJN(MB, 0);                  \
  }
  if (mb <= 3) PCF_ROTADJ(3)
  else if (mb <= 6) PCF_ROTADJ(6)
  else if (mb <= 12) PCF_ROTADJ(12)
  else PCF_ROTADJ(18)
#undef PCF_ROTADJ
#undef PCF_ROTADJN
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rot_adj_mhd(const Geo &g, const NodalMats<T> &nm, const void *Yf, Workspace &ws,
                              const FFTPlan &pz, hipStream_t s) {
  using V = typename VT<T>::V;
  const int nunits = g.n0 * (g.my / 2);
  const size_t lds = ((size_t)18 * g.mz + g.mz) * sizeof(V);
  if (lds > 65536) return hipErrorInvalidValue;
  const int mb = max_bf(pz, 18);
  const V *yf = reinterpret_cast<const V *>(Yf);
  const V *hz = reinterpret_cast<const V *>(ws.Hz);
  V *yo = reinterpret_cast<V *>(ws.Y);
#define PCF_ROTADJM_N(MB, NZ) \
  hipLaunchKernelGGL((k_rot_adj_mhd<T, MB, NZ>), dim3(nunits), dim3(256), lds, s, g, nm, pz, yf, hz, yo)
#define PCF_ROTADJM(MB)                           \
  switch (g.mz) {                                 \
    case 24: PCF_ROTADJM_N(MB, 24); break;        \
    case 48: PCF_ROTADJM_N(MB, 48); break;        \
    case 96: PCF_ROTADJM_N(MB, 96); break;        \
    case 144: PCF_ROTADJM_N(MB, 144); break;      \
    case 192: PCF_ROTADJM_N(MB, 192); break;      \
    default: PCF_ROTADJM_N(MB, 0);                \
  }
  if (mb <= 3) PCF_ROTADJM(3)
  else if (mb <= 6) PCF_ROTADJM(6)
  else if (mb <= 12) PCF_ROTADJM(12)
  else if (mb <= 18) PCF_ROTADJM(18)
  else return hipErrorInvalidValue;
#undef PCF_ROTADJM
#undef PCF_ROTADJM_N
  return hipGetLastError();
}
template hipError_t launch_rot_adj_mhd<double>(const Geo &, const NodalMats<double> &, const void *, Workspace &,
                                               const FFTPlan &, hipStream_t);

hipError_t launch_prim_sum(const Geo &g, const PrimSum &in, Prim out, hipStream_t s) {
  dim3 grid((g.m + 255) / 256, g.nd);
  hipLaunchKernelGGL(k_prim_sum, grid, dim3(256), 0, s, g, in, out);
  return hipGetLastError();
}

hipError_t launch_energy_seed(const Geo &g, const double *QB, const double *QD, double scale,
                              const double2 *u0, const double2 *u1, const double2 *u2, double2 *mu0,
                              double2 *mu1, double2 *mu2, hipStream_t s) {
  dim3 grid((g.m + 255) / 256, g.nd);
  hipLaunchKernelGGL(k_energy_seed, grid, dim3(256), 0, s, g, QB, QD, scale, u0, u1, u2, mu0, mu1,
                     mu2);
  return hipGetLastError();
}

bool rot_adj_fits(const Geo &g, const FFTPlan &pz, int elem_bytes) {
  const int upb = (elem_bytes == 16) ? rot_adj_units<double>(g) : rot_adj_units<float>(g);
  if (((size_t)upb * 9 * g.mz + g.mz) * elem_bytes > 65536) return false;
  return max_bf(pz, upb * 9) <= 18;
}

#define PCF_ADJ_INSTANTIATE(T)                                                                     \
  template hipError_t launch_rot_adj<T>(const Geo &, const NodalMats<T> &, const void *,           \
                                        Workspace &, const FFTPlan &, hipStream_t);
PCF_ADJ_INSTANTIATE(double)
PCF_ADJ_INSTANTIATE(float)
#undef PCF_ADJ_INSTANTIATE

}  // namespace pcf
