// Cadence diagnostics of the rotating-shear MHD family on the device
// (reference examples/pcf_mhd_mri_shearpy_jax.py:178-513): Reynolds and
// Maxwell stresses, the mean-field split, the kz = 1 channel amplitudes,
// viscous and resistive dissipation, and the electromagnetic health set
// (ideal / resistive / total electric field, their divergence after the TC
// projection, div A, wall traces, the Faraday tendencies of the mean field).
//
// Layout: "nodal-spectral" arrays are [point][m], point over the n0 Gauss
// nodes followed by the walls x = -1, +1 (P = n0 + 2 rows), m the half
// spectrum iy * kzs + kz.  Quadratic integrals over the (y, z) grid follow from
// the discrete Parseval identity on these arrays (weight 1 for kz = 0, 2 for
// the other stored columns; the spanwise and streamwise Nyquist modes are
// zero); only the electric field u_tot x b_tot, its wall values and the
// maxima need the physical grid, which the small DFT kernels below produce.
// Everything runs in FP64: the state is FP64 in every precision mode.
#include "pcf_diag.hpp"

#include <algorithm>

namespace pcf {

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Streamwise wavenumber of row iy, zero on the Nyquist row (the step's
// contractions use the same value for every wavenumber factor).
__device__ __forceinline__ double ky_of(const Geo &g, int iy) {
  if (g.ny % 2 == 0 && iy == g.ny / 2) return 0.0;
  return (double)(iy < (g.ny + 1) / 2 ? iy : iy - g.ny) * g.df1;
}

// Block-wide sum of v over 256 threads (fixed tree, so the order of the
// additions does not depend on timing).
__device__ double block_sum256(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ double block_max256(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] = fmax(red[t], red[t + s]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

}  // namespace

// ------------------------------------------------------- k_eval_terms ----
// out[p][m] = sum_t f_t(m) sum_i M_t[p][i] F_t[i][m], f_t a mode factor
// (EvalFactor), for blockIdx.z's output of the batch.  A block covers 64
// consecutive modes (one per lane) and EVAL_RW consecutive rows; each of its
// four waves sums a quarter of every term's i range for all EVAL_RW rows, so a
// field value a lane loads serves EVAL_RW rows, the matrix entries are
// wave-uniform loads, and four times as many waves keep loads in flight as
// with whole sums per wave.  The four partial sums meet in LDS in a fixed
// order; wave r stores row p0 + r.
constexpr int EVAL_RW = 4;

__global__ void __launch_bounds__(256) k_eval_terms(Geo g, EvalBatch b) {
  __shared__ double2 part[4][EVAL_RW][64];
  const EvalOut &o = b.o[blockIdx.z];
  const int p0 = blockIdx.y * EVAL_RW;
  if (p0 >= o.rows) return;  // uniform across the block
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int nr = min(EVAL_RW, o.rows - p0);
  const int md = blockIdx.x * 64 + lane;
  const bool live = md < g.m;
  const int mc = live ? md : g.m - 1;  // lanes past the last mode load a valid column
  const int iy = mc / g.kzs, kz = mc - iy * g.kzs;
  const double ky = ky_of(g, iy), kzv = kz * g.df2;
  double ar[EVAL_RW], ai[EVAL_RW];
#pragma unroll
  for (int r = 0; r < EVAL_RW; ++r) ar[r] = ai[r] = 0.0;
  for (int t = 0; t < o.nt; ++t) {
    const EvalTerm &T = o.t[t];
    const int K = T.K, kc = (K + 3) / 4;
    const int i0 = min(K, wave * kc), i1 = min(K, i0 + kc);
    const double *Mr[EVAL_RW];
#pragma unroll
    for (int r = 0; r < EVAL_RW; ++r) Mr[r] = T.M + (size_t)(p0 + min(r, nr - 1)) * K;  // rows past nr: not stored
    const double2 *F = T.F + mc;
    double sr[EVAL_RW], si[EVAL_RW];
#pragma unroll
    for (int r = 0; r < EVAL_RW; ++r) sr[r] = si[r] = 0.0;
    int i = i0;
    for (; i + 4 <= i1; i += 4) {
      double2 f[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) f[u] = F[(size_t)(i + u) * g.m];
#pragma unroll
      for (int r = 0; r < EVAL_RW; ++r)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double a = Mr[r][i + u];
          sr[r] = fma(a, f[u].x, sr[r]);
          si[r] = fma(a, f[u].y, si[r]);
        }
    }
    for (; i < i1; ++i) {
      const double2 f = F[(size_t)i * g.m];
#pragma unroll
      for (int r = 0; r < EVAL_RW; ++r) {
        const double a = Mr[r][i];
        sr[r] = fma(a, f.x, sr[r]);
        si[r] = fma(a, f.y, si[r]);
      }
    }
    if (T.kf == KF_IKY || T.kf == KF_IKZ) {  // c i k (sr + i si) = c k (-si + i sr)
      const double ck = T.c * ((T.kf == KF_IKY) ? ky : kzv);
#pragma unroll
      for (int r = 0; r < EVAL_RW; ++r) {
        ar[r] -= ck * si[r];
        ai[r] += ck * sr[r];
      }
    } else {
      double f = 1.0;
      if (T.kf == KF_KYY) f = ky * ky;
      else if (T.kf == KF_KZZ) f = kzv * kzv;
      else if (T.kf == KF_KYZ) f = ky * kzv;
      else if (T.kf == KF_K2) f = ky * ky + kzv * kzv;
      const double cf = T.c * f;
#pragma unroll
      for (int r = 0; r < EVAL_RW; ++r) {
        ar[r] += cf * sr[r];
        ai[r] += cf * si[r];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < EVAL_RW; ++r) part[wave][r][lane] = make_double2(ar[r], ai[r]);
  __syncthreads();
  static_assert(EVAL_RW <= 4, "one storing wave per row");
  if (wave >= nr || !live) return;
  double2 v = part[0][wave][lane];
#pragma unroll
  for (int w = 1; w < 4; ++w) {
    const double2 q = part[w][wave][lane];
    v.x += q.x;
    v.y += q.y;
  }
  o.dst[(size_t)(p0 + wave) * g.m + md] = v;
}

// ---------------------------------------------------------- k_curl_b ----
// J = curl b at the points from b and its x-derivative:
// Jx = i ky bz - i kz by, Jy = i kz bx - dbz/dx, Jz = dby/dx - i ky bx.
__global__ void __launch_bounds__(256) k_curl_b(Geo g, int P, const double2 *__restrict__ b,
                                                const double2 *__restrict__ db, double2 *__restrict__ J) {
  const int p = blockIdx.y;
  const int md = blockIdx.x * 256 + threadIdx.x;
  if (md >= g.m || p >= P) return;
  const size_t f = (size_t)P * g.m, at = (size_t)p * g.m + md;
  const int iy = md / g.kzs, kz = md - iy * g.kzs;
  const double2 ik1 = make_double2(0.0, ky_of(g, iy)), ik2 = make_double2(0.0, kz * g.df2);
  const double2 bx = b[at], by = b[f + at], bz = b[2 * f + at];
  const double2 dby = db[f + at], dbz = db[2 * f + at];
  const double2 a = cmul(ik1, bz), c = cmul(ik2, by), d = cmul(ik2, bx), e = cmul(ik1, bx);
  J[at] = make_double2(a.x - c.x, a.y - c.y);
  J[f + at] = make_double2(d.x - dbz.x, d.y - dbz.y);
  J[2 * f + at] = make_double2(dby.x - e.x, dby.y - e.y);
}

// ------------------------------------------------------------ k_dft_y ----
// Streamwise DFT of every stored spanwise column of `planes` planes
// [plane][row][kzs], row = iy (spectral) or y (physical), ny rows:
//   INV:  out[y][kz]  = sum_iy in[iy][kz] e^{+2 pi i iy y / ny}
//   !INV: out[iy][kz] = (1/ny) sum_y in[y][kz] e^{-2 pi i iy y / ny}, with the
//         Nyquist row iy = ny/2 set to zero (mask_nyquist).
// tw[t] = e^{2 pi i t / ny}.
template <bool INV>
__global__ void __launch_bounds__(256) k_dft_y(Geo g, const double2 *__restrict__ tw, const double2 *__restrict__ in,
                                               double2 *__restrict__ out) {
  const int plane = blockIdx.y;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= g.m) return;
  const int r = idx / g.kzs, kz = idx - r * g.kzs;
  const double2 *src = in + (size_t)plane * g.m + kz;
  double sr = 0.0, si = 0.0;
  if (INV || !(g.ny % 2 == 0 && r == g.ny / 2)) {
    int t = 0;  // (j * r) mod ny
    for (int j = 0; j < g.ny; ++j) {
      const double2 v = src[(size_t)j * g.kzs];
      const double2 w = tw[t];
      if (INV) {
        sr += v.x * w.x - v.y * w.y;
        si += v.x * w.y + v.y * w.x;
      } else {
        sr += v.x * w.x + v.y * w.y;
        si += v.y * w.x - v.x * w.y;
      }
      t += r;
      if (t >= g.ny) t -= g.ny;
    }
    if (!INV) {
      sr /= g.ny;
      si /= g.ny;
    }
  }
  out[(size_t)plane * g.m + idx] = make_double2(sr, si);
}

// ------------------------------------------------------ k_dft_z_c2r ----
// Spanwise c2r of [plane][y][kzs] -> [plane][y][nz]: the Hermitian
// completion of the stored columns (the Nyquist column is zero),
//   f(z) = Re c_0 + 2 Re sum_{kz >= 1} c_kz e^{2 pi i kz z / nz}.
// tw[t] = e^{2 pi i t / nz}.
__global__ void __launch_bounds__(256) k_dft_z_c2r(Geo g, const double2 *__restrict__ tw,
                                                   const double2 *__restrict__ in, double *__restrict__ out) {
  const int plane = blockIdx.y;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  const int npl = g.ny * g.nz;
  if (idx >= npl) return;
  const int y = idx / g.nz, z = idx - y * g.nz;
  const double2 *src = in + (size_t)plane * g.m + (size_t)y * g.kzs;
  double s = src[0].x, acc = 0.0;
  int t = 0;
  for (int k = 1; k < g.kzs; ++k) {
    t += z;
    if (t >= g.nz) t -= g.nz;
    const double2 v = src[k], w = tw[t];
    acc += v.x * w.x - v.y * w.y;
  }
  out[(size_t)plane * npl + idx] = s + 2.0 * acc;
}

// ------------------------------------------------------ k_dft_z_r2c ----
// Spanwise r2c of [plane][y][nz] -> [plane][y][kzs]:
//   c_kz = (1/nz) sum_z f(z) e^{-2 pi i kz z / nz}, kz < kzs.
__global__ void __launch_bounds__(256) k_dft_z_r2c(Geo g, const double2 *__restrict__ tw,
                                                   const double *__restrict__ in, double2 *__restrict__ out) {
  const int plane = blockIdx.y;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= g.m) return;
  const int y = idx / g.kzs, k = idx - y * g.kzs;
  const double *src = in + (size_t)plane * g.ny * g.nz + (size_t)y * g.nz;
  double sr = 0.0, si = 0.0;
  int t = 0;
  for (int z = 0; z < g.nz; ++z) {
    const double v = src[z];
    const double2 w = tw[t];
    sr += v * w.x;
    si -= v * w.y;
    t += k;
    if (t >= g.nz) t -= g.nz;
  }
  out[(size_t)plane * g.m + idx] = make_double2(sr / g.nz, si / g.nz);
}

// ----------------------------------------------------------- k_emf ----
// One block per point row p of the physical fields [10][P][ny][nz]
// (u 0-2, b 3-5, J 6-8, div A 9).  u_tot = u + (0, U x_p, 0), b_tot = b + B0,
// ideal = -(u_tot x b_tot), resistive = eta J, total = ideal + resistive
// (examples/pcf_mhd_mri_shearpy_jax.py:265-276).  Interior rows write the
// ideal and resistive fields [6][n0][ny][nz] for their TC projection and
// their weighted sums; the two wall rows give the tangential maxima, the
// spanwise-streamwise sums of the total field's tangential components and
// max |div A| (:296-353, :355-429).  Results: part[p][EMF_Q].
__global__ void __launch_bounds__(256) k_emf(Geo g, int P, double eta, const double *__restrict__ xp,
                                             const double *__restrict__ wq, const double *__restrict__ phys,
                                             double *__restrict__ E, double *__restrict__ part) {
  __shared__ double red[256];
  const int p = blockIdx.x;
  const int npl = g.ny * g.nz;
  const size_t fs = (size_t)P * npl;
  const bool wall = p >= g.n0;
  const double U = g.uwall * xp[p], w = wall ? 0.0 : wq[p];
  double S[EMF_Q];
  for (int q = 0; q < EMF_Q; ++q) S[q] = 0.0;
  for (int i = threadIdx.x; i < npl; i += 256) {
    const size_t at = (size_t)p * npl + i;
    const double u0 = phys[at], u1 = phys[fs + at] + U, u2 = phys[2 * fs + at];
    const double b0 = phys[3 * fs + at], b1 = phys[4 * fs + at], b2 = phys[5 * fs + at];
    const double t0 = b0 + g.b0[0], t1 = b1 + g.b0[1], t2 = b2 + g.b0[2];
    const double j0 = phys[6 * fs + at], j1 = phys[7 * fs + at], j2 = phys[8 * fs + at];
    const double e0 = -(u1 * t2 - u2 * t1), e1 = -(u2 * t0 - u0 * t2), e2 = -(u0 * t1 - u1 * t0);
    const double r0 = eta * j0, r1 = eta * j1, r2 = eta * j2;
    const double s0 = e0 + r0, s1 = e1 + r1, s2 = e2 + r2;
    if (!wall) {
      const size_t ef = (size_t)g.n0 * npl, ea = (size_t)p * npl + i;
      E[ea] = e0;
      E[ef + ea] = e1;
      E[2 * ef + ea] = e2;
      E[3 * ef + ea] = r0;
      E[4 * ef + ea] = r1;
      E[5 * ef + ea] = r2;
      S[EMF_IDEAL2] += w * (e0 * e0 + e1 * e1 + e2 * e2);
      S[EMF_RES2] += w * (r0 * r0 + r1 * r1 + r2 * r2);
      S[EMF_TOT2] += w * (s0 * s0 + s1 * s1 + s2 * s2);
      S[EMF_MEAN + 0] += w * e0;
      S[EMF_MEAN + 1] += w * e1;
      S[EMF_MEAN + 2] += w * e2;
      S[EMF_MEAN + 3] += w * r0;
      S[EMF_MEAN + 4] += w * r1;
      S[EMF_MEAN + 5] += w * r2;
      S[EMF_MEAN + 6] += w * s0;
      S[EMF_MEAN + 7] += w * s1;
      S[EMF_MEAN + 8] += w * s2;
      S[EMF_BMAX] = fmax(S[EMF_BMAX], fmax(fabs(b0), fmax(fabs(b1), fabs(b2))));
      S[EMF_BTMAX] = fmax(S[EMF_BTMAX], fmax(fabs(t0), fmax(fabs(t1), fabs(t2))));
    } else {
      S[EMF_WIDEAL] = fmax(S[EMF_WIDEAL], fmax(fabs(e1), fabs(e2)));
      S[EMF_WRES] = fmax(S[EMF_WRES], fmax(fabs(r1), fabs(r2)));
      S[EMF_WTOT] = fmax(S[EMF_WTOT], fmax(fabs(s1), fabs(s2)));
      S[EMF_WSUM1] += s1;
      S[EMF_WSUM2] += s2;
      S[EMF_WDIVA] = fmax(S[EMF_WDIVA], fabs(phys[9 * fs + at]));
    }
  }
  for (int q = 0; q < EMF_Q; ++q) {
    const bool is_max = q == EMF_BMAX || q == EMF_BTMAX || q == EMF_WIDEAL || q == EMF_WRES || q == EMF_WTOT ||
                        q == EMF_WDIVA;
    const double r = is_max ? block_max256(S[q], red) : block_sum256(S[q], red);
    if (threadIdx.x == 0) part[(size_t)p * EMF_Q + q] = r;
  }
}

// ------------------------------------------------------- k_parseval ----
// One block per (node x, job): the sum over the stored modes of one
// quadratic or linear form of nodal-spectral arrays at node x.  part[job][x].
__global__ void __launch_bounds__(256) k_parseval(Geo g, const ParsevalJob *__restrict__ jobs,
                                                  double *__restrict__ part) {
  __shared__ double red[256];
  const int x = blockIdx.x, j = blockIdx.y;
  const ParsevalJob J = jobs[j];
  const double2 *a = J.a + (size_t)x * g.m, *b = J.b + (size_t)x * g.m;
  double s = 0.0;
  for (int md = threadIdx.x; md < g.m; md += 256) {
    const int iy = md / g.kzs, kz = md - iy * g.kzs;
    const double wt = kz == 0 ? 1.0 : 2.0;
    double2 va = a[md];
    if (md == 0) va.x += J.add0a;
    switch (J.kind) {
      case PV_DOT: {
        double2 vb = b[md];
        if (md == 0) vb.x += J.add0b;
        s += wt * (va.x * vb.x + va.y * vb.y);
        break;
      }
      case PV_K2: {
        const double ky = ky_of(g, iy), kzv = kz * g.df2;
        s += wt * (ky * ky + kzv * kzv) * (va.x * va.x + va.y * va.y);
        break;
      }
      case PV_MEAN:
        if (md == 0) s += va.x;
        break;
      case PV_KZ1:
        if (md == 1) s += va.x * va.x + va.y * va.y;
        break;
      case PV_SUM2: {
        const double2 vb = b[md];
        const double re = va.x + vb.x, im = va.y + vb.y;
        s += wt * (re * re + im * im);
        break;
      }
      default: break;
    }
  }
  s = block_sum256(s, red);
  if (threadIdx.x == 0) part[(size_t)j * g.n0 + x] = s;
}

// --------------------------------------------------------- k_seed_b ----
// Cotangent of Emag = Ly Lz sum_x fejer_x sum_k |b(x, k)|^2 (full spectrum)
// at the nodal fields X = (b, j) [6][n0][m]: mu_b = scale fejer_x b, mu_j = 0.
__global__ void __launch_bounds__(256) k_seed_b(Geo g, const double *__restrict__ fejer, double scale,
                                                double2 *__restrict__ X) {
  const int x = blockIdx.y;
  const int md = blockIdx.x * 256 + threadIdx.x;
  if (md >= g.m) return;
  const size_t pl = (size_t)g.n0 * g.m, at = (size_t)x * g.m + md;
  const double w = scale * fejer[x];
  for (int f = 0; f < 3; ++f) {
    const double2 v = X[f * pl + at];
    X[f * pl + at] = make_double2(w * v.x, w * v.y);
    X[(3 + f) * pl + at] = make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------- launchers ----
hipError_t launch_seed_b(const Geo &g, const double *fejer, double scale, double2 *X, hipStream_t s) {
  hipLaunchKernelGGL(k_seed_b, dim3((g.m + 255) / 256, g.n0), dim3(256), 0, s, g, fejer, scale, X);
  return hipGetLastError();
}

hipError_t launch_eval_terms(const Geo &g, const EvalOut *o, int n, hipStream_t s) {
  for (int b0 = 0; b0 < n; b0 += EVAL_BATCH) {
    EvalBatch b;
    b.n = std::min(EVAL_BATCH, n - b0);
    int rows = 0;
    for (int j = 0; j < b.n; ++j) {
      const EvalOut &e = o[b0 + j];
      if (!e.dst || e.rows < 1 || e.nt < 1 || e.nt > EVAL_TERMS) return hipErrorInvalidValue;
      for (int t = 0; t < e.nt; ++t)
        if (!e.t[t].M || !e.t[t].F || e.t[t].K < 1) return hipErrorInvalidValue;
      b.o[j] = e;
      rows = std::max(rows, e.rows);
    }
    hipLaunchKernelGGL(k_eval_terms, dim3((g.m + 63) / 64, (rows + EVAL_RW - 1) / EVAL_RW, b.n), dim3(256), 0, s, g,
                       b);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_curl_b(const Geo &g, int P, const double2 *b, const double2 *db, double2 *J, hipStream_t s) {
  hipLaunchKernelGGL(k_curl_b, dim3((g.m + 255) / 256, P), dim3(256), 0, s, g, P, b, db, J);
  return hipGetLastError();
}

hipError_t launch_dft_y(const Geo &g, bool inverse, int planes, const double2 *tw, const double2 *in, double2 *out,
                        hipStream_t s) {
  const dim3 grid((g.m + 255) / 256, planes);
  if (inverse)
    hipLaunchKernelGGL(k_dft_y<true>, grid, dim3(256), 0, s, g, tw, in, out);
  else
    hipLaunchKernelGGL(k_dft_y<false>, grid, dim3(256), 0, s, g, tw, in, out);
  return hipGetLastError();
}

hipError_t launch_dft_z_c2r(const Geo &g, int planes, const double2 *tw, const double2 *in, double *out,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_dft_z_c2r, dim3((g.ny * g.nz + 255) / 256, planes), dim3(256), 0, s, g, tw, in, out);
  return hipGetLastError();
}

hipError_t launch_dft_z_r2c(const Geo &g, int planes, const double2 *tw, const double *in, double2 *out,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_dft_z_r2c, dim3((g.m + 255) / 256, planes), dim3(256), 0, s, g, tw, in, out);
  return hipGetLastError();
}

hipError_t launch_emf(const Geo &g, int P, double eta, const double *xp, const double *wq, const double *phys,
                      double *E, double *part, hipStream_t s) {
  hipLaunchKernelGGL(k_emf, dim3(P), dim3(256), 0, s, g, P, eta, xp, wq, phys, E, part);
  return hipGetLastError();
}

hipError_t launch_parseval(const Geo &g, const ParsevalJob *jobs, int njobs, double *part, hipStream_t s) {
  hipLaunchKernelGGL(k_parseval, dim3(g.n0, njobs), dim3(256), 0, s, g, jobs, part);
  return hipGetLastError();
}

}  // namespace pcf
