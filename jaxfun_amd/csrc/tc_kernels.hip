// Kernels of the axisymmetric Taylor-Couette step
// (reference: examples/taylor_couette_dns_jax.py:703-756, 885-946, 1036-1069).
//
// Layout: fields are [component][k][i] complex, k the stored axial mode
// (0..kh-1, the Hermitian half), i the radial Dirichlet index; the step is
// small (Nr ~ 40, Nz ~ 8-32), so each kernel keeps one radial point (the
// product kernel) or one axial mode (the solve kernel) per block with the
// per-block data in LDS, and dot products run one wave per output row with
// the lanes over the contracted index (coalesced reads of the row).
#include "tc_kernels.hpp"

namespace tc {
namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---------------------------------------------------------- k_tc_nonlinear --
// One block per padded radial point j: u, du/dr and du/dz of the three
// components at (r_j, every k), the axial synthesis on mz points, the
// cylindrical advection products (:716-739)
//   n_r = u_r du_r/dr + u_z du_r/dz - u_t^2 / r
//   n_t = u_r du_t/dr + u_z du_t/dz + u_r u_t / r
//   n_z = u_r du_z/dr + u_z du_z/dz,
// and their axial analysis G[c][k][j] = (Lz / mz) sum_m n_c(z_m) e^{-i k z_m}.
__global__ void __launch_bounds__(256) k_tc_nonlinear(TcGeo g, const double2 *__restrict__ u,
                                                      double2 *__restrict__ G) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int j = blockIdx.x, kh = g.kh, n = g.n, mz = g.mz;
  double2 *val = reinterpret_cast<double2 *>(smem);        // [9][kh]: (c, v/dr/dz)
  double *phys = reinterpret_cast<double *>(val + 9 * kh);  // [9][mz]
  double *prod = phys + 9 * mz;                             // [3][mz]
  const int tid = threadIdx.x;
  for (int t = tid; t < 6 * kh; t += blockDim.x) {
    const int c = t / (2 * kh), rem = t - c * 2 * kh, typ = rem / kh, k = rem - typ * kh;
    const double *E = (typ == 0 ? g.Ev : g.Ed) + (size_t)j * n;
    const double2 *x = u + ((size_t)c * kh + k) * n;
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < n; ++i) {
      sr = fma(E[i], x[i].x, sr);
      si = fma(E[i], x[i].y, si);
    }
    val[(c * 3 + typ) * kh + k] = make_double2(sr, si);
    if (typ == 0) val[(c * 3 + 2) * kh + k] = make_double2(-g.kz[k] * si, g.kz[k] * sr);  // i kz u
  }
  __syncthreads();
  // Hermitian synthesis: f(z_m) = Re c_0 + 2 Re sum_{k>0} c_k e^{i k z_m}
  for (int t = tid; t < 9 * mz; t += blockDim.x) {
    const int f = t / mz, m = t - f * mz;
    const double2 *c = val + f * kh;
    double s = 0.0;
    for (int k = 1; k < kh; ++k) {
      const double2 w = g.twz[(k * m) % mz];
      s += c[k].x * w.x - c[k].y * w.y;
    }
    phys[t] = c[0].x + 2.0 * s;
  }
  __syncthreads();
  const double ir = g.invr[j];
  for (int m = tid; m < mz; m += blockDim.x) {
    const double ur = phys[0 * mz + m], urr = phys[1 * mz + m], urz = phys[2 * mz + m];
    const double ut = phys[3 * mz + m], utr = phys[4 * mz + m], utz = phys[5 * mz + m];
    const double uz = phys[6 * mz + m], uzr = phys[7 * mz + m], uzz = phys[8 * mz + m];
    prod[m] = ur * urr + uz * urz - ut * ut * ir;
    prod[mz + m] = ur * utr + uz * utz + ur * ut * ir;
    prod[2 * mz + m] = ur * uzr + uz * uzz;
  }
  __syncthreads();
  const double sc = g.lz / mz;
  for (int t = tid; t < 3 * kh; t += blockDim.x) {
    const int c = t / kh, k = t - c * kh;
    const double *p = prod + c * mz;
    double sr = 0.0, si = 0.0;
    for (int m = 0; m < mz; ++m) {
      const double2 w = g.twz[(k * m) % mz];
      sr = fma(p[m], w.x, sr);
      si = fma(-p[m], w.y, si);
    }
    G[((size_t)c * kh + k) * g.mr + j] = make_double2(sc * sr, sc * si);
  }
}

// --------------------------------------------------------- k_tc_rhs_solve --
// One block per stored axial mode k: N = Pr G (the dealiased scalar products
// of the products, :733-739), the right-hand side of the integrator as a sum
// of operator terms, then (u, p) = Ainv_k rhs with the pinned inverse of the
// mode's saddle-point block (:614-637; the pressure rows of rhs are zero, so
// only the velocity columns of the inverse are stored).
__global__ void __launch_bounds__(256) k_tc_rhs_solve(TcGeo g, TcRhs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = blockIdx.x, kh = g.kh, n = g.n, n3 = 3 * n, n4 = 4 * n;
  double2 *Ncur = reinterpret_cast<double2 *>(smem);  // [3n]
  double2 *xs = Ncur + n3;                              // [3n]
  double2 *rhs = xs + n3;                               // [3n]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  auto at = [&](const double2 *f, int row) {  // component row / n, radial row % n of mode k
    const int c = row / n;
    return f[((size_t)c * kh + k) * n + (row - c * n)];
  };
  if (a.G) {
    for (int row = wave; row < n3; row += nw) {
      const int c = row / n, i = row - c * n;
      const double *P = g.Pr + (size_t)i * g.mr;
      const double2 *Gc = a.G + ((size_t)c * kh + k) * g.mr;
      double sr = 0.0, si = 0.0;
      for (int jj = lane; jj < g.mr; jj += 64) {
        sr = fma(P[jj], Gc[jj].x, sr);
        si = fma(P[jj], Gc[jj].y, si);
      }
      sr = wave_sum(sr);
      si = wave_sum(si);
      if (lane == 0) {
        Ncur[row] = make_double2(sr, si);
        if (a.nout) a.nout[((size_t)c * kh + k) * n + i] = make_double2(sr, si);
      }
    }
  }
  __syncthreads();
  for (int row = threadIdx.x; row < n3; row += blockDim.x)
    rhs[row] = a.G ? make_double2(a.cN * Ncur[row].x, a.cN * Ncur[row].y) : make_double2(0.0, 0.0);
  for (int t = 0; t < a.nt; ++t) {
    __syncthreads();
    for (int row = threadIdx.x; row < n3; row += blockDim.x) xs[row] = at(a.x[t], row);
    __syncthreads();
    const double ct = a.c[t];
    if (a.op[t] == TC_OP_ID) {
      for (int row = threadIdx.x; row < n3; row += blockDim.x) {
        rhs[row].x += ct * xs[row].x;
        rhs[row].y += ct * xs[row].y;
      }
      continue;
    }
    // dense rows: the mass block of the row's component, or the 3n x 3n
    // operator of mode k
    const bool mass = a.op[t] == TC_OP_MASS;
    const double *Mk = (a.op[t] == TC_OP_LIN ? g.Lin : g.Lexp) + (size_t)k * n3 * n3;
    for (int row = wave; row < n3; row += nw) {
      const int c = row / n, i = row - c * n;
      const double *M = mass ? g.Mass + (size_t)i * n : Mk + (size_t)row * n3;
      const double2 *x = mass ? xs + c * n : xs;
      const int len = mass ? n : n3;
      double sr = 0.0, si = 0.0;
      for (int jj = lane; jj < len; jj += 64) {
        sr = fma(M[jj], x[jj].x, sr);
        si = fma(M[jj], x[jj].y, si);
      }
      sr = wave_sum(sr);
      si = wave_sum(si);
      if (lane == 0) {
        rhs[row].x += ct * sr;
        rhs[row].y += ct * si;
      }
    }
  }
  __syncthreads();
  const double2 *A = a.Ainv + (size_t)k * n4 * n3;
  for (int row = wave; row < n4; row += nw) {
    const double2 *Ar = A + (size_t)row * n3;
    double sr = 0.0, si = 0.0;
    for (int jj = lane; jj < n3; jj += 64) {
      const double2 m = Ar[jj], r = rhs[jj];
      sr += m.x * r.x - m.y * r.y;
      si += m.x * r.y + m.y * r.x;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) {
      if (row < n3) {
        const int c = row / n, i = row - c * n;
        a.uout[((size_t)c * kh + k) * n + i] = make_double2(sr, si);
      } else {
        a.pout[(size_t)k * n + (row - n3)] = make_double2(sr, si);
      }
    }
  }
}

// --------------------------------------------------------------- k_tc_diag --
// One block: E = (1/2) int |u|^2 r dr dz and Eth = int |u_t|^2 r dr dz by
// Parseval in z (Fejer-I weights in r), the largest |div u| on the standard
// grid, the continuity residual of the weak divergence rows (:1036-1069).
__device__ double block_sum(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ double block_max(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = fmax(red[t], red[t + s]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(256) k_tc_diag(TcGeo g, const double2 *__restrict__ u, double *__restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int kh = g.kh, n = g.n, nr = g.nr, nz = g.nz;
  double *red = reinterpret_cast<double *>(smem);               // [blockDim]
  double2 *v = reinterpret_cast<double2 *>(red + blockDim.x);   // [4][kh][nr]: ur, ut, uz, dur/dr
  const int tid = threadIdx.x;
  for (int t = tid; t < 4 * kh * nr; t += blockDim.x) {
    const int f = t / (kh * nr), rem = t - f * kh * nr, k = rem / nr, q = rem - k * nr;
    const double *E = (f == 3 ? g.Esd : g.Es) + (size_t)q * n;
    const double2 *x = u + ((size_t)(f == 3 ? 0 : f) * kh + k) * n;
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < n; ++i) {
      sr = fma(E[i], x[i].x, sr);
      si = fma(E[i], x[i].y, si);
    }
    v[t] = make_double2(sr, si);
  }
  __syncthreads();
  double e = 0.0, eth = 0.0;
  for (int t = tid; t < kh * nr; t += blockDim.x) {
    const int k = t / nr, q = t - k * nr;
    const double w = (k == 0 ? 1.0 : 2.0) * g.rstd[q] * g.fejer[q] * g.lz;
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
      const double2 a = v[(c * kh + k) * nr + q];
      const double m2 = a.x * a.x + a.y * a.y;
      s += m2;
      if (c == 1) eth += w * m2;
    }
    e += 0.5 * w * s;
  }
  e = block_sum(e, red);
  eth = block_sum(eth, red);
  double dmax = 0.0;
  for (int t = tid; t < nz * nr; t += blockDim.x) {
    const int m = t / nr, q = t - m * nr;
    double s0 = 0.0, s = 0.0;
    for (int k = 0; k < kh; ++k) {
      const double2 ur = v[(0 * kh + k) * nr + q], uz = v[(2 * kh + k) * nr + q], dr = v[(3 * kh + k) * nr + q];
      const double2 d = make_double2(dr.x + ur.x / g.rstd[q] - g.kz[k] * uz.y, dr.y + ur.y / g.rstd[q] + g.kz[k] * uz.x);
      if (k == 0) {
        s0 = d.x;
      } else {
        const double2 w = g.twn[(k * m) % nz];
        s += d.x * w.x - d.y * w.y;
      }
    }
    dmax = fmax(dmax, fabs(s0 + 2.0 * s));
  }
  dmax = block_max(dmax, red);
  double cres = 0.0;
  for (int t = tid; t < kh * n; t += blockDim.x) {
    const int k = t / n, i = t - k * n;
    const double2 *ur = u + ((size_t)0 * kh + k) * n, *uz = u + ((size_t)2 * kh + k) * n;
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
    for (int jj = 0; jj < n; ++jj) {
      ar = fma(g.Drc[(size_t)i * n + jj], ur[jj].x, ar);
      ai = fma(g.Drc[(size_t)i * n + jj], ur[jj].y, ai);
      br = fma(g.Dzc[(size_t)i * n + jj], uz[jj].x, br);
      bi = fma(g.Dzc[(size_t)i * n + jj], uz[jj].y, bi);
    }
    const double rr = ar - g.kz[k] * bi, ri = ai + g.kz[k] * br;
    cres += (k == 0 ? 1.0 : 2.0) * (rr * rr + ri * ri);
  }
  cres = block_sum(cres, red);
  if (tid == 0) {
    out[0] = e;
    out[1] = dmax;
    out[2] = sqrt(cres);
    out[3] = 0.0;  // the Dirichlet modes vanish at both walls: T_i(+-1) = T_{i+2}(+-1)
    out[4] = eth;
  }
}

}  // namespace

hipError_t launch_nonlinear(const TcGeo &g, const double2 *u, double2 *G, hipStream_t s) {
  const size_t lds = 9 * g.kh * sizeof(double2) + 12 * (size_t)g.mz * sizeof(double);
  hipLaunchKernelGGL(k_tc_nonlinear, dim3(g.mr), dim3(256), lds, s, g, u, G);
  return hipGetLastError();
}

hipError_t launch_rhs_solve(const TcGeo &g, const TcRhs &a, hipStream_t s) {
  if (a.nt < 0 || a.nt > TC_TERMS || !a.Ainv || !a.uout || !a.pout) return hipErrorInvalidValue;
  const size_t lds = 9 * (size_t)g.n * sizeof(double2);
  hipLaunchKernelGGL(k_tc_rhs_solve, dim3(g.kh), dim3(256), lds, s, g, a);
  return hipGetLastError();
}

hipError_t launch_diag(const TcGeo &g, const double2 *u, double *out, hipStream_t s) {
  const size_t lds = 256 * sizeof(double) + 4 * (size_t)g.kh * g.nr * sizeof(double2);
  hipLaunchKernelGGL(k_tc_diag, dim3(1), dim3(256), lds, s, g, u, out);
  return hipGetLastError();
}

}  // namespace tc
