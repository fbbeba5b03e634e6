// Kernels of the axisymmetric Taylor-Couette step
// (reference: examples/taylor_couette_dns_jax.py:703-756, 885-946, 1036-1069).
//
// Layout: fields are [component][k][i] complex, k the stored axial mode
// (0..kh-1, the Hermitian half), i the radial Dirichlet index; the step is
// small (Nr ~ 40, Nz ~ 8-32), so each kernel keeps one radial point (the
// product kernel) or one axial mode (the solve kernel) per block with the
// per-block data in LDS.
#include "tc_kernels.hpp"

#include <algorithm>

namespace tc {
namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// ---------------------------------------------------------- k_tc_nonlinear --
// One block per padded radial point j: value, d/dr and d/dz of every field at
// (r_j, every k), the axial synthesis on mz points, the products, and their
// axial analysis G[c][k][j] = (Lz / mz) sum_m n_c(z_m) e^{-i k z_m}.
// Hydrodynamic (NF = 3, :716-739): the cylindrical advection
//   n_r = u_r du_r/dr + u_z du_r/dz - u_t^2 / r
//   n_t = u_r du_t/dr + u_z du_t/dz + u_r u_t / r
//   n_z = u_r du_z/dr + u_z du_z/dz.
// MHD (NF = 6, :1603-1649): the advection minus the same form of b (the
// Lorentz force) in G[0..2] and the EMF u x b in G[3..5]; the rows of b are
// formed from the latter in the solve kernel.
// The twiddles sit in LDS and every transform loop steps its table index
// instead of reducing k m modulo mz.
template <int NF>
__global__ void __launch_bounds__(256) k_tc_nonlinear(TcGeo g, const double2 *__restrict__ u,
                                                      double2 *__restrict__ G) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NV = 3 * NF, NG = NF == 6 ? 6 : 3;
  const int j = blockIdx.x, kh = g.kh, n = g.n, mz = g.mz, mr = g.mr;
  double2 *tw = reinterpret_cast<double2 *>(smem);             // [mz]
  double2 *val = tw + mz;                                      // [NV][kh]: (c, v/dr/dz)
  double *phys = reinterpret_cast<double *>(val + NV * kh);    // [NV][mz]
  double *prod = phys + NV * mz;                               // [NG][mz]
  double2 *part = reinterpret_cast<double2 *>(prod + NG * mz);  // [nch][2 NF kh] partial sums
  const int tid = threadIdx.x;
  for (int t = tid; t < mz; t += blockDim.x) tw[t] = g.twz[t];
  // 2 NF kh dot products of length n (value and d/dr of every field at every
  // mode), each split into nch chunks so that every thread has work
  const int nd = 2 * NF * kh, nch = max(1, (int)blockDim.x / nd), len = (n + nch - 1) / nch;
  for (int t = tid; t < nd * nch; t += blockDim.x) {
    const int d = t % nd, ch = t / nd;
    const int c = d / (2 * kh), rem = d - c * 2 * kh, typ = rem / kh, k = rem - typ * kh;
    const int basis = c < 4 ? 0 : c - 3;  // b_theta: Robin, b_z: Neumann
    const double *E = (typ == 0 ? g.Ev : g.Ed) + (size_t)basis * mr * n + (size_t)j * n;
    const double2 *x = u + ((size_t)c * kh + k) * n;
    const int i0 = ch * len, i1 = min(n, i0 + len);
    double sr = 0.0, si = 0.0;
    for (int i = i0; i < i1; ++i) {
      sr = fma(E[i], x[i].x, sr);
      si = fma(E[i], x[i].y, si);
    }
    part[t] = make_double2(sr, si);
  }
  __syncthreads();
  for (int d = tid; d < nd; d += blockDim.x) {
    double sr = 0.0, si = 0.0;
    for (int ch = 0; ch < nch; ++ch) {
      sr += part[ch * nd + d].x;
      si += part[ch * nd + d].y;
    }
    const int c = d / (2 * kh), rem = d - c * 2 * kh, typ = rem / kh, k = rem - typ * kh;
    val[(c * 3 + typ) * kh + k] = make_double2(sr, si);
    if (typ == 0) val[(c * 3 + 2) * kh + k] = make_double2(-g.kz[k] * si, g.kz[k] * sr);  // i kz u
  }
  __syncthreads();
  // Hermitian synthesis: f(z_m) = Re c_0 + 2 Re sum_{k>0} c_k e^{i k z_m}
  for (int t = tid; t < NV * mz; t += blockDim.x) {
    const int f = t / mz, m = t - f * mz;
    const double2 *c = val + f * kh;
    double s = 0.0;
    int idx = 0;
    for (int k = 1; k < kh; ++k) {
      idx += m;
      if (idx >= mz) idx -= mz;
      const double2 w = tw[idx];
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
    double nr = ur * urr + uz * urz - ut * ut * ir;
    double nt = ur * utr + uz * utz + ur * ut * ir;
    double nz = ur * uzr + uz * uzz;
    if constexpr (NF == 6) {
      const double br = phys[9 * mz + m], brr = phys[10 * mz + m], brz = phys[11 * mz + m];
      const double bt = phys[12 * mz + m], btr = phys[13 * mz + m], btz = phys[14 * mz + m];
      const double bz = phys[15 * mz + m], bzr = phys[16 * mz + m], bzz = phys[17 * mz + m];
      nr -= br * brr + bz * brz - bt * bt * ir;
      nt -= br * btr + bz * btz + br * bt * ir;
      nz -= br * bzr + bz * bzz;
      prod[3 * mz + m] = ut * bz - uz * bt;
      prod[4 * mz + m] = uz * br - ur * bz;
      prod[5 * mz + m] = ur * bt - ut * br;
    }
    prod[m] = nr;
    prod[mz + m] = nt;
    prod[2 * mz + m] = nz;
  }
  __syncthreads();
  const double sc = g.lz / mz;
  for (int t = tid; t < NG * kh; t += blockDim.x) {
    const int c = t / kh, k = t - c * kh;
    const double *p = prod + c * mz;
    double sr = 0.0, si = 0.0;
    int idx = 0;
    for (int m = 0; m < mz; ++m) {
      const double2 w = tw[idx];
      sr = fma(p[m], w.x, sr);
      si = fma(-p[m], w.y, si);
      idx += k;
      if (idx >= mz) idx -= mz;
    }
    G[((size_t)c * kh + k) * mr + j] = make_double2(sc * sr, sc * si);
  }
}

constexpr int RHS_THREADS = 256;

// The nonlinear row i of field c from the axial analyses G of the mode: the
// dealiased projection of the products (velocity rows), or with the MHD
// fields minus the curl of the EMF (G components 3..5 hold its r, theta, z
// parts): b_r i kz Qbr G_t, b_theta -i kz Qbt G_r + Qbtd G_z, b_z Qbz G_t.
__device__ __forceinline__ double2 nl_row(const TcGeo &g, const double2 *gs, int c, int i, double kz, double m) {
  const int n = g.n, mr = g.mr;
  auto dot = [&](const double *T, const double2 *Gc) {
    double sr = 0.0, si = 0.0;
#pragma unroll 8
    for (int jj = 0; jj < mr; ++jj) {
      const double p = T[(size_t)jj * n + i];
      sr = fma(p, Gc[jj].x, sr);
      si = fma(p, Gc[jj].y, si);
    }
    return make_double2(sr, si);
  };
  if (c < 3) return dot(g.PrT, gs + c * mr);
  if (c == 3) {  // i kz Qbr G_t - i m Qbrr G_z
    const double2 v = dot(g.QbrT, gs + 4 * mr);
    double2 o = make_double2(-kz * v.y, kz * v.x);
    if (m != 0.0) {
      const double2 w = dot(g.QbrrT, gs + 5 * mr);
      o.x += m * w.y;
      o.y -= m * w.x;
    }
    return o;
  }
  if (c == 4) {
    const double2 v = dot(g.QbtT, gs + 3 * mr), w = dot(g.QbtdT, gs + 5 * mr);
    return make_double2(w.x + kz * v.y, w.y - kz * v.x);
  }
  double2 o = dot(g.QbzT, gs + 4 * mr);  // Qbz G_t + i m Qbzr G_r
  if (m != 0.0) {
    const double2 w = dot(g.QbzrT, gs + 3 * mr);
    o.x -= m * w.y;
    o.y += m * w.x;
  }
  return o;
}

// --------------------------------------------------------- k_tc_rhs_solve --
// One block per stored Fourier mode: the nonlinear rows N from the product
// kernel's analyses (:733-739; the MHD rows :1603-1649), the right-hand side
// of the integrator, then (fields, p) = Ainv rhs with the pinned inverse of
// the mode's saddle-point block (:614-637; the pressure rows of rhs are zero,
// so only the field columns of the inverse are stored).  The right-hand
// side's terms are first combined per operator (sum c_t X_t for the mass
// terms, the L terms, the explicit terms), so each operator is applied once.
// One thread per right-hand-side row; the tables are stored transposed, so
// consecutive rows read consecutive addresses, and the vector operands sit
// in LDS.  gridDim.y blocks share a mode: each forms the whole right-hand
// side and solves a slice of the rows.
__global__ void __launch_bounds__(RHS_THREADS) k_tc_rhs_solve(TcGeo g, TcRhs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = blockIdx.x, kh = g.nm, n = g.n, nfn = g.nf * n, n4 = nfn + n, mr = g.mr;
  const int ng = g.nf == 6 ? 6 : 3;                 // G components
  double2 *gs = reinterpret_cast<double2 *>(smem);  // [ng mr]: G of mode k
  double2 *xm = gs + ng * mr;                       // [max(nf n, threads)]: combined mass operand
  double2 *xl = xm + max(nfn, RHS_THREADS);         // [nf n]: combined L / explicit operand
  double2 *rhs = xl + nfn;                          // [nf n]
  const int tid = threadIdx.x;
  bool has_mass = false, has_lin = false;
  int lop = TC_OP_LIN;
  for (int t = 0; t < a.nt; ++t) {
    has_mass |= a.op[t] == TC_OP_MASS;
    if (a.op[t] == TC_OP_LIN || a.op[t] == TC_OP_LEXP) {
      has_lin = true;
      lop = a.op[t];
    }
  }
  if (a.G)
    for (int t = tid; t < ng * mr; t += blockDim.x) {
      const int c = t / mr, jj = t - c * mr;
      gs[t] = a.G[((size_t)c * kh + k) * mr + jj];
    }
  for (int row = tid; row < nfn; row += blockDim.x) {
    const int c = row / n;
    const size_t o = ((size_t)c * kh + k) * n + (row - c * n);
    double2 m = make_double2(0.0, 0.0), l = m, d = m;
    for (int t = 0; t < a.nt; ++t) {
      const double2 x = a.x[t][o];
      const double ct = a.c[t];
      double2 &dst = (a.op[t] == TC_OP_MASS) ? m : (a.op[t] == TC_OP_ID) ? d : l;
      dst.x = fma(ct, x.x, dst.x);
      dst.y = fma(ct, x.y, dst.y);
    }
    xm[row] = m;
    xl[row] = l;
    rhs[row] = d;
  }
  __syncthreads();
  const double kz = g.kz[k], mo = g.mq_odd[k];
  for (int row = tid; row < nfn; row += blockDim.x) {
    const int c = row / n, i = row - c * n;
    double2 r = rhs[row];
    if (a.G) {
      const double2 v = nl_row(g, gs, c, i, kz, mo);
      if (a.nout && blockIdx.y == 0) a.nout[((size_t)c * kh + k) * n + i] = v;
      r.x = fma(a.cN, v.x, r.x);
      r.y = fma(a.cN, v.y, r.y);
    }
    if (has_mass) {  // block diagonal: the mass of the row's field
      const double2 *x = xm + c * n;
      const double *MT = g.MassT + (size_t)c * n * n;
      double sr = 0.0, si = 0.0;
#pragma unroll 8
      for (int jj = 0; jj < n; ++jj) {
        const double mm = MT[(size_t)jj * n + i];
        sr = fma(mm, x[jj].x, sr);
        si = fma(mm, x[jj].y, si);
      }
      r.x += sr;
      r.y += si;
    }
    if (has_lin) {  // complex per-mode operator
      const double2 *M = (lop == TC_OP_LIN ? g.LinT : g.LexpT) + (size_t)k * nfn * nfn;
      double sr = 0.0, si = 0.0;
#pragma unroll 8
      for (int jj = 0; jj < nfn; ++jj) {
        const double2 mm = M[(size_t)jj * nfn + row], x = xl[jj];
        sr += mm.x * x.x - mm.y * x.y;
        si += mm.x * x.y + mm.y * x.x;
      }
      r.x += sr;
      r.y += si;
    }
    rhs[row] = r;
  }
  __syncthreads();
  // this block's slice of the solution rows, the contracted index split into
  // chunks over the threads (partial sums meet in LDS in a fixed order)
  const double2 *A = a.AinvT + (size_t)k * nfn * n4;  // [j < nf n][r < (nf + 1) n]
  const int per = (n4 + gridDim.y - 1) / gridDim.y, r0 = blockIdx.y * per, nrow = min(per, n4 - r0);
  double2 *part = xm;  // the mass operand is no longer needed: [RHS_THREADS]
  const int nch = max(1, (int)blockDim.x / max(nrow, 1)), clen = (nfn + nch - 1) / nch;
  for (int t = tid; t < nrow * nch; t += blockDim.x) {
    const int row = r0 + t % nrow, ch = t / nrow, j0 = ch * clen, j1 = min(nfn, j0 + clen);
    double sr = 0.0, si = 0.0;
#pragma unroll 4
    for (int jj = j0; jj < j1; ++jj) {
      const double2 m = A[(size_t)jj * n4 + row], r = rhs[jj];
      sr += m.x * r.x - m.y * r.y;
      si += m.x * r.y + m.y * r.x;
    }
    part[t] = make_double2(sr, si);
  }
  __syncthreads();
  for (int t = tid; t < nrow; t += blockDim.x) {
    double sr = 0.0, si = 0.0;
    for (int ch = 0; ch < nch; ++ch) {
      sr += part[ch * nrow + t].x;
      si += part[ch * nrow + t].y;
    }
    const int row = r0 + t;
    if (row < nfn) {
      const int c = row / n, i = row - c * n;
      a.uout[((size_t)c * kh + k) * n + i] = make_double2(sr, si);
    } else {
      a.pout[(size_t)k * n + (row - nfn)] = make_double2(sr, si);
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
  const int kh = g.kh, nm = g.nm, n = g.n, nr = g.nr, nz = g.nz, nth = g.nth;
  double *red = reinterpret_cast<double *>(smem);  // [blockDim]
  double2 *dv = g.dv;                              // [nm][nr]: the divergence per mode (global scratch)
  const int tid = threadIdx.x;
  double e = 0.0, eth = 0.0;
  for (int t = tid; t < nm * nr; t += blockDim.x) {
    const int q = t / nr, r = t - q * nr, k = q % kh;
    const double *E = g.Es + (size_t)r * n, *Ed = g.Esd + (size_t)r * n;
    double2 v[4];
    for (int f = 0; f < 4; ++f) {
      const double *M = f == 3 ? Ed : E;
      const double2 *x = u + ((size_t)(f == 3 ? 0 : f) * nm + q) * n;
      double sr = 0.0, si = 0.0;
      for (int i = 0; i < n; ++i) {
        sr = fma(M[i], x[i].x, sr);
        si = fma(M[i], x[i].y, si);
      }
      v[f] = make_double2(sr, si);
    }
    // Parseval over (theta, z): every stored k > 0 mode stands for its conjugate too
    const double w = (k == 0 ? 1.0 : 2.0) * g.rstd[r] * g.fejer[r] * g.fz;
    double s = 0.0;
    for (int c = 0; c < 3; ++c) s += v[c].x * v[c].x + v[c].y * v[c].y;
    e += 0.5 * w * s;
    eth += w * (v[1].x * v[1].x + v[1].y * v[1].y);
    const double ir = 1.0 / g.rstd[r], mo = g.mq_odd[q], kz = g.kz[q];
    // d u_r / dr + u_r / r + (i m / r) u_theta + i kz u_z
    dv[t] = make_double2(v[3].x + v[0].x * ir - mo * ir * v[1].y - kz * v[2].y,
                         v[3].y + v[0].y * ir + mo * ir * v[1].x + kz * v[2].x);
  }
  e = block_sum(e, red);
  eth = block_sum(eth, red);
  __threadfence_block();
  __syncthreads();
  double dmax = 0.0;
  for (int t = tid; t < nth * nz * nr; t += blockDim.x) {
    const int a = t / (nz * nr), rem = t - a * nz * nr, b = rem / nr, r = rem - b * nr;
    double s = 0.0;
    for (int q = 0; q < nm; ++q) {
      const int im = q / kh, k = q - im * kh;
      const double2 d = dv[q * nr + r];
      const int mi = (int)g.mq[q];
      const double2 wt = g.twn_t[((mi * a) % nth + nth) % nth], wz = g.twn[(k * b) % nz];
      const double2 w = make_double2(wt.x * wz.x - wt.y * wz.y, wt.x * wz.y + wt.y * wz.x);
      s += (k == 0 ? 1.0 : 2.0) * (d.x * w.x - d.y * w.y);
    }
    dmax = fmax(dmax, fabs(s));
  }
  dmax = block_max(dmax, red);
  double cres = 0.0;
  for (int t = tid; t < nm * n; t += blockDim.x) {
    const int q = t / n, i = t - q * n, k = q % kh;
    const double2 *ur = u + ((size_t)0 * nm + q) * n, *ut = u + ((size_t)1 * nm + q) * n,
                  *uz = u + ((size_t)2 * nm + q) * n;
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0, cr = 0.0, ci = 0.0;
    for (int jj = 0; jj < n; ++jj) {
      const double dr = g.Drc[(size_t)i * n + jj], dz = g.Dzc[(size_t)i * n + jj], dt = g.Dtc[(size_t)i * n + jj];
      ar = fma(dr, ur[jj].x, ar);
      ai = fma(dr, ur[jj].y, ai);
      br = fma(dz, uz[jj].x, br);
      bi = fma(dz, uz[jj].y, bi);
      cr = fma(dt, ut[jj].x, cr);
      ci = fma(dt, ut[jj].y, ci);
    }
    const double kz = g.kz[q], m = g.mq[q];
    const double rr = ar - kz * bi - m * ci, ri = ai + kz * br + m * cr;
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

// ----------------------------------------------------------- k_tc_mhd_diag --
// One block, the MHD set (:1844-1886; 3D :2232-2245): Ekin and Emag by
// Parseval over the Fourier axes (Fejer-I weights in r), the largest |div u|
// and |div b| on the standard grid (d/dr + 1/r of the radial component,
// (1/r) d/dtheta of the azimuthal one and d/dz of the axial one, b_theta and
// b_z in their wall bases), and the wall trace of u (zero: Dirichlet modes).
__global__ void __launch_bounds__(256) k_tc_mhd_diag(TcGeo g, const double2 *__restrict__ u,
                                                     double *__restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int kh = g.kh, nm = g.nm, n = g.n, nr = g.nr, nz = g.nz, nth = g.nth;
  const size_t se = (size_t)nr * n;
  double *red = reinterpret_cast<double *>(smem);  // [blockDim]
  double2 *dv = g.dv;                              // [2][nm][nr]: div u, div b per mode
  const int tid = threadIdx.x;
  double ek = 0.0, em = 0.0;
  for (int t = tid; t < nm * nr; t += blockDim.x) {
    const int q = t / nr, r = t - q * nr, k = q % kh;
    double2 v[6], dr[2];
    for (int c = 0; c < 6; ++c) {
      const int basis = c < 4 ? 0 : c - 3;
      const double *E = g.Es + basis * se + (size_t)r * n;
      const double2 *x = u + ((size_t)c * nm + q) * n;
      double sr = 0.0, si = 0.0;
      for (int i = 0; i < n; ++i) {
        sr = fma(E[i], x[i].x, sr);
        si = fma(E[i], x[i].y, si);
      }
      v[c] = make_double2(sr, si);
    }
    for (int hh = 0; hh < 2; ++hh) {  // d/dr of u_r and b_r (Dirichlet)
      const double *E = g.Esd + (size_t)r * n;
      const double2 *x = u + ((size_t)(3 * hh) * nm + q) * n;
      double sr = 0.0, si = 0.0;
      for (int i = 0; i < n; ++i) {
        sr = fma(E[i], x[i].x, sr);
        si = fma(E[i], x[i].y, si);
      }
      dr[hh] = make_double2(sr, si);
    }
    const double w = (k == 0 ? 1.0 : 2.0) * g.rstd[r] * g.fejer[r] * g.fz;
    double su = 0.0, sb = 0.0;
    for (int c = 0; c < 3; ++c) {
      su += v[c].x * v[c].x + v[c].y * v[c].y;
      sb += v[c + 3].x * v[c + 3].x + v[c + 3].y * v[c + 3].y;
    }
    ek += 0.5 * w * su;
    em += 0.5 * w * sb;
    const double ir = 1.0 / g.rstd[r], kz = g.kz[q], mo = g.mq_odd[q];
    for (int hh = 0; hh < 2; ++hh) {
      const double2 a = v[3 * hh], th = v[3 * hh + 1], z = v[3 * hh + 2];
      dv[((size_t)hh * nm + q) * nr + r] = make_double2(dr[hh].x + a.x * ir - mo * ir * th.y - kz * z.y,
                                                        dr[hh].y + a.y * ir + mo * ir * th.x + kz * z.x);
    }
  }
  ek = block_sum(ek, red);
  em = block_sum(em, red);
  __threadfence_block();
  __syncthreads();
  double dmax[2] = {0.0, 0.0};
  for (int t = tid; t < 2 * nth * nz * nr; t += blockDim.x) {
    const int hh = t / (nth * nz * nr), rem0 = t - hh * nth * nz * nr;
    const int a = rem0 / (nz * nr), rem = rem0 - a * nz * nr, b = rem / nr, r = rem - b * nr;
    double s = 0.0;
    for (int q = 0; q < nm; ++q) {
      const int k = q % kh, mi = (int)g.mq[q];
      const double2 d = dv[((size_t)hh * nm + q) * nr + r];
      const double2 wt = g.twn_t[((mi * a) % nth + nth) % nth], wz = g.twn[(k * b) % nz];
      const double2 wv = make_double2(wt.x * wz.x - wt.y * wz.y, wt.x * wz.y + wt.y * wz.x);
      s += (k == 0 ? 1.0 : 2.0) * (d.x * wv.x - d.y * wv.y);
    }
    dmax[hh] = fmax(dmax[hh], fabs(s));
  }
  const double du = block_max(dmax[0], red), db = block_max(dmax[1], red);
  if (tid == 0) {
    out[0] = ek;
    out[1] = em;
    out[2] = ek + em;
    out[3] = du;
    out[4] = db;
    out[5] = 0.0;  // the Dirichlet modes vanish at both walls
  }
}

// ------------------------------------------------------- k_tc3_nonlinear --
// The 3D product kernel (:1234-1262): one block per padded radial point j.
// u, du/dr, (1/r-free) du/dtheta and du/dz of the three components at every
// stored (m, kz) mode, the azimuthal synthesis g_f,k(theta) = sum_m c e^{i m
// theta} on mth points, the axial synthesis to the (mth x mz) grid (the
// stored kz > 0 modes count twice for their conjugates), the products
//   n_r = u_r du_r/dr + (u_t/r) du_r/dtheta + u_z du_r/dz - u_t^2 / r
//   n_t = u_r du_t/dr + (u_t/r) du_t/dtheta + u_z du_t/dz + u_r u_t / r
//   n_z = u_r du_z/dr + (u_t/r) du_z/dtheta + u_z du_z/dz,
// then the axial and azimuthal analyses truncated to the stored modes, times
// (2 pi / mth)(Lz / mz); the azimuthal Nyquist rows are masked.  With the MHD
// fields (NF = 6, :2141-2190) the products are the advection minus the
// Lorentz force and the EMF u x b, as in the axisymmetric kernel.  Both
// twiddle tables sit in LDS; the axial synthesis runs k-outer so that one
// twiddle serves all fields of a point.
constexpr int TC3_THREADS = 512;

__device__ __forceinline__ int fft_wavenumber(int im, int nth) { return im < (nth + 1) / 2 ? im : im - nth; }

template <int NF>
__global__ void __launch_bounds__(TC3_THREADS) k_tc3_nonlinear(TcGeo g, const double2 *__restrict__ u,
                                                               double2 *__restrict__ G) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NV = 4 * NF, NG = NF == 6 ? 6 : 3;
  const int j = blockIdx.x, kh = g.kh, nm = g.nm, n = g.n, mz = g.mz, mth = g.mth, nth = g.nth, mr = g.mr;
  double2 *tz = reinterpret_cast<double2 *>(smem);  // [mz]
  double2 *tt = tz + mz;                             // [mth]
  double *base = reinterpret_cast<double *>(tt + mth);
  double2 *val = reinterpret_cast<double2 *>(base);  // [NV][nm]; later prod [NG][mth mz] and h [NG][mth][kh]
  const size_t region = max((size_t)NV * nm * 2, (size_t)NG * mth * mz + (size_t)NG * mth * kh * 2);
  double2 *gs = reinterpret_cast<double2 *>(base + region);  // [NV][kh][mth]
  double *prod = base;
  double2 *h = reinterpret_cast<double2 *>(prod + NG * mth * mz);
  constexpr int NB = NF == 6 ? 3 : 1;
  double *er = reinterpret_cast<double *>(gs + (size_t)NV * kh * mth);  // [NB][2][n]: this point's basis rows
  const int tid = threadIdx.x;
  for (int t = tid; t < mz; t += blockDim.x) tz[t] = g.twz[t];
  for (int t = tid; t < mth; t += blockDim.x) tt[t] = g.twt[t];
  for (int t = tid; t < NB * 2 * n; t += blockDim.x) {
    const int bt = t / n, i = t - bt * n, basis = bt >> 1;
    er[t] = ((bt & 1) ? g.Ed : g.Ev)[(size_t)basis * mr * n + (size_t)j * n + i];
  }
  __syncthreads();
  for (int d = tid; d < 2 * NF * nm; d += blockDim.x) {
    const int c = d / (2 * nm), rem = d - c * 2 * nm, typ = rem / nm, q = rem - typ * nm;
    const int basis = c < 4 ? 0 : c - 3;  // b_theta: Robin, b_z: Neumann
    const double *E = er + (size_t)(basis * 2 + typ) * n;
    const double2 *x = u + ((size_t)c * nm + q) * n;
    // two partial sums and unrolled loads: the row's global loads in flight together
    double sr = 0.0, si = 0.0, sr2 = 0.0, si2 = 0.0;
    int i = 0;
#pragma unroll 4
    for (; i + 1 < n; i += 2) {
      const double2 x0 = x[i], x1 = x[i + 1];
      sr = fma(E[i], x0.x, sr);
      si = fma(E[i], x0.y, si);
      sr2 = fma(E[i + 1], x1.x, sr2);
      si2 = fma(E[i + 1], x1.y, si2);
    }
    if (i < n) {
      sr = fma(E[i], x[i].x, sr);
      si = fma(E[i], x[i].y, si);
    }
    sr += sr2;
    si += si2;
    val[(c * 4 + typ) * nm + q] = make_double2(sr, si);
    if (typ == 0) {
      const double m = g.mq_odd[q], kz = g.kz[q];
      val[(c * 4 + 2) * nm + q] = make_double2(-m * si, m * sr);    // i m u
      val[(c * 4 + 3) * nm + q] = make_double2(-kz * si, kz * sr);  // i kz u
    }
  }
  __syncthreads();
  for (int t = tid; t < NV * kh * mth; t += blockDim.x) {
    const int f = t / (kh * mth), rem = t - f * kh * mth, k = rem / mth, a = rem - k * mth;
    double sr = 0.0, si = 0.0;
    // table index (m a) mod mth stepped by a; it restarts once at the first
    // negative wavenumber
    const int hpos = (nth + 1) / 2;
    int idx = 0;
    for (int im = 0; im < nth; ++im) {
      if (im == hpos) idx = (((hpos - nth) * a) % mth + mth) % mth;
      const double2 c = val[f * nm + im * kh + k];
      const double2 w = tt[idx];
      sr += c.x * w.x - c.y * w.y;
      si += c.x * w.y + c.y * w.x;
      idx += a;
      if (idx >= mth) idx -= mth;
    }
    gs[t] = make_double2(sr, si);
  }
  __syncthreads();
  const double ir = g.invr[j];
  for (int p = tid; p < mth * mz; p += blockDim.x) {
    const int a = p / mz, b = p - a * mz;
    double v[NV];
#pragma unroll
    for (int f = 0; f < NV; ++f) v[f] = 0.0;
    int idx = 0;
    for (int k = 1; k < kh; ++k) {
      idx += b;
      if (idx >= mz) idx -= mz;
      const double2 w = tz[idx];
#pragma unroll
      for (int f = 0; f < NV; ++f) {
        const double2 c = gs[((size_t)f * kh + k) * mth + a];
        v[f] += c.x * w.x - c.y * w.y;
      }
    }
#pragma unroll
    for (int f = 0; f < NV; ++f) v[f] = gs[(size_t)f * kh * mth + a].x + 2.0 * v[f];
    const double ur = v[0], urr = v[1], urt = v[2], urz = v[3];
    const double ut = v[4], utr = v[5], utt = v[6], utz = v[7];
    const double uz = v[8], uzr = v[9], uzt = v[10], uzz = v[11];
    const double a_t = ut * ir;
    double nr = ur * urr + a_t * urt + uz * urz - ut * ut * ir;
    double nt = ur * utr + a_t * utt + uz * utz + ur * ut * ir;
    double nz = ur * uzr + a_t * uzt + uz * uzz;
    if constexpr (NF == 6) {
      const double br = v[12], brr = v[13], brt = v[14], brz = v[15];
      const double bt = v[16], btr = v[17], btt = v[18], btz = v[19];
      const double bz = v[20], bzr = v[21], bzt = v[22], bzz = v[23];
      const double b_t = bt * ir;
      nr -= br * brr + b_t * brt + bz * brz - bt * bt * ir;
      nt -= br * btr + b_t * btt + bz * btz + br * bt * ir;
      nz -= br * bzr + b_t * bzt + bz * bzz;
      prod[3 * mth * mz + p] = ut * bz - uz * bt;
      prod[4 * mth * mz + p] = uz * br - ur * bz;
      prod[5 * mth * mz + p] = ur * bt - ut * br;
    }
    prod[p] = nr;
    prod[mth * mz + p] = nt;
    prod[2 * mth * mz + p] = nz;
  }
  __syncthreads();
  for (int t = tid; t < NG * mth * kh; t += blockDim.x) {
    const int c = t / (mth * kh), rem = t - c * mth * kh, a = rem / kh, k = rem - a * kh;
    const double *pr = prod + (size_t)c * mth * mz + (size_t)a * mz;
    double sr = 0.0, si = 0.0;
    int idx = 0;
    for (int b = 0; b < mz; ++b) {
      const double2 w = tz[idx];
      sr = fma(pr[b], w.x, sr);
      si = fma(-pr[b], w.y, si);
      idx += k;
      if (idx >= mz) idx -= mz;
    }
    h[t] = make_double2(sr, si);  // [c][a][k]
  }
  __syncthreads();
  const double sc = (2.0 * 3.141592653589793 / mth) * (g.lz / mz);
  for (int t = tid; t < NG * nm; t += blockDim.x) {
    const int c = t / nm, q = t - c * nm, im = q / kh, k = q - im * kh;
    double sr = 0.0, si = 0.0;
    if (!(nth % 2 == 0 && im == nth / 2)) {  // the azimuthal Nyquist rows are masked
      const int step = (fft_wavenumber(im, nth) % mth + mth) % mth;
      int idx = 0;
      for (int a = 0; a < mth; ++a) {
        const double2 hv = h[((size_t)c * mth + a) * kh + k], w = tt[idx];
        sr += hv.x * w.x + hv.y * w.y;  // h e^{-i m theta_a}
        si += hv.y * w.x - hv.x * w.y;
        idx += step;
        if (idx >= mth) idx -= mth;
      }
    }
    G[((size_t)c * nm + q) * mr + j] = make_double2(sc * sr, sc * si);
  }
}

}  // namespace

size_t nonlinear_lds(const TcGeo &g) {
  if (g.nth > 1) {
    const size_t nv = 4 * (size_t)g.nf, ng = g.nf == 6 ? 6 : 3;
    const size_t region = std::max(nv * g.nm * 2, ng * g.mth * g.mz + ng * g.mth * g.kh * 2);
    const size_t rows = (g.nf == 6 ? 3 : 1) * 2 * (size_t)g.n;
    return ((size_t)g.mz + g.mth) * sizeof(double2) + (region + nv * g.kh * g.mth * 2 + rows) * sizeof(double);
  }
  const size_t nf = g.nf, ng = g.nf == 6 ? 6 : 3;
  const int nd = 2 * g.nf * g.kh, nch = std::max(1, 256 / nd);
  return ((size_t)g.mz + 3 * nf * g.kh) * sizeof(double2) + (3 * nf + ng) * (size_t)g.mz * sizeof(double) +
         (size_t)nd * nch * sizeof(double2);
}

size_t rhs_lds(const TcGeo &g) {
  // gs, xm (also the solve's partial sums), xl, rhs
  const size_t nfn = (size_t)g.nf * g.n, ng = g.nf == 6 ? 6 : 3;
  return (ng * g.mr + std::max(nfn, (size_t)RHS_THREADS) + 2 * nfn) * sizeof(double2);
}

// Least blocks per mode of the solve kernel: row slices of the solution rows.
constexpr int RHS_SLICES = 4;

size_t diag_lds(const TcGeo &) { return 256 * sizeof(double); }

hipError_t allow_lds(size_t bytes) {
  // gfx950 gives a block up to 160 KB of LDS; raise every kernel's dynamic
  // limit to what the handle needs
  const int v = (int)bytes;
  const void *fns[] = {reinterpret_cast<const void *>(&k_tc_nonlinear<3>),
                       reinterpret_cast<const void *>(&k_tc_nonlinear<6>),
                       reinterpret_cast<const void *>(&k_tc3_nonlinear<3>),
                       reinterpret_cast<const void *>(&k_tc3_nonlinear<6>),
                       reinterpret_cast<const void *>(&k_tc_rhs_solve),
                       reinterpret_cast<const void *>(&k_tc_diag),
                       reinterpret_cast<const void *>(&k_tc_mhd_diag)};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, v);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_nonlinear(const TcGeo &g, const double2 *u, double2 *G, hipStream_t s) {
  const size_t lds = nonlinear_lds(g);
  if (g.nth > 1 && g.nf == 6)
    hipLaunchKernelGGL(k_tc3_nonlinear<6>, dim3(g.mr), dim3(TC3_THREADS), lds, s, g, u, G);
  else if (g.nth > 1)
    hipLaunchKernelGGL(k_tc3_nonlinear<3>, dim3(g.mr), dim3(TC3_THREADS), lds, s, g, u, G);
  else if (g.nf == 6)
    hipLaunchKernelGGL(k_tc_nonlinear<6>, dim3(g.mr), dim3(256), lds, s, g, u, G);
  else
    hipLaunchKernelGGL(k_tc_nonlinear<3>, dim3(g.mr), dim3(256), lds, s, g, u, G);
  return hipGetLastError();
}

hipError_t launch_rhs_solve(const TcGeo &g, const TcRhs &a, hipStream_t s) {
  if (a.nt < 0 || a.nt > TC_TERMS || !a.AinvT || !a.uout || !a.pout) return hipErrorInvalidValue;
  bool lin = false, lexp = false;
  for (int t = 0; t < a.nt; ++t) {
    if (!a.x[t] || a.op[t] < TC_OP_ID || a.op[t] > TC_OP_LEXP) return hipErrorInvalidValue;
    lin |= a.op[t] == TC_OP_LIN;
    lexp |= a.op[t] == TC_OP_LEXP;
  }
  if (lin && lexp) return hipErrorInvalidValue;  // one combined operand per call
  const size_t lds = rhs_lds(g);
  // at least RHS_SLICES row slices per mode, one more per 24 solution rows
  const int n4 = (g.nf + 1) * g.n, slices = std::max(RHS_SLICES, (n4 + 23) / 24);
  hipLaunchKernelGGL(k_tc_rhs_solve, dim3(g.nm, slices), dim3(RHS_THREADS), lds, s, g, a);
  return hipGetLastError();
}

hipError_t launch_diag(const TcGeo &g, const double2 *u, double *out, hipStream_t s) {
  const size_t lds = diag_lds(g);
  if (g.nf == 6)
    hipLaunchKernelGGL(k_tc_mhd_diag, dim3(1), dim3(256), lds, s, g, u, out);
  else
    hipLaunchKernelGGL(k_tc_diag, dim3(1), dim3(256), lds, s, g, u, out);
  return hipGetLastError();
}

}  // namespace tc
