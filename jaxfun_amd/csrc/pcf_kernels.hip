// HIP kernels of the plane-Couette KMM SBDF3 step for gfx950 (MI355X).
//
// Stage map of one step (reference call stack examples/channelflow_kmm.py:1016-1064):
//   k_nodal_backward   wall-normal transform of u and curl(u) for every mode
//                      (TensorProductSpace.backward/backward_primitive along x,
//                      tensorproductspace.py:459-594 + Chebyshev.py:263-355)
//   k_fft_y (inverse)  zero-padded streamwise c2c FFT per (field, node, kz)
//                      (Fourier.py:128-151)
//   k_rotational       spanwise c2r FFTs of two y-lines per unit, the
//                      rotational product H = -(u_tot x w_tot)
//                      (channelflow_kmm.py:754-762, pcf_fluctuations_jax.py:108-112,
//                      nonlinear.py:17-29) and spanwise r2c FFTs with truncation
//   k_fft_y (forward)  streamwise FFT with truncation + Nyquist mask
//                      (Fourier.py:170-185, tensorproductspace.py:276-293)
//   k_forward_rhs      Dirichlet projection of H fused with the coefficient-space
//                      KMM rows Nu, Ng, Nv00, Nw00 (channelflow_kmm.py:810-829)
//   k_solve            SBDF3/IMEXRK3 right-hand side, FP64 per-mode Helmholtz and
//                      biharmonic solves, mean-mode solve, velocity
//                      reconstruction (sbdf3.py:44-62, tpmatrix.py:686-799,
//                      channelflow_kmm.py:855-870)
#include "pcf_kernels.hpp"

namespace pcf {

template <typename T>
struct VT;
template <>
struct VT<double> {
  using V = double2;
};
template <>
struct VT<float> {
  using V = float2;
};

template <typename V>
__device__ __forceinline__ V cadd(V a, V b) {
  return V{a.x + b.x, a.y + b.y};
}
template <typename V>
__device__ __forceinline__ V csub(V a, V b) {
  return V{a.x - b.x, a.y - b.y};
}
template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
  return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename V>
__device__ __forceinline__ V cconj(V a) {
  return V{a.x, -a.y};
}

// ------------------------------------------------------------------ FFT ----
// In-place Stockham passes over LDS.  Each thread gathers its butterflies into
// registers, the block synchronises, then writes the permuted outputs; so one
// buffer suffices.  Forward = exp(-i...), inverse = exp(+i...) without scaling.
// Twiddles come from a table tw[t] = (cos, -sin)(2 pi t / n).

template <typename V, typename T, bool INV>
__device__ __forceinline__ V tw_get(const V *tw, int t) {
  V w = tw[t];
  return INV ? V{w.x, -w.y} : w;
}

template <typename V, typename T, int R, bool INV>
__device__ __forceinline__ void dft_small(V *a) {
  if constexpr (R == 2) {
    V t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  } else if constexpr (R == 4) {
    V s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
    V s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
    // forward: y1 = d02 - i d13, y3 = d02 + i d13
    V jd13 = INV ? V{-d13.y, d13.x} : V{d13.y, -d13.x};  // (+-i) d13
    a[0] = cadd(s02, s13);
    a[2] = csub(s02, s13);
    a[1] = cadd(d02, jd13);
    a[3] = csub(d02, jd13);
  } else if constexpr (R == 3) {
    const T c = (T)-0.5;
    const T s = INV ? (T)0.86602540378443864676372317075294 : (T)-0.86602540378443864676372317075294;
    V t1 = cadd(a[1], a[2]), t2 = csub(a[1], a[2]);
    V m = V{a[0].x + c * t1.x, a[0].y + c * t1.y};
    V n = V{-s * t2.y, s * t2.x};  // i s t2
    a[0] = cadd(a[0], t1);
    a[1] = cadd(m, n);
    a[2] = csub(m, n);
  }
}

// batch address: base(b) = b * bstride; element e at base + e * estride.
// BATCH_FAST: consecutive threads take consecutive batches (coalesced for
// column-interleaved layouts); otherwise consecutive butterflies.
template <typename V, typename T, int R, bool INV, bool BATCH_FAST, int MAXB>
__device__ void fft_pass(V *buf, const V *tw, int n, int p, int nbatch, int bstride,
                         int estride, int tid, int nthr) {
  const int nbf = n / R;
  const int total = nbf * nbatch;
  const int twstep = n / (p * R);
  V v[MAXB][R];
#pragma unroll
  for (int t = 0; t < MAXB; ++t) {
    int idx = tid + t * nthr;
    if (idx < total) {
      int b = BATCH_FAST ? idx % nbatch : idx / nbf;
      int i = BATCH_FAST ? idx / nbatch : idx % nbf;
      V *base = buf + (size_t)b * bstride;
#pragma unroll
      for (int j = 0; j < R; ++j) v[t][j] = base[(i + j * nbf) * estride];
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < MAXB; ++t) {
    int idx = tid + t * nthr;
    if (idx < total) {
      int b = BATCH_FAST ? idx % nbatch : idx / nbf;
      int i = BATCH_FAST ? idx / nbatch : idx % nbf;
      int k = i % p;
      if (p > 1) {
#pragma unroll
        for (int j = 1; j < R; ++j) v[t][j] = cmul(v[t][j], tw_get<V, T, INV>(tw, j * k * twstep));
      }
      dft_small<V, T, R, INV>(v[t]);
      V *base = buf + (size_t)b * bstride;
      int o = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) base[(o + j * p) * estride] = v[t][j];
    }
  }
  __syncthreads();
}

template <typename V, typename T, bool INV, bool BATCH_FAST, int MAXB>
__device__ void fft_lds(V *buf, const V *tw, const FFTPlan &plan, int nbatch, int bstride,
                        int estride) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  int p = 1;
  for (int s = 0; s < plan.nst; ++s) {
    const int r = plan.radix[s];
    if (r == 4)
      fft_pass<V, T, 4, INV, BATCH_FAST, MAXB>(buf, tw, plan.n, p, nbatch, bstride, estride, tid, nthr);
    else if (r == 2)
      fft_pass<V, T, 2, INV, BATCH_FAST, MAXB>(buf, tw, plan.n, p, nbatch, bstride, estride, tid, nthr);
    else
      fft_pass<V, T, 3, INV, BATCH_FAST, MAXB>(buf, tw, plan.n, p, nbatch, bstride, estride, tid, nthr);
    p *= r;
  }
}

// ---------------------------------------------------- wavenumber helpers ----
__device__ __forceinline__ void mode_k(const Geo &g, int m, double &ky, double &kz, double &ky_odd) {
  const int iy = m / g.kzs, iz = m - iy * g.kzs;
  const int kyi = (iy < g.ny / 2) ? iy : iy - g.ny;  // FFT order
  ky = kyi * g.df1;
  ky_odd = (iy == g.ny / 2) ? 0.0 : ky;  // odd derivatives drop the Nyquist wavenumber
  kz = iz * g.df2;
}

// -------------------------------------------------- k_nodal_backward ----
// For each mode (lane) and a chunk of XC nodes: U0 = VB u0, U1 = VD u1,
// U2 = VD u2, D1 = VD1 u1, D2 = VD1 u2, then
// w0 = i ky U2 - i kz U1, w1 = i kz U0 - D2, w2 = D1 - i ky U0.
template <typename T, int XC>
__global__ void __launch_bounds__(256) k_nodal_backward(Geo g, NodalMats<T> nm,
                                                        const double2 *__restrict__ u0,
                                                        const double2 *__restrict__ u1,
                                                        const double2 *__restrict__ u2,
                                                        typename VT<T>::V *__restrict__ X) {
  using V = typename VT<T>::V;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int x0 = blockIdx.y * XC;
  if (m >= g.m) return;
  V aU0[XC], aU1[XC], aU2[XC], aD1[XC], aD2[XC];
#pragma unroll
  for (int q = 0; q < XC; ++q) aU0[q] = aU1[q] = aU2[q] = aD1[q] = aD2[q] = V{0, 0};
  for (int i = 0; i < g.nb; ++i) {
    double2 cd = u0[(size_t)i * g.m + m];
    V c{(T)cd.x, (T)cd.y};
    const T *vb = nm.VBt + (size_t)i * nm.ldx + x0;
#pragma unroll
    for (int q = 0; q < XC; ++q) {
      aU0[q].x += vb[q] * c.x;
      aU0[q].y += vb[q] * c.y;
    }
  }
  for (int i = 0; i < g.nd; ++i) {
    double2 ad = u1[(size_t)i * g.m + m], bd = u2[(size_t)i * g.m + m];
    V a{(T)ad.x, (T)ad.y}, b{(T)bd.x, (T)bd.y};
    const T *vd = nm.VDt + (size_t)i * nm.ldx + x0;
    const T *vd1 = nm.VD1t + (size_t)i * nm.ldx + x0;
#pragma unroll
    for (int q = 0; q < XC; ++q) {
      T s = vd[q], d = vd1[q];
      aU1[q].x += s * a.x;
      aU1[q].y += s * a.y;
      aU2[q].x += s * b.x;
      aU2[q].y += s * b.y;
      aD1[q].x += d * a.x;
      aD1[q].y += d * a.y;
      aD2[q].x += d * b.x;
      aD2[q].y += d * b.y;
    }
  }
  double kyd, kzd, kyo;
  mode_k(g, m, kyd, kzd, kyo);
  const T ky = (T)kyo, kz = (T)kzd;
  const size_t plane = (size_t)g.n0 * g.m;
#pragma unroll
  for (int q = 0; q < XC; ++q) {
    const int x = x0 + q;
    if (x < g.n0) {
      const size_t o = (size_t)x * g.m + m;
      V U0 = aU0[q], U1 = aU1[q], U2 = aU2[q];
      // i k c = (-k c.y, k c.x)
      V w0{-ky * U2.y + kz * U1.y, ky * U2.x - kz * U1.x};
      V w1{-kz * U0.y - aD2[q].x, kz * U0.x - aD2[q].y};
      V w2{aD1[q].x + ky * U0.y, aD1[q].y - ky * U0.x};
      X[o] = U0;
      X[plane + o] = U1;
      X[2 * plane + o] = U2;
      X[3 * plane + o] = w0;
      X[4 * plane + o] = w1;
      X[5 * plane + o] = w2;
    }
  }
}

// ---------------------------------------------------------- k_fft_y ----
// One block per (kz chunk, node x, field): the streamwise lines of KC spanwise
// modes.  INV: spectral [iy][kz] -> zero-padded [y][kz] lines, inverse FFT.
// !INV: [y][kz] lines -> forward FFT -> truncated [iy][kz], Nyquist row zeroed.
template <typename T, bool INV>
__global__ void __launch_bounds__(256) k_fft_y(Geo g, const typename VT<T>::V *__restrict__ tw,
                                               FFTPlan plan, int KC,
                                               const typename VT<T>::V *__restrict__ src,
                                               typename VT<T>::V *__restrict__ dst) {
  using V = typename VT<T>::V;
  extern __shared__ __align__(16) unsigned char smem[];
  V *buf = reinterpret_cast<V *>(smem);
  const int kc0 = blockIdx.x * KC, x = blockIdx.y, f = blockIdx.z;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int MY = g.my, NY = g.ny, KZ = g.kzs;
  if (INV) {
    const V *s = src + ((size_t)f * g.n0 + x) * g.m;
    for (int idx = tid; idx < MY * KC; idx += nthr) buf[idx] = V{0, 0};
    __syncthreads();
    for (int idx = tid; idx < NY * KC; idx += nthr) {
      int iy = idx / KC, c = idx - iy * KC;
      int y = (iy < NY / 2) ? iy : iy + MY - NY;
      buf[y * KC + c] = s[(size_t)iy * KZ + kc0 + c];
    }
    __syncthreads();
    fft_lds<V, T, true, true, 8>(buf, tw, plan, KC, 1, KC);
    V *d = dst + ((size_t)f * g.n0 + x) * (size_t)MY * KZ;
    for (int idx = tid; idx < MY * KC; idx += nthr) {
      int y = idx / KC, c = idx - y * KC;
      d[(size_t)y * KZ + kc0 + c] = buf[idx];
    }
  } else {
    const V *s = src + ((size_t)f * g.n0 + x) * (size_t)MY * KZ;
    for (int idx = tid; idx < MY * KC; idx += nthr) {
      int y = idx / KC, c = idx - y * KC;
      buf[idx] = s[(size_t)y * KZ + kc0 + c];
    }
    __syncthreads();
    fft_lds<V, T, false, true, 8>(buf, tw, plan, KC, 1, KC);
    V *d = dst + ((size_t)f * g.n0 + x) * g.m;
    for (int idx = tid; idx < NY * KC; idx += nthr) {
      int iy = idx / KC, c = idx - iy * KC;
      int y = (iy < NY / 2) ? iy : iy + MY - NY;
      V val = (iy == NY / 2) ? V{0, 0} : buf[y * KC + c];
      d[(size_t)iy * KZ + kc0 + c] = val;
    }
  }
}

// ------------------------------------------------------- k_rotational ----
// One unit = two neighbouring y-lines (x, 2*y2) and (x, 2*y2+1) of all six
// nodal fields.  The two real lines of a field ride in one complex FFT
// (line A as the real part, line B as the imaginary part), both for the
// c2r synthesis and the r2c analysis.
template <typename T>
__global__ void __launch_bounds__(256) k_rotational(Geo g, NodalMats<T> nm, FFTPlan plan, int UPB,
                                                    const typename VT<T>::V *__restrict__ Y,
                                                    typename VT<T>::V *__restrict__ Hz) {
  using V = typename VT<T>::V;
  extern __shared__ __align__(16) unsigned char smem[];
  V *buf = reinterpret_cast<V *>(smem);  // [UPB][6][MZ]
  V *out = buf + (size_t)UPB * 6 * g.mz; // [UPB][3][MZ]
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int MZ = g.mz, KZ = g.kzs, MY = g.my;
  const int nunits = g.n0 * (MY / 2);
  const int ubase = blockIdx.x * UPB;
  const size_t fstride = (size_t)g.n0 * MY * KZ;  // field stride in Y / Hz
  const V *tw = reinterpret_cast<const V *>(nm.twz);

  for (int idx = tid; idx < UPB * 6 * MZ; idx += nthr) {
    int lu = idx / (6 * MZ), rem = idx - lu * 6 * MZ;
    int f = rem / MZ, k = rem - f * MZ;
    int u = ubase + lu;
    V z{0, 0};
    if (u < nunits) {
      int x = u / (MY / 2), y2 = u - x * (MY / 2);
      const V *ra = Y + f * fstride + ((size_t)x * MY + 2 * y2) * KZ;
      const V *rb = ra + KZ;
      if (k == 0) {
        z = V{ra[0].x, rb[0].x};
      } else if (k < KZ) {
        V a = ra[k], b = rb[k];
        z = V{a.x - b.y, a.y + b.x};
      } else if (k > MZ - KZ) {
        V a = ra[MZ - k], b = rb[MZ - k];
        z = V{a.x + b.y, b.x - a.y};
      }
    }
    buf[idx] = z;
  }
  __syncthreads();
  fft_lds<V, T, true, false, 12>(buf, tw, plan, UPB * 6, MZ, 1);

  const T U = (T)g.uwall;
  for (int idx = tid; idx < UPB * MZ; idx += nthr) {
    int lu = idx / MZ, n = idx - lu * MZ;
    int u = ubase + lu;
    int x = (u < nunits) ? u / (MY / 2) : 0;
    const T xv = nm.xnode[x];
    const V *b = buf + (size_t)lu * 6 * MZ + n;
    V vu0 = b[0], vu1 = b[MZ], vu2 = b[2 * MZ], vw0 = b[3 * MZ], vw1 = b[4 * MZ], vw2 = b[5 * MZ];
    V h0, h1, h2;
    {
      T a0 = vu0.x, a1 = vu1.x + U * xv, a2 = vu2.x, o0 = vw0.x, o1 = vw1.x, o2 = vw2.x + U;
      h0.x = -(a1 * o2 - a2 * o1);
      h1.x = -(a2 * o0 - a0 * o2);
      h2.x = -(a0 * o1 - a1 * o0);
    }
    {
      T a0 = vu0.y, a1 = vu1.y + U * xv, a2 = vu2.y, o0 = vw0.y, o1 = vw1.y, o2 = vw2.y + U;
      h0.y = -(a1 * o2 - a2 * o1);
      h1.y = -(a2 * o0 - a0 * o2);
      h2.y = -(a0 * o1 - a1 * o0);
    }
    V *o = out + (size_t)lu * 3 * MZ + n;
    o[0] = h0;
    o[MZ] = h1;
    o[2 * MZ] = h2;
  }
  __syncthreads();
  fft_lds<V, T, false, false, 12>(out, tw, plan, UPB * 3, MZ, 1);

  const T scale = (T)(1.0 / ((double)MY * (double)MZ));
  for (int idx = tid; idx < UPB * 3 * KZ; idx += nthr) {
    int lu = idx / (3 * KZ), rem = idx - lu * 3 * KZ;
    int f = rem / KZ, k = rem - f * KZ;
    int u = ubase + lu;
    if (u >= nunits) continue;
    int x = u / (MY / 2), y2 = u - x * (MY / 2);
    const V *P = out + ((size_t)lu * 3 + f) * MZ;
    V p = P[k], q = cconj(P[(MZ - k) % MZ]);
    V ha{(p.x + q.x) * (T)0.5 * scale, (p.y + q.y) * (T)0.5 * scale};
    V hb{(p.y - q.y) * (T)0.5 * scale, -(p.x - q.x) * (T)0.5 * scale};
    V *dst = Hz + f * fstride + ((size_t)x * MY + 2 * y2) * KZ;
    dst[k] = ha;
    dst[KZ + k] = hb;
  }
}

// ----------------------------------------------------- k_forward_rhs ----
// For each mode (lane) and a chunk of IC output rows:
//   Nu = W (A1 (i ky H1 + i kz H2) + A0 (k^2 H0)),  Ng = W AG (i kz H1 - i ky H2),
// and for the mean mode Nv00 = -AG Re H1, Nw00 = -AG Re H2 (1-D mass rows).
template <typename T, int IC>
__global__ void __launch_bounds__(256) k_forward_rhs(Geo g, NodalMats<T> nm,
                                                     const typename VT<T>::V *__restrict__ Hx,
                                                     double2 *__restrict__ Nu,
                                                     double2 *__restrict__ Ng,
                                                     double *__restrict__ Nv,
                                                     double *__restrict__ Nw) {
  using V = typename VT<T>::V;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = blockIdx.y * IC;
  if (m >= g.m) return;
  double kyd, kzd, kyo;
  mode_k(g, m, kyd, kzd, kyo);
  const T ky = (T)kyo, kz = (T)kzd, k2 = (T)(kyd * kyd + kzd * kzd);
  V aU[IC], aG[IC];
  T aV[IC], aW[IC];
#pragma unroll
  for (int q = 0; q < IC; ++q) {
    aU[q] = aG[q] = V{0, 0};
    aV[q] = aW[q] = 0;
  }
  const size_t plane = (size_t)g.n0 * g.m;
  for (int x = 0; x < g.n0; ++x) {
    const size_t o = (size_t)x * g.m + m;
    V h0 = Hx[o], h1 = Hx[plane + o], h2 = Hx[2 * plane + o];
    // c1 = i ky h1 + i kz h2 ; cg = i kz h1 - i ky h2 ; c0 = k^2 h0
    V c1{-ky * h1.y - kz * h2.y, ky * h1.x + kz * h2.x};
    V cg{-kz * h1.y + ky * h2.y, kz * h1.x - ky * h2.x};
    V c0{k2 * h0.x, k2 * h0.y};
    const T *a1 = nm.A1t + (size_t)x * nm.ldi + i0;
    const T *a0 = nm.A0t + (size_t)x * nm.ldi + i0;
    const T *ag = nm.AGt + (size_t)x * nm.ldi + i0;
#pragma unroll
    for (int q = 0; q < IC; ++q) {
      aU[q].x += a1[q] * c1.x + a0[q] * c0.x;
      aU[q].y += a1[q] * c1.y + a0[q] * c0.y;
      aG[q].x += ag[q] * cg.x;
      aG[q].y += ag[q] * cg.y;
      aV[q] += ag[q] * h1.x;
      aW[q] += ag[q] * h2.x;
    }
  }
  const double W = g.W;
#pragma unroll
  for (int q = 0; q < IC; ++q) {
    const int i = i0 + q;
    if (i < g.nb) Nu[(size_t)i * g.m + m] = double2{W * (double)aU[q].x, W * (double)aU[q].y};
    if (i < g.nd) {
      Ng[(size_t)i * g.m + m] = double2{W * (double)aG[q].x, W * (double)aG[q].y};
      if (m == 0) {
        Nv[i] = -(double)aV[q];
        Nw[i] = -(double)aW[q];
      }
    }
  }
}

// ------------------------------------------------------------ k_solve ----
__device__ __forceinline__ double readlane_d(double v, int lane) {
  unsigned long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffull), lane);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One parity block of one per-mode system, solved by one wavefront.
//   A = s0 C0 + s2 C2 + s4 C4                 (n x n, lower bandwidth pbw)
//   rhs_i = sum_j C0_ij delta_J + C2_ij alpha_J + C4_ij beta_J + N_I
// (J = 2 j + par is the full Shen index).  LU without pivoting, as the
// reference's banded kernel (diamatrix.py:1937-1973): rows are reduced in
// order with lanes holding columns; the L multipliers are wave-uniform so the
// forward substitution is a scalar recurrence; the back substitution runs
// column by column with lanes holding rows.
__device__ void wave_solve(int n, int par, int pbw, const double *r0, const double *r2,
                           const double *r4, const double *c0, const double *c2,
                           const double *c4, double s0, double s2, double s4,
                           const double2 *delta, const double2 *alpha, const double2 *beta,
                           const double2 *Nc, double *U, int ldu, double2 *y, double *ipiv,
                           double2 *xout) {
  const int lane = threadIdx.x & 63;
  // right-hand side, lanes = rows
  if (lane < n) {
    double2 r = Nc[2 * lane + par];
    for (int j = 0; j < n; ++j) {
      const int J = 2 * j + par;
      const double a0 = c0[(size_t)j * n + lane], a2 = c2[(size_t)j * n + lane];
      double2 dl = delta[J], al = alpha[J];
      r.x += a0 * dl.x + a2 * al.x;
      r.y += a0 * dl.y + a2 * al.y;
      if (c4) {
        const double a4 = c4[(size_t)j * n + lane];
        double2 be = beta[J];
        r.x += a4 * be.x;
        r.y += a4 * be.y;
      }
    }
    y[lane] = r;
  }
  wave_sync();
  // forward elimination, lanes = columns
  double prev1 = 0.0, prev2 = 0.0, ip1 = 0.0, ip2 = 0.0;
  double2 y1{0, 0}, y2{0, 0};
  for (int i = 0; i < n; ++i) {
    double cur = 0.0;
    if (lane < n) {
      const size_t o = (size_t)i * n + lane;
      cur = s0 * r0[o] + s2 * r2[o];
      if (r4) cur += s4 * r4[o];
    }
    double2 yi = y[i];
    if (pbw >= 2 && i >= 2) {
      const double l = readlane_d(cur, i - 2) * ip2;
      cur -= l * prev2;
      yi.x -= l * y2.x;
      yi.y -= l * y2.y;
    }
    if (i >= 1) {
      const double l = readlane_d(cur, i - 1) * ip1;
      cur -= l * prev1;
      yi.x -= l * y1.x;
      yi.y -= l * y1.y;
    }
    const double ip = 1.0 / readlane_d(cur, i);
    if (lane < n) U[(size_t)i * ldu + lane] = cur;
    if (lane == 0) {
      y[i] = yi;
      ipiv[i] = ip;
    }
    prev2 = prev1;
    prev1 = cur;
    ip2 = ip1;
    ip1 = ip;
    y2 = y1;
    y1 = yi;
  }
  wave_sync();
  // back substitution, lanes = rows
  double2 yv = (lane < n) ? y[lane] : double2{0, 0};
  double2 xv{0, 0};
  for (int j = n - 1; j >= 0; --j) {
    const double ipj = ipiv[j];
    double2 xj{readlane_d(yv.x, j) * ipj, readlane_d(yv.y, j) * ipj};
    if (lane == j) xv = xj;
    if (lane < j) {
      const double u = U[(size_t)lane * ldu + j];
      yv.x -= u * xj.x;
      yv.y -= u * xj.y;
    }
  }
  if (lane < n) xout[2 * lane + par] = xv;
}

// Per-mode implicit update.  One block of four wavefronts per mode:
// wave 0/1 = clamped (u0) system even/odd rows, wave 2/3 = Dirichlet (g)
// system even/odd rows.  The mean-mode block additionally solves the 1-D
// Dirichlet system for (v00, w00) packed as one complex right-hand side.
__global__ void __launch_bounds__(256) k_solve(Geo g, SolveMats sm, SolveArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nb = g.nb, nd = g.nd;
  const int m = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6;
  double2 *alpha_u = reinterpret_cast<double2 *>(smem);
  double2 *beta_u = alpha_u + nb;
  double2 *delta_u = beta_u + nb;
  double2 *N_u = delta_u + nb;
  double2 *alpha_g = N_u + nb;
  double2 *delta_g = alpha_g + nd;
  double2 *N_g = delta_g + nd;
  double2 *x_u = N_g + nd;
  double2 *x_g = x_u + nb;
  double2 *x_m = x_g + nd;         // mean-mode (v00 + i w00) solution
  const int npm = g.npmax, ldu = g.npmax + 1;
  double2 *ybuf = x_m + nd;        // [4][npm]
  double *ipiv = reinterpret_cast<double *>(ybuf + 4 * npm);  // [4][npm]
  double *Ubase = ipiv + 4 * npm;  // [4][npm * ldu]

  double kyd, kzd, kyo;
  mode_k(g, m, kyd, kzd, kyo);
  const double k2 = kyd * kyd + kzd * kzd;
  const double W = g.W, nu = g.nu, gm = a.gamma;

  // z / zL / N combinations into the operator-coefficient arrays.
  for (int idx = tid; idx < nb + nd; idx += blockDim.x) {
    const bool isu = idx < nb;
    const int i = isu ? idx : idx - nb;
    const size_t o = (size_t)i * g.m + m;
    double2 z{0, 0}, zl{0, 0}, nn{0, 0};
    for (int l = 0; l < a.nx; ++l) {
      double2 v = isu ? a.u0[l][o] : a.g[l][o];
      z.x += a.c[l] * v.x;
      z.y += a.c[l] * v.y;
      zl.x += a.d[l] * v.x;
      zl.y += a.d[l] * v.y;
    }
    for (int l = 0; l < a.nn; ++l) {
      double2 v = isu ? a.Nu[l][o] : a.Ng[l][o];
      nn.x += a.e[l] * v.x;
      nn.y += a.e[l] * v.y;
    }
    if (isu) {
      // M = W (C2 - k2 C0), L = W nu (C4 - 2 k2 C2 + k2^2 C0)
      const double wn = W * nu;
      alpha_u[i] = double2{W * z.x - 2 * k2 * wn * zl.x, W * z.y - 2 * k2 * wn * zl.y};
      beta_u[i] = double2{wn * zl.x, wn * zl.y};
      delta_u[i] = double2{-k2 * W * z.x + k2 * k2 * wn * zl.x, -k2 * W * z.y + k2 * k2 * wn * zl.y};
      N_u[i] = nn;
    } else {
      // M = W C0, L = W nu (C2 - k2 C0)
      const double wn = W * nu;
      alpha_g[i] = double2{wn * zl.x, wn * zl.y};
      delta_g[i] = double2{W * z.x - k2 * wn * zl.x, W * z.y - k2 * wn * zl.y};
      N_g[i] = nn;
    }
  }
  __syncthreads();

  {
    const int s = wave >> 1, par = wave & 1;
    const int n = sm.np[s][par];
    double *U = Ubase + (size_t)wave * npm * ldu;
    double2 *y = ybuf + wave * npm;
    double *ip = ipiv + wave * npm;
    if (s == 0) {
      // Su = M - gamma L = W[(1 + 2 gamma nu k2) C2 - gamma nu C4 - (k2 + gamma nu k2^2) C0]
      const double s2 = W * (1 + 2 * gm * nu * k2), s4 = -W * gm * nu, s0 = -W * (k2 + gm * nu * k2 * k2);
      wave_solve(n, par, 2, sm.row[0][par][0], sm.row[0][par][1], sm.row[0][par][2],
                 sm.col[0][par][0], sm.col[0][par][1], sm.col[0][par][2], s0, s2, s4, delta_u,
                 alpha_u, beta_u, N_u, U, ldu, y, ip, x_u);
    } else {
      // Sg = M - gamma L = W[(1 + gamma nu k2) C0 - gamma nu C2]
      const double s0 = W * (1 + gm * nu * k2), s2 = -W * gm * nu;
      wave_solve(n, par, 1, sm.row[1][par][0], sm.row[1][par][1], nullptr, sm.col[1][par][0],
                 sm.col[1][par][1], nullptr, s0, s2, 0.0, delta_g, alpha_g, nullptr, N_g, U, ldu,
                 y, ip, x_g);
    }
  }
  __syncthreads();

  // Mean mode: S00 (v + i w) = M00 z + L00 zL + N with M00 = C0_D, L00 = nu C2_D,
  // S00 = C0_D - gamma nu C2_D (no Fourier weight: a 1-D system).  The g-system
  // coefficient arrays are free again and hold its right-hand-side terms.
  const bool mean = (m == 0);
  if (mean) {
    for (int i = tid; i < nd; i += blockDim.x) {
      double2 z{0, 0}, zl{0, 0}, nn{0, 0};
      for (int l = 0; l < a.nx; ++l) {
        z.x += a.c[l] * a.v[l][i];
        z.y += a.c[l] * a.w[l][i];
        zl.x += a.d[l] * a.v[l][i];
        zl.y += a.d[l] * a.w[l][i];
      }
      for (int l = 0; l < a.nn; ++l) {
        nn.x += a.e[l] * a.Nv[l][i];
        nn.y += a.e[l] * a.Nw[l][i];
      }
      delta_g[i] = z;                              // coefficient of C0
      alpha_g[i] = double2{nu * zl.x, nu * zl.y};  // coefficient of C2
      N_g[i] = nn;
    }
    __syncthreads();
    if (wave >= 2) {
      const int par = wave & 1;
      const int n = sm.np[1][par];
      wave_solve(n, par, 1, sm.row[1][par][0], sm.row[1][par][1], nullptr, sm.col[1][par][0],
                 sm.col[1][par][1], nullptr, 1.0, -gm * nu, 0.0, delta_g, alpha_g, nullptr, N_g,
                 Ubase + (size_t)wave * npm * ldu, ldu, ybuf + wave * npm, ipiv + wave * npm, x_m);
    }
    __syncthreads();
    for (int i = tid; i < nd; i += blockDim.x) {
      a.v_out[i] = x_m[i].x;
      a.w_out[i] = x_m[i].y;
    }
  }
  for (int i = tid; i < nb; i += blockDim.x) a.u0_out[(size_t)i * g.m + m] = x_u[i];

  // reconstruction: f = R u0_new; u1 = i (ky f + kz g)/k2, u2 = i (kz f - ky g)/k2
  const double den = (k2 == 0.0) ? 1.0 : k2;
  const double kyo2 = kyd / den, kzo2 = kzd / den;
  const double2 *xu = x_u;
  for (int i = tid; i < nd; i += blockDim.x) {
    const size_t o = (size_t)i * g.m + m;
    const double2 gi = x_g[i];
    a.g_out[o] = gi;
    if (mean) {
      a.u1[o] = double2{x_m[i].x, 0.0};
      a.u2[o] = double2{x_m[i].y, 0.0};
    } else {
      double2 f{0, 0};
      for (int j = 0; j < nb; ++j) {
        const double r = sm.Rt[(size_t)j * nd + i];
        f.x += r * xu[j].x;
        f.y += r * xu[j].y;
      }
      // i (a) = (-a.y, a.x)
      double2 s1{kyo2 * f.x + kzo2 * gi.x, kyo2 * f.y + kzo2 * gi.y};
      double2 s2{kzo2 * f.x - kyo2 * gi.x, kzo2 * f.y - kyo2 * gi.y};
      a.u1[o] = double2{-s1.y, s1.x};
      a.u2[o] = double2{-s2.y, s2.x};
    }
  }
}

// -------------------------------------------------------------- k_diag ----
__global__ void __launch_bounds__(256) k_diag(Geo g, const double *__restrict__ VB,
                                              const double *__restrict__ VD,
                                              const double *__restrict__ VB1,
                                              const double *__restrict__ VD1,
                                              const double *__restrict__ fejer,
                                              const double *__restrict__ xnode,
                                              const double2 *__restrict__ u0,
                                              const double2 *__restrict__ u1,
                                              const double2 *__restrict__ u2,
                                              double *__restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0, dv = 0.0, cross = 0.0, msh = 0.0;
  if (m < g.m) {
    double kyd, kzd, kyo;
    mode_k(g, m, kyd, kzd, kyo);
    const double wt = ((m % g.kzs) == 0) ? 1.0 : 2.0;
    for (int x = 0; x < g.n0; ++x) {
      double2 U0{0, 0}, U1{0, 0}, U2{0, 0}, D0{0, 0}, D1{0, 0};
      for (int i = 0; i < g.nb; ++i) {
        double2 c = u0[(size_t)i * g.m + m];
        double vb = VB[(size_t)x * g.nb + i], vb1 = VB1[(size_t)x * g.nb + i];
        U0.x += vb * c.x;
        U0.y += vb * c.y;
        D0.x += vb1 * c.x;
        D0.y += vb1 * c.y;
      }
      for (int i = 0; i < g.nd; ++i) {
        double2 a = u1[(size_t)i * g.m + m], b = u2[(size_t)i * g.m + m];
        double vd = VD[(size_t)x * g.nd + i], vd1 = VD1[(size_t)x * g.nd + i];
        U1.x += vd * a.x;
        U1.y += vd * a.y;
        U2.x += vd * b.x;
        U2.y += vd * b.y;
        D1.x += vd1 * a.x;
        D1.y += vd1 * a.y;
      }
      const double w = fejer[x];
      e += w * wt * (U0.x * U0.x + U0.y * U0.y + U1.x * U1.x + U1.y * U1.y + U2.x * U2.x + U2.y * U2.y);
      // div = dU0/dx + i ky U1 + i kz U2
      double2 dvv{D0.x - kyo * U1.y - kzd * U2.y, D0.y + kyo * U1.x + kzd * U2.x};
      dv += w * wt * (dvv.x * dvv.x + dvv.y * dvv.y);
      if (m == 0) {
        cross += w * U1.x * xnode[x];
        msh += w * D1.x;
      }
    }
  }
  // block reduction
  __shared__ double red[4][256];
  red[0][threadIdx.x] = e;
  red[1][threadIdx.x] = dv;
  red[2][threadIdx.x] = cross;
  red[3][threadIdx.x] = msh;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; ++q) atomicAdd(&out[q], red[q][0]);
}

// ------------------------------------------------------------ launchers ----
namespace {
constexpr int XC_NODAL = 8;
constexpr int IC_RHS = 8;

int pick_kc(int kzs) {
  int kc = 16;
  while (kc > 1 && kzs % kc) kc >>= 1;
  return kc;
}
}  // namespace

template <typename T>
hipError_t launch_nodal(const Geo &g, const NodalMats<T> &nm, const double2 *u0, const double2 *u1,
                        const double2 *u2, Workspace &ws, hipStream_t s) {
  using V = typename VT<T>::V;
  dim3 grid((g.m + 255) / 256, (g.n0 + XC_NODAL - 1) / XC_NODAL);
  hipLaunchKernelGGL((k_nodal_backward<T, XC_NODAL>), grid, dim3(256), 0, s, g, nm, u0, u1, u2,
                     reinterpret_cast<V *>(ws.X));
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_y(const Geo &g, const NodalMats<T> &nm, Workspace &ws, const FFTPlan &py,
                        bool inverse, hipStream_t s) {
  using V = typename VT<T>::V;
  const int kc = pick_kc(g.kzs);
  size_t lds = (size_t)g.my * kc * sizeof(V);
  if (inverse) {
    hipLaunchKernelGGL((k_fft_y<T, true>), dim3(g.kzs / kc, g.n0, 6), dim3(256), lds, s, g,
                       reinterpret_cast<const V *>(nm.twy), py, kc,
                       reinterpret_cast<const V *>(ws.X), reinterpret_cast<V *>(ws.Y));
  } else {
    hipLaunchKernelGGL((k_fft_y<T, false>), dim3(g.kzs / kc, g.n0, 3), dim3(256), lds, s, g,
                       reinterpret_cast<const V *>(nm.twy), py, kc,
                       reinterpret_cast<const V *>(ws.Hz), reinterpret_cast<V *>(ws.Hx));
  }
  return hipGetLastError();
}

template <typename T>
int rot_units(const Geo &g) {
  using V = typename VT<T>::V;
  // keep LDS (9 lines of MZ per unit) under 64 KB
  int upb = 4;
  while (upb > 1 && (size_t)upb * 9 * g.mz * sizeof(V) > 65536) upb >>= 1;
  return upb;
}

template <typename T>
hipError_t launch_rotational(const Geo &g, const NodalMats<T> &nm, Workspace &ws,
                             const FFTPlan &pz, hipStream_t s) {
  using V = typename VT<T>::V;
  const int upb = rot_units<T>(g);
  const int nunits = g.n0 * (g.my / 2);
  size_t lds = (size_t)upb * 9 * g.mz * sizeof(V);
  hipLaunchKernelGGL((k_rotational<T>), dim3((nunits + upb - 1) / upb), dim3(256), lds, s, g, nm,
                     pz, upb, reinterpret_cast<const V *>(ws.Y), reinterpret_cast<V *>(ws.Hz));
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rhs(const Geo &g, const NodalMats<T> &nm, Workspace &ws, double2 *Nu,
                      double2 *Ng, double *Nv, double *Nw, hipStream_t s) {
  using V = typename VT<T>::V;
  dim3 grid((g.m + 255) / 256, (g.nd + IC_RHS - 1) / IC_RHS);
  hipLaunchKernelGGL((k_forward_rhs<T, IC_RHS>), grid, dim3(256), 0, s, g, nm,
                     reinterpret_cast<const V *>(ws.Hx), Nu, Ng, Nv, Nw);
  return hipGetLastError();
}

size_t solve_lds_bytes(const Geo &g) {
  return (size_t)(4 * g.nb + 3 * g.nd + g.nb + 2 * g.nd + 4 * g.npmax) * sizeof(double2) +
         (size_t)4 * g.npmax * sizeof(double) +
         (size_t)4 * g.npmax * (g.npmax + 1) * sizeof(double);
}

hipError_t launch_solve(const Geo &g, const SolveMats &sm, const SolveArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_solve, dim3(g.m), dim3(256), solve_lds_bytes(g), s, g, sm, a);
  return hipGetLastError();
}

hipError_t launch_diag(const Geo &g, const double *VB, const double *VD, const double *VB1,
                       const double *VD1, const double *fejer, const double *xnode,
                       const double2 *u0, const double2 *u1, const double2 *u2, double *out,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_diag, dim3((g.m + 255) / 256), dim3(256), 0, s, g, VB, VD, VB1, VD1, fejer,
                     xnode, u0, u1, u2, out);
  return hipGetLastError();
}


bool fft_fits(const Geo &g, const FFTPlan &py, const FFTPlan &pz, int elem_bytes) {
  const int kc = pick_kc(g.kzs);
  if ((size_t)g.my * kc * elem_bytes > 65536) return false;
  for (int s = 0; s < py.nst; ++s) {
    const int total = py.n / py.radix[s] * kc;
    if ((total + 255) / 256 > 8) return false;
  }
  const int upb = (elem_bytes == 16) ? rot_units<double>(g) : rot_units<float>(g);
  if ((size_t)upb * 9 * g.mz * elem_bytes > 65536) return false;
  for (int s = 0; s < pz.nst; ++s) {
    const int total = pz.n / pz.radix[s] * upb * 6;
    if ((total + 255) / 256 > 12) return false;
  }
  return true;
}

#define PCF_INSTANTIATE(T)                                                                     \
  template hipError_t launch_nodal<T>(const Geo &, const NodalMats<T> &, const double2 *,        \
                                      const double2 *, const double2 *, Workspace &, hipStream_t); \
  template hipError_t launch_fft_y<T>(const Geo &, const NodalMats<T> &, Workspace &,            \
                                      const FFTPlan &, bool, hipStream_t);                        \
  template hipError_t launch_rotational<T>(const Geo &, const NodalMats<T> &, Workspace &,       \
                                           const FFTPlan &, hipStream_t);                        \
  template hipError_t launch_rhs<T>(const Geo &, const NodalMats<T> &, Workspace &, double2 *,   \
                                    double2 *, double *, double *, hipStream_t);
PCF_INSTANTIATE(double)
PCF_INSTANTIATE(float)
#undef PCF_INSTANTIATE

}  // namespace pcf
