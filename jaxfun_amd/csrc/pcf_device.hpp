// Device helpers shared by the forward and adjoint kernels: complex
// arithmetic, the LDS Stockham FFT passes and the mode -> wavenumber map.
#pragma once

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
  } else if constexpr (R == 8) {
    // two 4-point DFTs of the even and odd samples joined by w^k, w = exp(-+2 pi i / 8)
    V e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
    dft_small<V, T, 4, INV>(e);
    dft_small<V, T, 4, INV>(o);
    const T h = (T)0.70710678118654752440084436210485;  // 1 / sqrt(2)
    const V w1 = INV ? V{(o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h} : V{(o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h};
    const V w2 = INV ? V{-o[2].y, o[2].x} : V{o[2].y, -o[2].x};
    const V w3 = INV ? V{-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h}
                     : V{(o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h};
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], w1);
    a[5] = csub(e[1], w1);
    a[2] = cadd(e[2], w2);
    a[6] = csub(e[2], w2);
    a[3] = cadd(e[3], w3);
    a[7] = csub(e[3], w3);
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
__device__ __forceinline__ void fft_pass(V *buf, const V *tw, int n, int p, int nbatch, int bstride,
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
__device__ __forceinline__ void fft_lds(V *buf, const V *tw, const FFTPlan &plan, int nbatch, int bstride,
                        int estride) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  int p = 1;
  for (int s = 0; s < plan.nst; ++s) {
    const int r = plan.radix[s];
    if (r == 8)
      fft_pass<V, T, 8, INV, BATCH_FAST, MAXB>(buf, tw, plan.n, p, nbatch, bstride, estride, tid, nthr);
    else if (r == 4)
      fft_pass<V, T, 4, INV, BATCH_FAST, MAXB>(buf, tw, plan.n, p, nbatch, bstride, estride, tid, nthr);
    else if (r == 2)
      fft_pass<V, T, 2, INV, BATCH_FAST, MAXB>(buf, tw, plan.n, p, nbatch, bstride, estride, tid, nthr);
    else
      fft_pass<V, T, 3, INV, BATCH_FAST, MAXB>(buf, tw, plan.n, p, nbatch, bstride, estride, tid, nthr);
    p *= r;
  }
}

// Compile-time-length variant: with N, the radix sequence and every stride
// known to the compiler, the index arithmetic of each pass is shifts and
// multiplications instead of integer divisions by runtime values.
struct CPlan {
  int nst;
  int r[12];
};
// Radix 8 first where that saves a pass (fewer LDS round trips for more
// registers per thread: 96 = 8.4.3, 64 = 8.8), else radices 4, 2, 3
// (144 = 4.4.3.3, 48 = 4.4.3); r8 = false keeps radix 4 throughout.
__host__ __device__ constexpr CPlan cplan_with(int n, bool r8) {
  CPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int m = n;
  while (r8 && m % 8 == 0 && p.nst < 12) p.r[p.nst++] = 8, m /= 8;
  while (m % 4 == 0 && p.nst < 12) p.r[p.nst++] = 4, m /= 4;
  while (m % 2 == 0 && p.nst < 12) p.r[p.nst++] = 2, m /= 2;
  while (m % 3 == 0 && p.nst < 12) p.r[p.nst++] = 3, m /= 3;
  return p;
}
__host__ __device__ constexpr CPlan cplan(int n, bool r8 = true) {
  return (r8 && cplan_with(n, true).nst < cplan_with(n, false).nst) ? cplan_with(n, true) : cplan_with(n, false);
}

template <typename V, typename T, int N, int R, int P, bool INV, bool BATCH_FAST, int MAXB>
__device__ __forceinline__ void fft_pass_c(V *buf, const V *tw, int nbatch, int bstride, int estride,
                                           int tid, int nthr) {
  constexpr int nbf = N / R;
  constexpr int twstep = N / (P * R);
  const int total = nbf * nbatch;
  V v[MAXB][R];
#pragma unroll
  for (int t = 0; t < MAXB; ++t) {
    const int idx = tid + t * nthr;
    if (idx < total) {
      const int b = BATCH_FAST ? idx % nbatch : idx / nbf;
      const int i = BATCH_FAST ? idx / nbatch : idx % nbf;
      const V *base = buf + (size_t)b * bstride;
#pragma unroll
      for (int j = 0; j < R; ++j) v[t][j] = base[(i + j * nbf) * estride];
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < MAXB; ++t) {
    const int idx = tid + t * nthr;
    if (idx < total) {
      const int b = BATCH_FAST ? idx % nbatch : idx / nbf;
      const int i = BATCH_FAST ? idx / nbatch : idx % nbf;
      const int k = i % P;
      if constexpr (P > 1) {
#pragma unroll
        for (int j = 1; j < R; ++j) v[t][j] = cmul(v[t][j], tw_get<V, T, INV>(tw, j * k * twstep));
      }
      dft_small<V, T, R, INV>(v[t]);
      V *base = buf + (size_t)b * bstride;
      const int o = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) base[(o + j * P) * estride] = v[t][j];
    }
  }
  __syncthreads();
}

template <typename V, typename T, int N, int S, int P, bool INV, bool BATCH_FAST, int MAXB, bool R8>
__device__ __forceinline__ void fft_stages_c(V *buf, const V *tw, int nbatch, int bstride, int estride,
                                             int tid, int nthr) {
  constexpr CPlan cp = cplan(N, R8);
  if constexpr (S < cp.nst) {
    constexpr int R = cp.r[S];
    fft_pass_c<V, T, N, R, P, INV, BATCH_FAST, MAXB>(buf, tw, nbatch, bstride, estride, tid, nthr);
    fft_stages_c<V, T, N, S + 1, P * R, INV, BATCH_FAST, MAXB, R8>(buf, tw, nbatch, bstride, estride, tid,
                                                                   nthr);
  }
}

template <typename V, typename T, int N, bool INV, bool BATCH_FAST, int MAXB, bool R8 = true>
__device__ __forceinline__ void fft_lds_c(V *buf, const V *tw, int nbatch, int bstride, int estride) {
  fft_stages_c<V, T, N, 0, 1, INV, BATCH_FAST, MAXB, R8>(buf, tw, nbatch, bstride, estride, threadIdx.x,
                                                          blockDim.x);
}

// Block-cooperative gather: element idx < total of the block's staging set is
// load(idx) (a global read, possibly several) and goes to store(idx, value)
// (an LDS write).  B elements per thread are loaded before any is stored, so
// B global round trips overlap instead of following one another.
template <int B, class Load, class Store>
__device__ __forceinline__ void gather_batched(int total, Load load, Store store) {
  using Val = decltype(load(0));
  for (int base = threadIdx.x; base < total; base += B * blockDim.x) {
    Val v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int idx = base + b * blockDim.x;
      if (idx < total) v[b] = load(idx);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int idx = base + b * blockDim.x;
      if (idx < total) store(idx, v[b]);
    }
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

}  // namespace pcf
