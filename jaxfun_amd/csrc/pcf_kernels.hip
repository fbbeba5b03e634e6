// HIP kernels of the plane-Couette KMM SBDF3 step for gfx950 (MI355X).
//
// Stage map of one step (reference call stack examples/channelflow_kmm.py:1016-1064):
//   k_nodal_mfma       wall-normal transform of u and curl(u) for every mode
//                      (TensorProductSpace.backward/backward_primitive along x,
//                      tensorproductspace.py:459-594 + Chebyshev.py:263-355);
//                      matrix cores, pcf_mfma.hip
//   k_fft_y (inverse)  zero-padded streamwise c2c FFT per (field, node, kz)
//                      (Fourier.py:128-151)
//   k_rotational       spanwise c2r FFTs of two y-lines per unit, the
//                      rotational product H = -(u_tot x w_tot)
//                      (channelflow_kmm.py:754-762, pcf_fluctuations_jax.py:108-112,
//                      nonlinear.py:17-29) and spanwise r2c FFTs with truncation
//   k_fft_y (forward)  streamwise FFT with truncation + Nyquist mask
//                      (Fourier.py:170-185, tensorproductspace.py:276-293)
//   k_rhs_mfma         Dirichlet projection of H fused with the coefficient-space
//                      KMM rows Nu, Ng, Nv00, Nw00 (channelflow_kmm.py:810-829);
//                      matrix cores, pcf_mfma.hip
//   k_solve_factored   (pcf_factored.hip) right-hand side and FP64 per-mode
//                      Helmholtz / biharmonic / mean-mode solves on stored
//                      factors (sbdf3.py:44-62, tpmatrix.py:686-799)
//   k_recon_mfma       (pcf_mfma.hip) velocity reconstruction (channelflow_kmm.py:855-870)
//   k_solve_bordered   insulating-wall A_y / A_z update
#include <algorithm>

#include "pcf_adjoint.hpp"
#include "pcf_device.hpp"

namespace pcf {

// ---------------------------------------------------------- k_fft_y ----
// One block per (kz chunk, node x, field): the streamwise lines of KC spanwise
// modes.  INV: spectral [iy][kz] -> zero-padded [y][kz] lines, inverse FFT.
// !INV: [y][kz] lines -> forward FFT -> truncated [iy][kz], Nyquist row zeroed.
// NYC, KCC > 0: padded streamwise length and column batch fixed at compile
// time (the launcher dispatches the common values); 0 reads them at run time.
template <typename T, bool INV, int MAXB, int NYC, int KCC>
__global__ void __launch_bounds__(256) k_fft_y(Geo g, const typename VT<T>::V *__restrict__ tw,
                                               FFTPlan plan, int KCr,
                                               const typename VT<T>::V *__restrict__ src,
                                               typename VT<T>::V *__restrict__ dst) {
  using V = typename VT<T>::V;
  extern __shared__ __align__(16) unsigned char smem[];
  V *buf = reinterpret_cast<V *>(smem);
  const int KC = KCC ? KCC : KCr;
  const int kc0 = blockIdx.x * KC, x = blockIdx.y, f = blockIdx.z;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int MY = NYC ? NYC : g.my, NY = g.ny, KZ = g.kzs;
  V *twl = buf + (size_t)MY * KC;  // twiddle table in LDS: the passes never touch global memory
  for (int i = tid; i < MY; i += nthr) twl[i] = tw[i];
  if (INV) {
    const V *s = src + ((size_t)f * g.n0 + x) * g.m;
    for (int idx = tid; idx < MY * KC; idx += nthr) buf[idx] = V{0, 0};
    __syncthreads();
    gather_batched<4>(
        NY * KC, [&](int idx) { return s[(size_t)(idx / KC) * KZ + kc0 + idx % KC]; },
        [&](int idx, const V &v) {
          const int iy = idx / KC, c = idx - iy * KC;
          const int y = (iy < NY / 2) ? iy : iy + MY - NY;
          buf[y * KC + c] = v;
        });
    __syncthreads();
    if constexpr (NYC > 0)
      fft_lds_c<V, T, NYC, true, true, MAXB>(buf, twl, KC, 1, KC);
    else
      fft_lds<V, T, true, true, MAXB>(buf, twl, plan, KC, 1, KC);
    V *d = dst + ((size_t)f * g.n0 + x) * (size_t)MY * KZ;
    for (int idx = tid; idx < MY * KC; idx += nthr) {
      int y = idx / KC, c = idx - y * KC;
      d[(size_t)y * KZ + kc0 + c] = buf[idx];
    }
  } else {
    const V *s = src + ((size_t)f * g.n0 + x) * (size_t)MY * KZ;
    gather_batched<4>(
        MY * KC, [&](int idx) { return s[(size_t)(idx / KC) * KZ + kc0 + idx % KC]; },
        [&](int idx, const V &v) { buf[idx] = v; });
    __syncthreads();
    if constexpr (NYC > 0)
      fft_lds_c<V, T, NYC, false, true, MAXB>(buf, twl, KC, 1, KC);
    else
      fft_lds<V, T, false, true, MAXB>(buf, twl, plan, KC, 1, KC);
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
// NZC > 0: the padded spanwise length is the compile-time NZC (the launcher
// dispatches the common lengths); NZC = 0 reads it from g.
template <typename T, int MAXB, int NZC>
__global__ void __launch_bounds__(256) k_rotational(Geo g, NodalMats<T> nm, FFTPlan plan, int UPB,
                                                    const typename VT<T>::V *__restrict__ Y,
                                                    typename VT<T>::V *__restrict__ Hz) {
  using V = typename VT<T>::V;
  extern __shared__ __align__(16) unsigned char smem[];
  const int MZ = NZC ? NZC : g.mz;
  V *buf = reinterpret_cast<V *>(smem);  // [UPB][6][MZ]
  V *out = buf + (size_t)UPB * 6 * MZ;   // [UPB][3][MZ]
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int KZ = g.kzs, MY = g.my;
  const int nunits = g.n0 * (MY / 2);
  const int ubase = blockIdx.x * UPB;
  const size_t fstride = (size_t)g.n0 * MY * KZ;  // field stride in Y / Hz
  V *tw = out + (size_t)UPB * 3 * g.mz;            // twiddle table in LDS
  for (int i = tid; i < MZ; i += nthr) tw[i] = reinterpret_cast<const V *>(nm.twz)[i];

  gather_batched<4>(
      UPB * 6 * MZ,
      [&](int idx) {
        const int lu = idx / (6 * MZ), rem = idx - lu * 6 * MZ;
        const int f = rem / MZ, k = rem - f * MZ;
        const int u = ubase + lu;
        V z{0, 0};
        if (u < nunits) {
          const int x = u / (MY / 2), y2 = u - x * (MY / 2);
          const V *ra = Y + f * fstride + ((size_t)x * MY + 2 * y2) * KZ;
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
  // radix-8 passes in FP32 only: in FP64 their registers cost more occupancy
  // than the saved LDS pass (compile-time lengths)
  if constexpr (NZC > 0)
    fft_lds_c<V, T, NZC, true, false, MAXB, sizeof(T) == 4>(buf, tw, UPB * 6, MZ, 1);
  else
    fft_lds<V, T, true, false, MAXB>(buf, tw, plan, UPB * 6, MZ, 1);

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
  if constexpr (NZC > 0)
    fft_lds_c<V, T, NZC, false, false, MAXB, sizeof(T) == 4>(out, tw, UPB * 3, MZ, 1);
  else
    fft_lds<V, T, false, false, MAXB>(out, tw, plan, UPB * 3, MZ, 1);

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

// ------------------------------------------------------- k_rot_wave ----
// k_rotational with the spanwise FFTs held in registers: KZ lanes (a power of
// two, KZ = nz / 2) own one y-line pair, and the padded length MZ = 3 KZ is
// factored as 3 x KZ.  Lane l holds entries l, l + KZ, l + 2 KZ of each line.
//   synthesis (exp +):  a 3-point DFT in the lane (the middle entry is the
//     zero padding), twiddles exp(+2 pi i l na / MZ), then a KZ-point DFT
//     across the lanes, radix 2 with decimation in frequency (__shfl_xor);
//     lane l ends with the physical point na + 3 bitrev(l) in register na;
//   the product H = -(u_tot x w_tot) point by point, in that order;
//   analysis (exp -):   the KZ-point DFT across the lanes with decimation in
//     time (bit-reversed input, natural output), twiddles exp(-2 pi i l na /
//     MZ) and the 3-point DFT in the lane give entries l and 2 KZ + l, the two
//     that survive the truncation to |kz| < KZ.
// The two real lines of a pair travel as the real and imaginary parts of one
// complex line (as in k_rotational).  No LDS and no barrier: a wave holds
// 64 / KZ pairs.
// The value of lane (lane ^ H), H a power of two up to 32.  FP32: without LDS
// traffic, by DPP quad permutes (H = 1, 2), row shifts (H = 4: the lower lane
// of a pair reads 4 lanes up, the upper 4 lanes down), a rotation of the
// 16-lane row by 8 (H = 8) and the gfx950 lane swaps v_permlane16_swap /
// v_permlane32_swap (H = 16, 32).
template <int H>
__device__ __forceinline__ unsigned xchg_u32(unsigned v, int lane) {
  if constexpr (H == 1) {
    return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
  } else if constexpr (H == 2) {
    return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
  } else if constexpr (H == 8) {
    return (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, true);  // row_ror:8 (half a row)
  } else if constexpr (H == 4) {
    // lanes shifted in from outside the row read 0 (bound_ctrl) and are not selected
    const unsigned up = (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x100 + H, 0xF, 0xF, true);  // row_shl
    const unsigned dn = (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x110 + H, 0xF, 0xF, true);  // row_shr
    return (lane & H) ? dn : up;
  } else if constexpr (H == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? r[0] : r[1];
  } else {
    static_assert(H == 32, "lane exchange distance");
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? r[0] : r[1];
  }
}

template <int H>
__device__ __forceinline__ float xchg(float v, int lane) {
  return __builtin_bit_cast(float, xchg_u32<H>(__builtin_bit_cast(unsigned, v), lane));
}
// FP64 values as two 32-bit halves through the same exchanges: with the
// branch-free butterflies below this ran 18.5 us in the 65x64x64 step against
// 20.3 us through ds_bpermute (the LDS round trips then bound the stages).
template <int H>
__device__ __forceinline__ double xchg(double v, int lane) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = xchg_u32<H>((unsigned)b, lane), hi = xchg_u32<H>((unsigned)(b >> 32), lane);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int H, typename V>
__device__ __forceinline__ V xchg_c(V v, int lane) {
  return V{xchg<H>(v.x, lane), xchg<H>(v.y, lane)};
}

// Radix-2 stages of the KZ-point DFT across the lanes.  Synthesis (DIF,
// natural order in, bit-reversed out) runs H = KZ/2 .. 1 on the 6 x 3 values of
// the lane; analysis (DIT, bit-reversed in, natural out) runs H = 1 .. KZ/2 on
// 3 x 3.  ws[log2 H] is the lane's twiddle exp(-2 pi i (l mod H) / (2 H)).
// Branch-free butterflies: the lower lane of a pair (bit H clear) and the upper
// one differ only in a sign s = +-1 and a twiddle that is 1 in the lower lane,
// both fixed per stage, so every lane runs the same instructions and a stage
// issues all its exchanges before the first result is needed.
//   DIF: t = p + s v, v' = w t       (lower a + b, upper (a - b) w)
//   DIT: t = w v exchanged as p, v' = p + s t   (lower a + w b, upper a - w b)
template <typename V, int H, int NF>
__device__ __forceinline__ void dif_stages(V (&v)[NF][3], const V *ws, int lane) {
  if constexpr (H >= 1) {
    using T = decltype(v[0][0].x);
    const bool up = (lane & H) != 0;
    const T sg = up ? (T)-1 : (T)1;
    const V w = up ? cconj(ws[__builtin_ctz(H)]) : V{(T)1, (T)0};
    V p[NF][3];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int na = 0; na < 3; ++na) p[f][na] = xchg_c<H>(v[f][na], lane);
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int na = 0; na < 3; ++na) {
        const V t{sg * v[f][na].x + p[f][na].x, sg * v[f][na].y + p[f][na].y};
        v[f][na] = cmul(t, w);
      }
    dif_stages<V, H / 2, NF>(v, ws, lane);
  }
}
template <typename V, int H, int KZ, int NF>
__device__ __forceinline__ void dit_stages(V (&v)[NF][3], const V *ws, int lane) {
  if constexpr (H < KZ) {
    using T = decltype(v[0][0].x);
    const bool up = (lane & H) != 0;
    const T sg = up ? (T)-1 : (T)1;
    const V w = up ? ws[__builtin_ctz(H)] : V{(T)1, (T)0};
    V p[NF][3];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int na = 0; na < 3; ++na) {
        v[f][na] = cmul(w, v[f][na]);
        p[f][na] = xchg_c<H>(v[f][na], lane);
      }
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int na = 0; na < 3; ++na)
        v[f][na] = V{sg * v[f][na].x + p[f][na].x, sg * v[f][na].y + p[f][na].y};
    dit_stages<V, 2 * H, KZ, NF>(v, ws, lane);
  }
}

template <typename T, int KZ>
__global__ void __launch_bounds__(128) k_rot_wave(Geo g, NodalMats<T> nm, const typename VT<T>::V *__restrict__ Y,
                                                  typename VT<T>::V *__restrict__ Hz) {
  using V = typename VT<T>::V;
  constexpr int MZ = 3 * KZ, GROUPS = 64 / KZ;
  const int lane = threadIdx.x & 63, l = lane & (KZ - 1), grp = lane / KZ;
  const int wave = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  const int MY = g.my;
  const int nunits = g.n0 * (MY / 2);
  const int u = wave * GROUPS + grp;
  const bool live = u < nunits;
  const int uu = live ? u : 0;  // idle lanes read a valid line and store nothing
  const int x = uu / (MY / 2), y2 = uu - x * (MY / 2);
  const size_t fstride = (size_t)g.n0 * MY * KZ;
  const size_t line = ((size_t)x * MY + 2 * y2) * KZ;
  const V *tw = reinterpret_cast<const V *>(nm.twz);  // tw[t] = exp(-2 pi i t / MZ)
  const V w1 = tw[l], w2 = tw[2 * l];
  // the lane's twiddle of each radix-2 stage, exp(-2 pi i (l mod h) / (2 h))
  V ws[6];
#pragma unroll
  for (int s = 0, h = 1; h < KZ; ++s, h <<= 1) ws[s] = tw[(l & (h - 1)) * (MZ / (2 * h))];
  const T c3 = (T)-0.5, s3 = (T)0.86602540378443864676372317075294;  // exp(+2 pi i / 3)

  // every load issued before the first use: one memory round trip
  const int lr = (KZ - l) & (KZ - 1);  // mirrored entry (lane 0 reads entry 0, unused)
  const T xv = nm.xnode[x];
  V la[6], lb[6], ma[6], mb[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const V *ra = Y + f * fstride + line, *rb = ra + KZ;
    la[f] = ra[l];
    lb[f] = rb[l];
    ma[f] = ra[lr];
    mb[f] = rb[lr];
  }
  V v[6][3];
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    const V a = la[f], b = lb[f], a2 = ma[f], b2 = mb[f];
    const V z0 = (l == 0) ? V{a.x, b.x} : V{a.x - b.y, a.y + b.x};
    const V z2 = (l == 0) ? V{0, 0} : V{a2.x + b2.y, b2.x - a2.y};
    // 3-point synthesis with a zero middle entry: t1 = z0 + w^2 z2, t2 = z0 + w z2
    const V wz2{c3 * z2.x - s3 * z2.y, c3 * z2.y + s3 * z2.x};   // exp(+2 pi i / 3) z2
    const V w2z2{c3 * z2.x + s3 * z2.y, c3 * z2.y - s3 * z2.x};  // exp(+4 pi i / 3) z2
    v[f][0] = cadd(z0, z2);
    v[f][1] = cmul(cadd(z0, w2z2), cconj(w1));
    v[f][2] = cmul(cadd(z0, wz2), cconj(w2));
  }
  // KZ-point synthesis across the lanes, decimation in frequency
  dif_stages<V, KZ / 2, 6>(v, ws, l);
  // the product at every point the lane holds
  const T U = (T)g.uwall;
  V hh[3][3];
#pragma unroll
  for (int na = 0; na < 3; ++na) {
    const V vu0 = v[0][na], vu1 = v[1][na], vu2 = v[2][na], vw0 = v[3][na], vw1 = v[4][na], vw2 = v[5][na];
    {
      const T a0 = vu0.x, a1 = vu1.x + U * xv, a2 = vu2.x, o0 = vw0.x, o1 = vw1.x, o2 = vw2.x + U;
      hh[0][na].x = -(a1 * o2 - a2 * o1);
      hh[1][na].x = -(a2 * o0 - a0 * o2);
      hh[2][na].x = -(a0 * o1 - a1 * o0);
    }
    {
      const T a0 = vu0.y, a1 = vu1.y + U * xv, a2 = vu2.y, o0 = vw0.y, o1 = vw1.y, o2 = vw2.y + U;
      hh[0][na].y = -(a1 * o2 - a2 * o1);
      hh[1][na].y = -(a2 * o0 - a0 * o2);
      hh[2][na].y = -(a0 * o1 - a1 * o0);
    }
  }
  // KZ-point analysis across the lanes, decimation in time
  dit_stages<V, 1, KZ, 3>(hh, ws, l);
  const T scale = (T)(0.5 / ((double)MY * (double)MZ));
  const int src = grp * KZ + ((KZ - l) & (KZ - 1));
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    // entries l (k = l) and 2 KZ + l of the analysed line
    const V q0 = hh[f][0], q1 = cmul(hh[f][1], w1), q2 = cmul(hh[f][2], w2);
    const V P0 = cadd(cadd(q0, q1), q2);
    const V wq1{c3 * q1.x - s3 * q1.y, c3 * q1.y + s3 * q1.x};   // exp(+2 pi i / 3) q1
    const V w2q2{c3 * q2.x + s3 * q2.y, c3 * q2.y - s3 * q2.x};  // exp(+4 pi i / 3) q2
    const V P2 = cadd(cadd(q0, wq1), w2q2);
    // P[MZ - k] = P2 of lane KZ - k (k >= 1); P[0] for k = 0
    const V r{__shfl(P2.x, src, 64), __shfl(P2.y, src, 64)};
    const V q = cconj(l == 0 ? P0 : r);
    const V ha{(P0.x + q.x) * scale, (P0.y + q.y) * scale};
    const V hb{(P0.y - q.y) * scale, -(P0.x - q.x) * scale};
    if (live) {
      V *dst = Hz + f * fstride + line;
      dst[l] = ha;
      dst[KZ + l] = hb;
    }
  }
}

// ------------------------------------------------------- k_absmax_z ----
// Largest |value| of nf physical fields on the unpadded grid (g is the
// unpadded geometry, MY = ny, MZ = nz), for the CFL numbers of
// production/health.py:120-155: the spanwise c2r FFT of y-line pairs as in
// k_rotational, plus the base flow U x on field 1 (streamwise velocity) and
// B0 on fields 3..5 (magnetic field), then a block maximum per field into
// blockmax[block][nf].  Units: the line pairs (x, 2 y2), (x, 2 y2 + 1).
template <typename T, int MAXB>
__global__ void __launch_bounds__(256) k_absmax_z(Geo g, const typename VT<T>::V *__restrict__ twz, FFTPlan plan,
                                                  int UPB, int nf, const T *__restrict__ xnode,
                                                  const typename VT<T>::V *__restrict__ Y,
                                                  double *__restrict__ blockmax, double *__restrict__ phys) {
  using V = typename VT<T>::V;
  extern __shared__ __align__(16) unsigned char smem[];
  const int MZ = g.mz, MY = g.my, KZ = g.kzs;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nunits = g.n0 * (MY / 2);
  const int ubase = blockIdx.x * UPB;
  const size_t fstride = (size_t)g.n0 * MY * KZ;
  V *buf = reinterpret_cast<V *>(smem);            // [UPB][nf][MZ]
  V *tw = buf + (size_t)UPB * nf * MZ;             // [MZ]
  double *red = reinterpret_cast<double *>(tw + MZ);  // [waves][6]
  for (int i = tid; i < MZ; i += nthr) tw[i] = twz[i];
  gather_batched<4>(
      UPB * nf * MZ,
      [&](int idx) {
        const int lu = idx / (nf * MZ), rem = idx - lu * nf * MZ;
        const int f = rem / MZ, k = rem - f * MZ;
        const int u = ubase + lu;
        V z{0, 0};
        if (u < nunits) {
          const int x = u / (MY / 2), y2 = u - x * (MY / 2);
          const V *ra = Y + f * fstride + ((size_t)x * MY + 2 * y2) * KZ;
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
  fft_lds<V, T, true, false, MAXB>(buf, tw, plan, UPB * nf, MZ, 1);
  if (phys) {
    // the fluctuation fields themselves, [field][x][y][z] (the two lines of a
    // unit are y = 2 y2 and 2 y2 + 1)
    for (int idx = tid; idx < UPB * nf * MZ; idx += nthr) {
      const int lu = idx / (nf * MZ), rem = idx - lu * nf * MZ, f = rem / MZ, z = rem - f * MZ;
      const int u = ubase + lu;
      if (u >= nunits) continue;
      const int x = u / (MY / 2), y2 = u - x * (MY / 2);
      const V v = buf[idx];
      double *o = phys + (((size_t)f * g.n0 + x) * MY + 2 * y2) * MZ + z;
      o[0] = (double)v.x;
      o[MZ] = (double)v.y;
    }
  }
  double mx[6] = {0, 0, 0, 0, 0, 0};
  for (int idx = tid; idx < UPB * nf * MZ; idx += nthr) {
    const int lu = idx / (nf * MZ), f = (idx - lu * nf * MZ) / MZ;
    const int u = ubase + lu;
    if (u >= nunits) continue;
    const int x = u / (MY / 2);
    const double add = (f == 1) ? g.uwall * (double)xnode[x] : (f >= 3) ? g.b0[f - 3] : 0.0;
    const V v = buf[idx];
    const double m = fmax(fabs((double)v.x + add), fabs((double)v.y + add));
#pragma unroll
    for (int q = 0; q < 6; ++q)
      if (q == f) mx[q] = fmax(mx[q], m);
  }
  const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double v = mx[q];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if (lane == 0) red[wave * 6 + q] = v;
  }
  __syncthreads();
  if (tid < nf) {
    double m = 0.0;
    for (int w = 0; w < (nthr >> 6); ++w) m = fmax(m, red[w * 6 + tid]);
    blockmax[(size_t)blockIdx.x * nf + tid] = m;
  }
}

// MHD counterpart of k_rotational: twelve input fields (u, w, b, j) per
// y-line and six outputs, H = -(u_tot x w_tot) + (-2 Omega u1, 2 Omega u0, 0)
// - j x b_tot and E = u_tot x b_tot with b_tot = b + B0
// (examples/pcf_mhd_jax.py:339-357; Coriolis force and background B0 of
// examples/pcf_mhd_mri_shearpy_jax.py:149-176).
template <typename T, int MAXB, int NZC>
__global__ void __launch_bounds__(256) k_rotational_mhd(Geo g, NodalMats<T> nm, FFTPlan plan, int UPB,
                                                        const typename VT<T>::V *__restrict__ Y,
                                                    typename VT<T>::V *__restrict__ Hz) {
  using V = typename VT<T>::V;
  constexpr int NIN = 12, NOUT = 6;
  extern __shared__ __align__(16) unsigned char smem[];
  const int MZ = NZC ? NZC : g.mz;
  V *buf = reinterpret_cast<V *>(smem);   // [UPB][NIN][MZ]
  V *out = buf + (size_t)UPB * NIN * MZ;  // [UPB][NOUT][MZ]
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int KZ = g.kzs, MY = g.my;
  const int nunits = g.n0 * (MY / 2);
  const int ubase = blockIdx.x * UPB;
  const size_t fstride = (size_t)g.n0 * MY * KZ;  // field stride in Y / Hz
  V *tw = out + (size_t)UPB * NOUT * g.mz;         // twiddle table in LDS
  for (int i = tid; i < MZ; i += nthr) tw[i] = reinterpret_cast<const V *>(nm.twz)[i];

  gather_batched<4>(
      UPB * NIN * MZ,
      [&](int idx) {
        const int lu = idx / (NIN * MZ), rem = idx - lu * NIN * MZ;
        const int f = rem / MZ, k = rem - f * MZ;
        const int u = ubase + lu;
        V z{0, 0};
        if (u < nunits) {
          const int x = u / (MY / 2), y2 = u - x * (MY / 2);
          const V *ra = Y + f * fstride + ((size_t)x * MY + 2 * y2) * KZ;
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
    fft_lds_c<V, T, NZC, true, false, MAXB, sizeof(T) == 4>(buf, tw, UPB * NIN, MZ, 1);
  else
    fft_lds<V, T, true, false, MAXB>(buf, tw, plan, UPB * NIN, MZ, 1);

  const T U = (T)g.uwall;
  for (int idx = tid; idx < UPB * MZ; idx += nthr) {
    int lu = idx / MZ, n = idx - lu * MZ;
    int u = ubase + lu;
    int x = (u < nunits) ? u / (MY / 2) : 0;
    const T xv = nm.xnode[x];
    const V *b = buf + (size_t)lu * NIN * MZ + n;
    V *o = out + (size_t)lu * NOUT * MZ + n;
    V in[NIN], res[NOUT];
#pragma unroll
    for (int f = 0; f < NIN; ++f) in[f] = b[f * MZ];
    // the two real lines ride in the real and imaginary parts
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      auto P = [&](int f) { return part ? in[f].y : in[f].x; };
      const T a0 = P(0), a1 = P(1) + U * xv, a2 = P(2), o0 = P(3), o1 = P(4), o2 = P(5) + U;
      T h0 = -(a1 * o2 - a2 * o1), h1 = -(a2 * o0 - a0 * o2), h2 = -(a0 * o1 - a1 * o0);
      T e0 = 0, e1 = 0, e2 = 0;
      {
        const T B0x = (T)g.b0[0], B0y = (T)g.b0[1], B0z = (T)g.b0[2], O2 = (T)g.omega2;
        // Coriolis force (-2 Omega u1, 2 Omega u0, 0) on the fluctuation velocity
        h0 -= O2 * P(1);
        h1 += O2 * P(0);
        const T c0 = P(6) + B0x, c1 = P(7) + B0y, c2 = P(8) + B0z, j0 = P(9), j1 = P(10), j2 = P(11);
        h0 -= j1 * c2 - j2 * c1;
        h1 -= j2 * c0 - j0 * c2;
        h2 -= j0 * c1 - j1 * c0;
        e0 = a1 * c2 - a2 * c1;
        e1 = a2 * c0 - a0 * c2;
        e2 = a0 * c1 - a1 * c0;
      }
      const T r[6] = {h0, h1, h2, e0, e1, e2};
#pragma unroll
      for (int f = 0; f < NOUT; ++f) {
        if (part)
          res[f].y = r[f];
        else
          res[f].x = r[f];
      }
    }
#pragma unroll
    for (int f = 0; f < NOUT; ++f) o[f * MZ] = res[f];
  }
  __syncthreads();
  if constexpr (NZC > 0)
    fft_lds_c<V, T, NZC, false, false, MAXB, sizeof(T) == 4>(out, tw, UPB * NOUT, MZ, 1);
  else
    fft_lds<V, T, false, false, MAXB>(out, tw, plan, UPB * NOUT, MZ, 1);

  const T scale = (T)(1.0 / ((double)MY * (double)MZ));
  for (int idx = tid; idx < UPB * NOUT * KZ; idx += nthr) {
    int lu = idx / (NOUT * KZ), rem = idx - lu * NOUT * KZ;
    int f = rem / KZ, k = rem - f * KZ;
    int u = ubase + lu;
    if (u >= nunits) continue;
    int x = u / (MY / 2), y2 = u - x * (MY / 2);
    const V *P = out + ((size_t)lu * NOUT + f) * MZ;
    V p = P[k], q = cconj(P[(MZ - k) % MZ]);
    V ha{(p.x + q.x) * (T)0.5 * scale, (p.y + q.y) * (T)0.5 * scale};
    V hb{(p.y - q.y) * (T)0.5 * scale, -(p.x - q.x) * (T)0.5 * scale};
    V *dst = Hz + f * fstride + ((size_t)x * MY + 2 * y2) * KZ;
    dst[k] = ha;
    dst[KZ + k] = hb;
  }
}

// Insulating walls: the A_y / A_z update (examples/pcf_mhd_mri_shearpy_jax.py:559-712).
// On the orthogonal Chebyshev basis the radial operator
// h - gamma eta ((T_i, T_j'') - k^2 h) is diagonal plus a strictly upper part
// that is rank 2 within each parity: (T_i, T_j'') = (pi/2) j (j^2 - i^2) for
// j > i, j - i even (both i = 0 and i > 0, since h_0 = pi carries the
// 1/c_0 = 1/2 of the derivative recurrence).  The two wall rows replacing the
// last two Galerkin rows separate by parity into sum_{j in parity} beta_j x_j = 0
// with beta_j = j^2 (P: P'(+-1) = 0) or j^2 + k (Q: Q'(-1) - k Q(-1) =
// Q'(1) + k Q(1) = 0), because T_j'(+-1) = (+-1)^{j+1} j^2 and
// T_j(+-1) = (+-1)^j.  So each parity block is upper triangular with its last
// row replaced by beta: with t = x_last, the back substitution gives
// x = a + b t, and beta . x = 0 fixes t — two top-down passes, O(n) each.
// Block: BS modes.  Phase 0 (all threads, coalesced over modes): the local part
// of the right-hand side h z - eta k^2 h zL + N and zL of both components,
// rotated to P = cy Ay + cz Az, Q = cz Ay - cy Az, into LDS.  Phase 1 (one
// thread per mode, parity and P/Q): the suffix-sum term
// eta (pi/2)(S3 - i^2 S1) of zL and the bordered recurrence, out of LDS.
// Phase 2 (all threads): rotate back, mask the Nyquist row, store.
template <int BS>
__global__ void __launch_bounds__(128) k_solve_bordered(Geo g, BorderedArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int n = g.n0, M = g.m;
  const int m0 = blockIdx.x * BS;
  const int tid = threadIdx.x, nthr = blockDim.x;
  double2 *Lx = reinterpret_cast<double2 *>(smem);  // [row][P/Q][lane]: local rhs, then a, then x
  double2 *Zl = Lx + (size_t)n * 2 * BS;             // [row][P/Q][lane]: zL
  const double eta = g.nu, ge = a.gamma * eta, C = ge * 0.5 * M_PI;
  const size_t cs = a.cstride;
  auto rot = [&](int m, double &cy, double &cz, double &k2, double &k) {
    double kyd, kzd, kyo;
    mode_k(g, m, kyd, kzd, kyo);
    k2 = kyd * kyd + kzd * kzd;
    k = sqrt(k2);
    cy = (k > 0.0) ? kyd / k : 1.0;
    cz = (k > 0.0) ? kzd / k : 0.0;
  };
  // phase 0
  for (int idx = tid; idx < n * BS; idx += nthr) {
    const int I = idx / BS, lane = idx - I * BS, m = m0 + lane;
    double2 lp{0, 0}, lq{0, 0}, zp{0, 0}, zq{0, 0};
    if (m < M) {
      double cy, cz, k2, k;
      rot(m, cy, cz, k2, k);
      const double hI = (I == 0) ? M_PI : 0.5 * M_PI;
      const size_t o = (size_t)I * M + m;
      double2 zy{0, 0}, zly{0, 0}, ny{0, 0}, zz{0, 0}, zlz{0, 0}, nz{0, 0};
      for (int l = 0; l < a.nx; ++l) {
        const double2 vy = a.x[l][cs + o], vz = a.x[l][2 * cs + o];
        zy.x += a.c[l] * vy.x;
        zy.y += a.c[l] * vy.y;
        zly.x += a.d[l] * vy.x;
        zly.y += a.d[l] * vy.y;
        zz.x += a.c[l] * vz.x;
        zz.y += a.c[l] * vz.y;
        zlz.x += a.d[l] * vz.x;
        zlz.y += a.d[l] * vz.y;
      }
      for (int l = 0; l < a.nn; ++l) {
        const double2 vy = a.n[l][cs + o], vz = a.n[l][2 * cs + o];
        ny.x += a.e[l] * vy.x;
        ny.y += a.e[l] * vy.y;
        nz.x += a.e[l] * vz.x;
        nz.y += a.e[l] * vz.y;
      }
      const double f = eta * k2 * hI;
      const double2 ly{hI * zy.x - f * zly.x + ny.x, hI * zy.y - f * zly.y + ny.y};
      const double2 lz{hI * zz.x - f * zlz.x + nz.x, hI * zz.y - f * zlz.y + nz.y};
      lp = double2{cy * ly.x + cz * lz.x, cy * ly.y + cz * lz.y};
      lq = double2{cz * ly.x - cy * lz.x, cz * ly.y - cy * lz.y};
      zp = double2{cy * zly.x + cz * zlz.x, cy * zly.y + cz * zlz.y};
      zq = double2{cz * zly.x - cy * zlz.x, cz * zly.y - cy * zlz.y};
    }
    Lx[((size_t)I * 2 + 0) * BS + lane] = lp;
    Lx[((size_t)I * 2 + 1) * BS + lane] = lq;
    Zl[((size_t)I * 2 + 0) * BS + lane] = zp;
    Zl[((size_t)I * 2 + 1) * BS + lane] = zq;
  }
  __syncthreads();
  // phase 1: thread = (lane, parity, P/Q)
  if (tid < 4 * BS) {
    const int lane = tid >> 2, par = (tid >> 1) & 1, pq = tid & 1, m = m0 + lane;
    if (m < M) {
      double cy, cz, k2, k;
      rot(m, cy, cz, k2, k);
      auto X = [&](int I) -> double2 & { return Lx[((size_t)I * 2 + pq) * BS + lane]; };
      auto Z = [&](int I) -> const double2 & { return Zl[((size_t)I * 2 + pq) * BS + lane]; };
      const int last = (n - 1 - par) / 2 * 2 + par;  // highest row of this parity: the wall row
      const double kb = pq ? k : 0.0;                // beta = i^2 (P) or i^2 + k (Q)
      double2 s3{0, 0}, s1{0, 0}, a3{0, 0}, a1{0, 0}, ba{0, 0};
      double b3 = 0, b1 = 0, bb = 0;
      for (int I = last; I >= 0; I -= 2) {
        const double hI = (I == 0) ? M_PI : 0.5 * M_PI, I2 = (double)I * I, I3 = I2 * I;
        const double2 zl = Z(I);
        double2 av;
        double bv;
        if (I == last) {
          av = double2{0, 0};
          bv = 1.0;
        } else {
          const double2 lr = X(I);
          const double2 r{lr.x + eta * 0.5 * M_PI * (s3.x - I2 * s1.x), lr.y + eta * 0.5 * M_PI * (s3.y - I2 * s1.y)};
          const double rd = 1.0 / (hI * (1.0 + ge * k2));
          av = double2{(r.x + C * (a3.x - I2 * a1.x)) * rd, (r.y + C * (a3.y - I2 * a1.y)) * rd};
          bv = C * (b3 - I2 * b1) * rd;
        }
        s3.x += I3 * zl.x;
        s3.y += I3 * zl.y;
        s1.x += I * zl.x;
        s1.y += I * zl.y;
        a3.x += I3 * av.x;
        a3.y += I3 * av.y;
        a1.x += I * av.x;
        a1.y += I * av.y;
        b3 += I3 * bv;
        b1 += I * bv;
        const double beta = I2 + kb;
        ba.x += beta * av.x;
        ba.y += beta * av.y;
        bb += beta * bv;
        X(I) = av;
      }
      const double2 t{-ba.x / bb, -ba.y / bb};
      b3 = b1 = 0;
      for (int I = last; I >= 0; I -= 2) {
        const double hI = (I == 0) ? M_PI : 0.5 * M_PI, I2 = (double)I * I, I3 = I2 * I;
        const double bv = (I == last) ? 1.0 : C * (b3 - I2 * b1) / (hI * (1.0 + ge * k2));
        b3 += I3 * bv;
        b1 += I * bv;
        const double2 av = X(I);
        X(I) = double2{av.x + bv * t.x, av.y + bv * t.y};
      }
    }
  }
  __syncthreads();
  // phase 2
  double2 *oy = a.out + cs, *oz = a.out + 2 * cs;
  for (int idx = tid; idx < n * BS; idx += nthr) {
    const int I = idx / BS, lane = idx - I * BS, m = m0 + lane;
    if (m >= M) continue;
    double cy, cz, k2, k;
    rot(m, cy, cz, k2, k);
    const bool nyq = (m / g.kzs == g.ny / 2);
    const double2 p = Lx[((size_t)I * 2 + 0) * BS + lane], q = Lx[((size_t)I * 2 + 1) * BS + lane];
    const size_t o = (size_t)I * M + m;
    oy[o] = nyq ? double2{0, 0} : double2{cy * p.x + cz * q.x, cy * p.y + cz * q.y};
    oz[o] = nyq ? double2{0, 0} : double2{cz * p.x - cy * q.x, cz * p.y - cy * q.y};
  }
}

// -------------------------------------------------- k_reduce_blockmax ----
// out[f] = max over the nblocks rows of blockmax[b][f], one block.
__global__ void __launch_bounds__(256) k_reduce_blockmax(const double *__restrict__ bm, int nblocks, int nf,
                                                         double *__restrict__ out) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  for (int f = 0; f < nf; ++f) {
    double m = 0.0;
    for (int b = t; b < nblocks; b += blockDim.x) m = fmax(m, bm[(size_t)b * nf + f]);
    red[t] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (t < w) red[t] = fmax(red[t], red[t + w]);
      __syncthreads();
    }
    if (t == 0) out[f] = red[0];
    __syncthreads();
  }
}

// -------------------------------------------------------- k_expand_half ----
// The Hermitian half spectrum [rows][ny][kzs] (kzs = nz / 2) in the boundary's
// full layout [rows][ny][nz]: c(ky, kz) for kz < kzs, c(ky, -kz) =
// conj(c(-ky, kz)) above the spanwise Nyquist plane, which is zero.  One
// element per thread, grid-stride.
__global__ void __launch_bounds__(256) k_expand_half(const double2 *__restrict__ half, double2 *__restrict__ full,
                                                     int rows, int ny, int nz, int kzs) {
  const size_t n = (size_t)rows * ny * nz;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t ry = i / nz;
    const int kz = (int)(i - ry * nz);
    const size_t r = ry / ny;
    const int iy = (int)(ry - r * ny);
    double2 v{0.0, 0.0};
    if (kz < kzs) {
      v = half[ry * kzs + kz];
    } else if (kz > kzs) {
      const double2 m = half[(r * ny + (ny - iy) % ny) * kzs + (nz - kz)];
      v = double2{m.x, -m.y};
    }
    full[i] = v;
  }
}

// -------------------------------------------------------- k_bits_differ ----
// diff |= 1 when a and b differ in any bit (n complex values).
__global__ void __launch_bounds__(256) k_bits_differ(const double2 *__restrict__ a, const double2 *__restrict__ b,
                                                     size_t n, unsigned *__restrict__ diff) {
  const unsigned long long *x = reinterpret_cast<const unsigned long long *>(a);
  const unsigned long long *y = reinterpret_cast<const unsigned long long *>(b);
  bool d = false;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < 2 * n; i += (size_t)gridDim.x * blockDim.x)
    d |= x[i] != y[i];
  if (__any(d) && (threadIdx.x & 63) == 0) atomicOr(diff, 1u);  // one atomic per wave
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
  const int x = blockIdx.y;  // one node per block row: the matrix rows are uniform
  double e = 0.0, dv = 0.0, cross = 0.0, msh = 0.0, e0 = 0.0, e1 = 0.0;
  if (m < g.m) {
    double kyd, kzd, kyo;
    mode_k(g, m, kyd, kzd, kyo);
    const double wt = ((m % g.kzs) == 0) ? 1.0 : 2.0;
    {
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
      e0 += w * wt * (U0.x * U0.x + U0.y * U0.y);
      e1 += w * wt * (U1.x * U1.x + U1.y * U1.y);
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
  __shared__ double red[6][256];
  red[0][threadIdx.x] = e;
  red[1][threadIdx.x] = dv;
  red[2][threadIdx.x] = cross;
  red[3][threadIdx.x] = msh;
  red[4][threadIdx.x] = e0;
  red[5][threadIdx.x] = e1;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) atomicAdd(&out[q], red[q][0]);
}

// Magnetic diagnostics of the vector potential A (three components cstride
// apart): out[0] += sum wF |b|^2, out[1] += sum wF |div b|^2 over the full
// spectrum (Hermitian weights), out[2..4] += sum_x wF b_c(x) of the mean mode,
// with b = curl A at the nodes: DX = VDX Ax, DY/DZ = VDY A{y,z}, EY/EZ = VEY A{y,z},
// bx = i ky DZ - i kz DY, by = i kz DX - EZ, bz = EY - i ky DX, and
// d/dx bx = i ky EZ - i kz EY (examples/pcf_mhd_jax.py:760-789).
__global__ void __launch_bounds__(256) k_diag_mhd(Geo g, const double *__restrict__ VDX, int kx,
                                                  const double *__restrict__ VDY,
                                                  const double *__restrict__ VEY, int kyz,
                                                  const double *__restrict__ fejer,
                                                  const double2 *__restrict__ A, size_t cstride,
                                                  double *__restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = blockIdx.y;  // one node per block row
  double eb = 0.0, dv = 0.0, mb[3] = {0.0, 0.0, 0.0};
  if (m < g.m) {
    double kyd, kzd, kyo;
    mode_k(g, m, kyd, kzd, kyo);
    const double ky = kyo, kz = kzd;
    const double wt = ((m % g.kzs) == 0) ? 1.0 : 2.0;
    {
      double2 DX{0, 0}, DY{0, 0}, DZ{0, 0}, EY{0, 0}, EZ{0, 0};
      for (int i = 0; i < kx; ++i) {
        const double2 ax = A[(size_t)i * g.m + m];
        const double vd = VDX[(size_t)x * kx + i];
        DX.x += vd * ax.x;
        DX.y += vd * ax.y;
      }
      for (int i = 0; i < kyz; ++i) {
        const size_t o = (size_t)i * g.m + m;
        const double2 ay = A[cstride + o], az = A[2 * cstride + o];
        const double vd = VDY[(size_t)x * kyz + i], ve = VEY[(size_t)x * kyz + i];
        DY.x += vd * ay.x;
        DY.y += vd * ay.y;
        DZ.x += vd * az.x;
        DZ.y += vd * az.y;
        EY.x += ve * ay.x;
        EY.y += ve * ay.y;
        EZ.x += ve * az.x;
        EZ.y += ve * az.y;
      }
      // i z = (-z.y, z.x)
      const double2 bx{-(ky * DZ.y - kz * DY.y), ky * DZ.x - kz * DY.x};
      const double2 by{-kz * DX.y - EZ.x, kz * DX.x - EZ.y};
      const double2 bz{EY.x + ky * DX.y, EY.y - ky * DX.x};
      const double2 dbx{-(ky * EZ.y - kz * EY.y), ky * EZ.x - kz * EY.x};
      const double2 d{dbx.x - ky * by.y - kz * bz.y, dbx.y + ky * by.x + kz * bz.x};
      const double w = fejer[x] * wt;
      eb += w * (bx.x * bx.x + bx.y * bx.y + by.x * by.x + by.y * by.y + bz.x * bz.x + bz.y * bz.y);
      dv += w * (d.x * d.x + d.y * d.y);
      if (m == 0) {
        mb[0] += fejer[x] * bx.x;
        mb[1] += fejer[x] * by.x;
        mb[2] += fejer[x] * bz.x;
      }
    }
  }
  __shared__ double red[5][256];
  red[0][threadIdx.x] = eb;
  red[1][threadIdx.x] = dv;
  for (int q = 0; q < 3; ++q) red[2 + q][threadIdx.x] = mb[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 5; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 5; ++q) atomicAdd(&out[q], red[q][0]);
}

// Vacuum-matching residual of insulating walls (A_y, A_z in TC):
// b = curl A at x = -1, +1 from T_j(+-1) = (+-1)^j, T_j'(+-1) = (+-1)^{j+1} j^2
// (the TD component Ax vanishes there): bx = i ky Az - i kz Ay,
// bx' = d/dx of it, by = -Az', bz = Ay'; residuals bx' + s k bx,
// by + s (i ky / k) bx, bz + s (i kz / k) bx with s = -1, +1 at the lower and
// upper walls (insulating_bc_residual, examples/pcf_mhd_mri_shearpy_jax.py:802-830).
// One thread per mode; blockmax[block] = the block's largest |residual|.
__global__ void __launch_bounds__(256) k_insul_residual(Geo g, const double2 *__restrict__ A, size_t cstride,
                                                        double *__restrict__ blockmax) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  double r = 0.0;
  if (m < g.m && m / g.kzs != g.ny / 2) {
    double kyd, kzd, kyo;
    mode_k(g, m, kyd, kzd, kyo);
    const double ky = kyo, kz = kzd, k = sqrt(kyd * kyd + kzd * kzd);
    const double kn = (k > 0.0) ? k : 1.0;
    // wall sums: v[w] = sum T_j(w) c_j, d[w] = sum T_j'(w) c_j for c = Ay, Az
    double2 vy[2] = {{0, 0}, {0, 0}}, vz[2] = {{0, 0}, {0, 0}}, dy[2] = {{0, 0}, {0, 0}}, dz[2] = {{0, 0}, {0, 0}};
    for (int j = 0; j < g.n0; ++j) {
      const size_t o = (size_t)j * g.m + m;
      const double2 ay = A[cstride + o], az = A[2 * cstride + o];
      const double sg = (j & 1) ? -1.0 : 1.0, jj = (double)j * j;
      const double t0[2] = {sg, 1.0}, t1[2] = {-sg * jj, jj};
      for (int w = 0; w < 2; ++w) {
        vy[w].x += t0[w] * ay.x;
        vy[w].y += t0[w] * ay.y;
        vz[w].x += t0[w] * az.x;
        vz[w].y += t0[w] * az.y;
        dy[w].x += t1[w] * ay.x;
        dy[w].y += t1[w] * ay.y;
        dz[w].x += t1[w] * az.x;
        dz[w].y += t1[w] * az.y;
      }
    }
    for (int w = 0; w < 2; ++w) {
      const double s = w ? 1.0 : -1.0;
      const double2 bx{-(ky * vz[w].y - kz * vy[w].y), ky * vz[w].x - kz * vy[w].x};
      const double2 bxp{-(ky * dz[w].y - kz * dy[w].y), ky * dz[w].x - kz * dy[w].x};
      const double2 by{-dz[w].x, -dz[w].y}, bz = dy[w];
      const double2 rx{bxp.x + s * k * bx.x, bxp.y + s * k * bx.y};
      // s (i ky / k) bx = s ky / k (-bx.y, bx.x)
      const double2 ry{by.x - s * ky / kn * bx.y, by.y + s * ky / kn * bx.x};
      const double2 rz{bz.x - s * kz / kn * bx.y, bz.y + s * kz / kn * bx.x};
      r = fmax(r, fmax(hypot(rx.x, rx.y), fmax(hypot(ry.x, ry.y), hypot(rz.x, rz.y))));
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = r;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) blockmax[blockIdx.x] = red[0];
}

// ------------------------------------------------------------ launchers ----
namespace {

// spanwise columns per block of the streamwise FFT: 8 (measured at 65x64x64:
// 9.6 / 5.6 us FP64 inverse / forward, against 10.2 / 6.9 with 16 and 12.8 /
// 8.6 with 4, which has no compile-time variant)
int pick_kc(int kzs) {
  int kc = 8;
  while (kc > 1 && kzs % kc) kc >>= 1;
  return kc;
}
}  // namespace

// Butterflies per thread of the largest pass of `plan` over `nbatch` sequences.
int max_butterflies(const FFTPlan &plan, int nbatch) {
  int mb = 0;
  for (int s = 0; s < plan.nst; ++s) mb = std::max(mb, (plan.n / plan.radix[s] * nbatch + 255) / 256);
  return mb;
}

// The same length without radix-8 passes (the FP64 k_rotational's
// compile-time plans); launch bounds cover both plans.
FFTPlan radix4_plan(const FFTPlan &p) {
  const CPlan c = cplan(p.n, false);
  FFTPlan q{};
  q.n = p.n;
  q.nst = c.nst;
  for (int s = 0; s < c.nst; ++s) q.radix[s] = c.r[s];
  return q;
}

int rot_butterflies(const FFTPlan &pz, int nbatch) {
  return std::max(max_butterflies(pz, nbatch), max_butterflies(radix4_plan(pz), nbatch));
}

template <typename T>
hipError_t launch_fft_y_fields(const Geo &g, const NodalMats<T> &nm, const FFTPlan &py, bool inverse,
                               int nf, const void *src_, void *dst_, hipStream_t s) {
  using V = typename VT<T>::V;
  const int kc = pick_kc(g.kzs);
  const size_t lds = ((size_t)g.my * kc + g.my) * sizeof(V);
  const int mb = max_butterflies(py, kc);
  const dim3 grid(g.kzs / kc, g.n0, nf);
  const V *tw = reinterpret_cast<const V *>(nm.twy);
  const V *src = reinterpret_cast<const V *>(src_);
  V *dst = reinterpret_cast<V *>(dst_);
#define PCF_FFTYN(INV, MB, NY, KC) \
  hipLaunchKernelGGL((k_fft_y<T, INV, MB, NY, KC>), grid, dim3(256), lds, s, g, tw, py, kc, src, dst)
#define PCF_FFTYK(INV, MB, NY)                       \
  if (kc == 16) PCF_FFTYN(INV, MB, NY, 16);          \
  else if (kc == 8) PCF_FFTYN(INV, MB, NY, 8);       \
  else PCF_FFTYN(INV, MB, NY, 0);
#define PCF_FFTY(INV, MB)                            \
  switch (g.my) {                                    \
    case 24: { PCF_FFTYK(INV, MB, 24) } break;       \
    case 48: { PCF_FFTYK(INV, MB, 48) } break;       \
    case 96: { PCF_FFTYK(INV, MB, 96) } break;       \
    case 144: { PCF_FFTYK(INV, MB, 144) } break;     \
    case 192: { PCF_FFTYK(INV, MB, 192) } break;     \
    default: PCF_FFTYN(INV, MB, 0, 0);               \
  }
  if (inverse) {
    if (mb <= 2) PCF_FFTY(true, 2)
    else if (mb <= 4) PCF_FFTY(true, 4)
    else PCF_FFTY(true, 8)
  } else {
    if (mb <= 2) PCF_FFTY(false, 2)
    else if (mb <= 4) PCF_FFTY(false, 4)
    else PCF_FFTY(false, 8)
  }
#undef PCF_FFTY
#undef PCF_FFTYK
#undef PCF_FFTYN
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_y(const Geo &g, const NodalMats<T> &nm, Workspace &ws, const FFTPlan &py,
                        bool inverse, bool mhd, hipStream_t s) {
  return inverse ? launch_fft_y_fields<T>(g, nm, py, true, mhd ? 12 : 6, ws.X, ws.Y, s)
                 : launch_fft_y_fields<T>(g, nm, py, false, mhd ? 6 : 3, ws.Hz, ws.Hx, s);
}

// y-line pairs per block; LDS holds (inputs + outputs) lines of MZ per unit
// plus the twiddle table.  FP64 and MHD: one pair (twice the blocks, half the
// LDS each: measured 20.6 instead of 23.3 us at 65x64x64 in FP64); FP32: two
// (14.3 instead of 15.5 us)
template <typename T>
int rot_units(const Geo &g, bool mhd) {
  using V = typename VT<T>::V;
  const int lines = mhd ? 18 : 9;
  int upb = (mhd || sizeof(T) == 8) ? 1 : 2;
  while (upb > 1 && ((size_t)upb * lines * g.mz + g.mz) * sizeof(V) > 65536) upb >>= 1;
  return upb;
}

template <typename T>
hipError_t launch_rotational(const Geo &g, const NodalMats<T> &nm, Workspace &ws,
                             const FFTPlan &pz, bool mhd, hipStream_t s) {
  using V = typename VT<T>::V;
  const int nunits = g.n0 * (g.my / 2);
  if (!mhd && g.mz == 3 * g.kzs && g.my % 2 == 0) {
    // spanwise FFTs in registers when nz / 2 is a power of two up to 64
    const V *Yw = reinterpret_cast<const V *>(ws.Y);
    V *Hw = reinterpret_cast<V *>(ws.Hz);
    auto grid_for = [&](int kz) { return dim3((unsigned)((nunits + 2 * (64 / kz) - 1) / (2 * (64 / kz)))); };
    switch (g.kzs) {
      case 8: hipLaunchKernelGGL((k_rot_wave<T, 8>), grid_for(8), dim3(128), 0, s, g, nm, Yw, Hw); return hipGetLastError();
      case 16: hipLaunchKernelGGL((k_rot_wave<T, 16>), grid_for(16), dim3(128), 0, s, g, nm, Yw, Hw); return hipGetLastError();
      case 32: hipLaunchKernelGGL((k_rot_wave<T, 32>), grid_for(32), dim3(128), 0, s, g, nm, Yw, Hw); return hipGetLastError();
      case 64: hipLaunchKernelGGL((k_rot_wave<T, 64>), grid_for(64), dim3(128), 0, s, g, nm, Yw, Hw); return hipGetLastError();
      default: break;
    }
  }
  const int upb = rot_units<T>(g, mhd);
  const size_t lds = ((size_t)upb * (mhd ? 18 : 9) * g.mz + g.mz) * sizeof(V);
  const int mb = rot_butterflies(pz, upb * (mhd ? 12 : 6));
  const dim3 grid((nunits + upb - 1) / upb);
  const V *Y = reinterpret_cast<const V *>(ws.Y);
  V *Hz = reinterpret_cast<V *>(ws.Hz);
#define PCF_ROTN(MB, NZ)                                                                                \
  if (mhd)                                                                                              \
    hipLaunchKernelGGL((k_rotational_mhd<T, MB, NZ>), grid, dim3(256), lds, s, g, nm, pz, upb, Y, Hz); \
  else                                                                                                  \
    hipLaunchKernelGGL((k_rotational<T, MB, NZ>), grid, dim3(256), lds, s, g, nm, pz, upb, Y, Hz)
#define PCF_ROT(MB)                              \
  switch (g.mz) {                                \
    case 24: PCF_ROTN(MB, 24); break;            \
    case 48: PCF_ROTN(MB, 48); break;            \
    case 96: PCF_ROTN(MB, 96); break;            \
    case 144: PCF_ROTN(MB, 144); break;          \
    case 192: PCF_ROTN(MB, 192); break;          \
    default: PCF_ROTN(MB, 0);                    \
  }
  if (mb <= 3) PCF_ROT(3)
  else if (mb <= 6) PCF_ROT(6)
  else PCF_ROT(12)
#undef PCF_ROT
#undef PCF_ROTN
  return hipGetLastError();
}

// Line pairs per block of k_absmax_z and its launch checks.
static int absmax_units(const Geo &g, int nf, int elem_bytes) {
  int upb = 4;
  while (upb > 1 && ((size_t)upb * nf * g.mz + g.mz) * elem_bytes + 4 * 6 * sizeof(double) > 65536) upb >>= 1;
  return upb;
}

int absmax_blocks(const Geo &g, int nf, int elem_bytes) {
  const int upb = absmax_units(g, nf, elem_bytes);
  return (g.n0 * (g.my / 2) + upb - 1) / upb;
}

bool absmax_fits(const Geo &g, const FFTPlan &pz, int nf, int elem_bytes) {
  const int upb = absmax_units(g, nf, elem_bytes);
  if (g.my % 2 || g.mz < 2 * g.kzs || nf < 1 || nf > 6) return false;
  if (((size_t)upb * nf * g.mz + g.mz) * elem_bytes + 4 * 6 * sizeof(double) > 65536) return false;
  return max_butterflies(pz, upb * nf) <= 12;
}

template <typename T>
hipError_t launch_absmax(const Geo &g, const T *xnode, const T *twz, const FFTPlan &pz, int nf, const void *Y,
                         double *blockmax, hipStream_t s, double *phys) {
  using V = typename VT<T>::V;
  if (!absmax_fits(g, pz, nf, sizeof(V))) return hipErrorInvalidValue;
  const int upb = absmax_units(g, nf, sizeof(V));
  const size_t lds = ((size_t)upb * nf * g.mz + g.mz) * sizeof(V) + 4 * 6 * sizeof(double);
  const int mb = max_butterflies(pz, upb * nf);
  const dim3 grid(absmax_blocks(g, nf, sizeof(V)));
  const V *tw = reinterpret_cast<const V *>(twz);
  const V *y = reinterpret_cast<const V *>(Y);
#define PCF_ABSMAX(MB) \
  hipLaunchKernelGGL((k_absmax_z<T, MB>), grid, dim3(256), lds, s, g, tw, pz, upb, nf, xnode, y, blockmax, phys)
  if (mb <= 3) PCF_ABSMAX(3);
  else if (mb <= 6) PCF_ABSMAX(6);
  else PCF_ABSMAX(12);
#undef PCF_ABSMAX
  return hipGetLastError();
}

size_t solve_scratch_doubles(const Geo &g) {
  return (size_t)4 * g.npmax * 7 * g.m + (size_t)2 * g.npmax * 7;
}

hipError_t launch_solve_bordered(const Geo &g, const BorderedArgs &a, hipStream_t s) {
  // modes per block: the rows of both LDS arrays within the default 64 KB
  const size_t per_mode = (size_t)g.n0 * 2 * 2 * sizeof(double2);
  if (8 * per_mode <= 65536)
    hipLaunchKernelGGL((k_solve_bordered<8>), dim3((g.m + 7) / 8), dim3(128), 8 * per_mode, s, g, a);
  else if (4 * per_mode <= 65536)
    hipLaunchKernelGGL((k_solve_bordered<4>), dim3((g.m + 3) / 4), dim3(128), 4 * per_mode, s, g, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_reduce_blockmax(const double *bm, int nblocks, int nf, double *out, hipStream_t s) {
  if (nblocks < 1 || nf < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_reduce_blockmax, dim3(1), dim3(256), 0, s, bm, nblocks, nf, out);
  return hipGetLastError();
}

hipError_t launch_expand_half(const double2 *half, double2 *full, int rows, int ny, int nz, int kzs, hipStream_t s) {
  const size_t n = (size_t)rows * ny * nz;
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>(8192, (n + 255) / 256);
  hipLaunchKernelGGL(k_expand_half, dim3((unsigned)blocks), dim3(256), 0, s, half, full, rows, ny, nz, kzs);
  return hipGetLastError();
}

hipError_t launch_bits_differ(const double2 *a, const double2 *b, size_t n, unsigned *diff, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>(1024, (2 * n + 255) / 256);
  hipLaunchKernelGGL(k_bits_differ, dim3((unsigned)blocks), dim3(256), 0, s, a, b, n, diff);
  return hipGetLastError();
}

hipError_t launch_diag(const Geo &g, const double *VB, const double *VD, const double *VB1,
                       const double *VD1, const double *fejer, const double *xnode,
                       const double2 *u0, const double2 *u1, const double2 *u2, double *out,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_diag, dim3((g.m + 255) / 256, g.n0), dim3(256), 0, s, g, VB, VD, VB1, VD1, fejer,
                     xnode, u0, u1, u2, out);
  return hipGetLastError();
}

hipError_t launch_diag_mhd(const Geo &g, const double *VDX, int kx, const double *VDY, const double *VEY, int kyz,
                           const double *fejer, const double2 *A, size_t cstride, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_mhd, dim3((g.m + 255) / 256, g.n0), dim3(256), 0, s, g, VDX, kx, VDY, VEY, kyz, fejer, A,
                     cstride, out);
  return hipGetLastError();
}

int insul_residual_blocks(const Geo &g) { return (g.m + 255) / 256; }

hipError_t launch_insul_residual(const Geo &g, const double2 *A, size_t cstride, double *blockmax, hipStream_t s) {
  hipLaunchKernelGGL(k_insul_residual, dim3(insul_residual_blocks(g)), dim3(256), 0, s, g, A, cstride, blockmax);
  return hipGetLastError();
}

bool fft_fits(const Geo &g, const FFTPlan &py, const FFTPlan &pz, int elem_bytes, bool mhd) {
  const int kc = pick_kc(g.kzs);
  if (((size_t)g.my * kc + g.my) * elem_bytes > 65536) return false;
  if (max_butterflies(py, kc) > 8) return false;
  const int upb = (elem_bytes == 16) ? rot_units<double>(g, mhd) : rot_units<float>(g, mhd);
  if (((size_t)upb * (mhd ? 18 : 9) * g.mz + g.mz) * elem_bytes > 65536) return false;
  return rot_butterflies(pz, upb * (mhd ? 12 : 6)) <= 12;
}

#define PCF_INSTANTIATE(T)                                                                     \
  template hipError_t launch_absmax<T>(const Geo &, const T *, const T *, const FFTPlan &, int,   \
                                       const void *, double *, hipStream_t, double *);           \
  template hipError_t launch_fft_y<T>(const Geo &, const NodalMats<T> &, Workspace &,            \
                                      const FFTPlan &, bool, bool, hipStream_t);                  \
  template hipError_t launch_fft_y_fields<T>(const Geo &, const NodalMats<T> &, const FFTPlan &,   \
                                             bool, int, const void *, void *, hipStream_t);       \
  template hipError_t launch_rotational<T>(const Geo &, const NodalMats<T> &, Workspace &,       \
                                           const FFTPlan &, bool, hipStream_t);                  \

PCF_INSTANTIATE(double)
PCF_INSTANTIATE(float)
#undef PCF_INSTANTIATE

}  // namespace pcf
