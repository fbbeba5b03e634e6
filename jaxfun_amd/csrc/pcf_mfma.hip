// Wall-normal contractions of the step and of its adjoint on the matrix cores.
//
// Every one of them is a small real GEMM per field, Out[r][c] = sum_k A[r][k]
// B[k][c]: A a (<= 65 x 65) basis/projection matrix, B the spectral or nodal
// field with its complex modes laid out as pairs of real columns (re, im).  The
// per-mode wavenumber factors (i ky, i kz, k^2) are applied to B as it is read
// and to the accumulators in the epilogue, where a lane takes its column's
// partner component (re <-> im) from the neighbouring lane.
//
// One wave computes one 16 x 16 tile (16 output rows x 8 modes) with
// v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (mixed precision).
// Operand layout of both: A[i][k] in lane 16k + i, B[k][j] in lane 16k + j.
// Result: FP64 lane l, register r holds D[4r + l/16][l%16]; FP32 holds
// D[4(l/16) + r][l%16].
#include <algorithm>

#include "pcf_adjoint.hpp"
#include "pcf_device.hpp"

namespace pcf {
namespace {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc = d4v;
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * r + (l >> 4); }
};
template <>
struct Mfma<float> {
  using acc = f4v;
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

template <typename T>
__device__ __forceinline__ T partner(T v) {
  return __shfl_xor(v, 1, 64);
}

template <typename T>
__device__ __forceinline__ T comp_of(const double2 &z, int comp) {
  return (T)(comp ? z.y : z.x);
}

// Lane geometry of one tile: real column c of the tile, its mode and component.
struct Lane {
  int l, c, mcol, comp;
  bool valid;  // mode inside the spectrum
};
__device__ __forceinline__ Lane lane_of(const Geo &g, int ct) {
  Lane L;
  L.l = threadIdx.x & 63;
  L.c = L.l & 15;
  L.mcol = 8 * ct + (L.c >> 1);
  L.comp = L.c & 1;
  L.valid = L.mcol < g.m;
  return L;
}

__device__ __forceinline__ bool nyq_mode(const Geo &g, int m) { return m / g.kzs == g.ny / 2; }

// ------------------------------------------------------------ staging ----
// A block owns one tile of 8 modes (16 real columns) and one wave per 16-row
// output tile.  Its B panels (NF fields x Kp rows x 16 columns, in T) are first
// staged into LDS with independent, coalesced loads by all threads, so the
// MFMA loop only waits on LDS and on the small (L2-resident) A matrices.
// Staging loads per thread of each kernel; the launchers check that
// blockDim * STAGE covers the kernel's panel.
constexpr int STAGE_NODAL = 6, STAGE_NODAL_PRIM = 4, STAGE_RHS = 3, STAGE_NODAL_ADJ = 3, STAGE_RHS_ADJ = 4,
              STAGE_RECON = 4;

// K loop of a wave's tile: nit steps of 4 K rows.  The A operands (NA values
// per step, from small L2-resident matrices) are read D steps ahead into a
// register ring, so a step waits on loads issued D steps earlier instead of
// one L2 round trip per step.  load(it, a) must be branch-free (clamped
// address, then select), which lets the ring unroll.
template <typename T, int NA, int D, class Load, class Step>
__device__ __forceinline__ void k_loop(int nit, Load load, Step step) {
  T ring[D][NA];
  // every load is issued unconditionally (step index clamped to nit - 1), so
  // the ring carries no branches and the waits stay partial
#pragma unroll
  for (int c = 0; c < D; ++c) load(c < nit ? c : nit - 1, ring[c]);
  int it0 = 0;
  for (; it0 + D <= nit; it0 += D) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      step(it0 + c, ring[c]);
      const int nx = it0 + c + D;
      load(nx < nit ? nx : nit - 1, ring[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < D; ++c)
    if (it0 + c < nit) step(it0 + c, ring[c]);
}

// Element k of column xa of a [K][ld] matrix.  The host pads every matrix with
// zero rows up to a multiple of 4 that covers n0, so k < kp needs no check.
template <typename T>
__device__ __forceinline__ T mat_at(const T *m, int ld, int k, int xa) {
  return m[(size_t)k * ld + xa];
}

template <typename T>
struct Panel {
  T *b;
  int kp;
  __device__ T &at(int f, int k, int c) const { return b[((size_t)f * kp + k) * 16 + c]; }
};

// ------------------------------------------------------------ forward ----
// U0 = VB u0, U1 = VD u1, U2 = VD u2, D1 = VD1 u1, D2 = VD1 u2 (rows: nodes x),
// then w0 = i ky U2 - i kz U1, w1 = i kz U0 - D2, w2 = D1 - i ky U0.
template <typename T>
__global__ void __launch_bounds__(512) k_nodal_mfma(Geo g, NodalMats<T> nm, const double2 *__restrict__ u0,
                                                    const double2 *__restrict__ u1,
                                                    const double2 *__restrict__ u2,
                                                    typename VT<T>::V *__restrict__ X) {
  using MM = Mfma<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ct = blockIdx.x;
  const int M = g.m, kp = (g.nd + 3) / 4 * 4;
  const Panel<T> B{reinterpret_cast<T *>(smem), kp};
  {
    // every load of the thread in flight before the first LDS store
    const int total = 3 * kp * 8;
    double2 z[STAGE_NODAL];
#pragma unroll
    for (int t = 0; t < STAGE_NODAL; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      const int f = idx / (kp * 8), rem = idx - f * kp * 8, k = rem >> 3, j = rem & 7;
      const int m = 8 * ct + j;
      z[t] = double2{0, 0};
      if (idx < total && m < M) {
        if (f == 0 && k < g.nb) z[t] = u0[(size_t)k * M + m];
        if (f == 1 && k < g.nd) z[t] = u1[(size_t)k * M + m];
        if (f == 2 && k < g.nd) z[t] = u2[(size_t)k * M + m];
      }
    }
#pragma unroll
    for (int t = 0; t < STAGE_NODAL; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      if (idx >= total) continue;
      const int f = idx / (kp * 8), rem = idx - f * kp * 8, k = rem >> 3, j = rem & 7;
      B.at(f, k, 2 * j) = (T)z[t].x;
      B.at(f, k, 2 * j + 1) = (T)z[t].y;
    }
  }
  __syncthreads();
  const Lane L = lane_of(g, ct);
  double kyd, kzd, kyo;
  mode_k(g, L.valid ? L.mcol : 0, kyd, kzd, kyo);
  const T ky = (T)kyo, kz = (T)kzd;
  const T s = L.comp ? (T)1 : (T)-1;  // re lane: i z -> -z.im ; im lane: i z -> +z.re
  const size_t plane = (size_t)g.n0 * 2 * M;
  T *Xr = reinterpret_cast<T *>(X);
  for (int x0 = (threadIdx.x >> 6) * 16; x0 < g.n0; x0 += (blockDim.x >> 6) * 16) {
  const int xa = x0 + (L.l & 15);
  typename MM::acc U0{}, U1{}, U2{}, D1{}, D2{};
  const int kq = L.l >> 4;
  k_loop<T, 3, 8>(
      kp / 4,
      [&](int it, T(&a)[3]) {
        const int k = 4 * it + kq;
        a[0] = mat_at(nm.VBt, nm.ldx, k, xa);
        a[1] = mat_at(nm.VDt, nm.ldx, k, xa);
        a[2] = mat_at(nm.VD1t, nm.ldx, k, xa);
      },
      [&](int it, const T(&a)[3]) {
        const int k = 4 * it + kq;
        const T b0 = B.at(0, k, L.c), b1 = B.at(1, k, L.c), b2 = B.at(2, k, L.c);
        U0 = MM::mma(a[0], b0, U0);
        U1 = MM::mma(a[1], b1, U1);
        U2 = MM::mma(a[1], b2, U2);
        D1 = MM::mma(a[2], b1, D1);
        D2 = MM::mma(a[2], b2, D2);
      });
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const T u0v = U0[r], u1v = U1[r], u2v = U2[r], d1 = D1[r], d2 = D2[r];
    const T p0 = partner(u0v), p1 = partner(u1v), p2 = partner(u2v);
    const T w0 = s * (ky * p2 - kz * p1);
    const T w1 = s * kz * p0 - d2;
    const T w2 = d1 - s * ky * p0;
    const int x = x0 + MM::row(L.l, r);
    if (L.valid && x < g.n0) {
      const size_t o = (size_t)x * 2 * M + 2 * L.mcol + L.comp;
      Xr[o] = u0v;
      Xr[plane + o] = u1v;
      Xr[2 * plane + o] = u2v;
      Xr[3 * plane + o] = w0;
      Xr[4 * plane + o] = w1;
      Xr[5 * plane + o] = w2;
    }
  }
  }
}

// The same nodal fields from the solver's own unknowns (u0, g, v00, w00), with
// the velocity reconstruction (channelflow_kmm.py:855-870)
//   u1 = i (ky f + kz g) / k^2,  u2 = i (kz f - ky g) / k^2,  f = R u0,
//   u1(0, 0) = v00, u2(0, 0) = w00
// folded into the contraction: P1 = (VD R) u0, P2 = (VD1 R) u0, P3 = VD g',
// P4 = VD1 g', where g' is g with the mean mode's column replaced by
// v00 + i w00 (g(0, 0) = 0).  Then U1 = i (ky P1 + kz P3) / k^2,
// U2 = i (kz P1 - ky P3) / k^2, D1 and D2 likewise from P2, P4, and
// w0 = i ky U2 - i kz U1 = P3 (zero at the mean mode), w1 = i kz U0 - D2,
// w2 = D1 - i ky U0.  Mean mode: U1 = Re P3, U2 = Im P3, D1 = Re P4, D2 = Im P4
// on the real lane, zero on the imaginary one.
template <typename T>
__global__ void __launch_bounds__(512) k_nodal_prim_mfma(Geo g, NodalMats<T> nm, const double2 *__restrict__ u0,
                                                         const double2 *__restrict__ gf,
                                                         const double *__restrict__ v00,
                                                         const double *__restrict__ w00,
                                                         typename VT<T>::V *__restrict__ X) {
  using MM = Mfma<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ct = blockIdx.x;
  const int M = g.m, kp = (g.nd + 3) / 4 * 4;
  const Panel<T> B{reinterpret_cast<T *>(smem), kp};
  {
    const int total = 2 * kp * 8;
    double2 z[STAGE_NODAL_PRIM];
#pragma unroll
    for (int t = 0; t < STAGE_NODAL_PRIM; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      const int f = idx / (kp * 8), rem = idx - f * kp * 8, k = rem >> 3, j = rem & 7;
      const int m = 8 * ct + j;
      z[t] = double2{0, 0};
      if (idx < total && m < M) {
        if (f == 0 && k < g.nb) z[t] = u0[(size_t)k * M + m];
        if (f == 1 && k < g.nd) z[t] = (m == 0) ? double2{v00[k], w00[k]} : gf[(size_t)k * M + m];
      }
    }
#pragma unroll
    for (int t = 0; t < STAGE_NODAL_PRIM; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      if (idx >= total) continue;
      const int f = idx / (kp * 8), rem = idx - f * kp * 8, k = rem >> 3, j = rem & 7;
      B.at(f, k, 2 * j) = (T)z[t].x;
      B.at(f, k, 2 * j + 1) = (T)z[t].y;
    }
  }
  __syncthreads();
  const Lane L = lane_of(g, ct);
  double kyd, kzd, kyo;
  mode_k(g, L.valid ? L.mcol : 0, kyd, kzd, kyo);
  const double k2d = kyd * kyd + kzd * kzd;
  const bool mean = L.valid && L.mcol == 0;
  // reconstruction factors ky/k^2, kz/k^2 (K_over_K2, tensorproductspace.py:80-99)
  const T ay = (T)(k2d > 0 ? kyd / k2d : 0.0), az = (T)(k2d > 0 ? kzd / k2d : 0.0);
  const T ky = (T)kyo, kz = (T)kzd;
  const T s = L.comp ? (T)1 : (T)-1;  // re lane: i z -> -z.im ; im lane: i z -> +z.re
  const size_t plane = (size_t)g.n0 * 2 * M;
  T *Xr = reinterpret_cast<T *>(X);
  for (int x0 = (threadIdx.x >> 6) * 16; x0 < g.n0; x0 += (blockDim.x >> 6) * 16) {
    const int xa = x0 + (L.l & 15);
    typename MM::acc U0{}, P1{}, P2{}, P3{}, P4{};
    const int kq = L.l >> 4;
    k_loop<T, 5, 5>(
        kp / 4,
        [&](int it, T(&a)[5]) {
          const int k = 4 * it + kq;
          a[0] = mat_at(nm.VBt, nm.ldx, k, xa);
          a[1] = mat_at(nm.VDRt, nm.ldx, k, xa);
          a[2] = mat_at(nm.VD1Rt, nm.ldx, k, xa);
          a[3] = mat_at(nm.VDt, nm.ldx, k, xa);
          a[4] = mat_at(nm.VD1t, nm.ldx, k, xa);
        },
        [&](int it, const T(&a)[5]) {
          const int k = 4 * it + kq;
          const T b0 = B.at(0, k, L.c), b1 = B.at(1, k, L.c);
          U0 = MM::mma(a[0], b0, U0);
          P1 = MM::mma(a[1], b0, P1);
          P2 = MM::mma(a[2], b0, P2);
          P3 = MM::mma(a[3], b1, P3);
          P4 = MM::mma(a[4], b1, P4);
        });
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const T u0v = U0[r], q1 = P1[r], q2 = P2[r], q3 = P3[r], q4 = P4[r];
      const T p0 = partner(u0v), p1 = partner(q1), p2 = partner(q2), p3 = partner(q3), p4 = partner(q4);
      T u1v = s * (ay * p1 + az * p3), u2v = s * (az * p1 - ay * p3);
      T d1 = s * (ay * p2 + az * p4), d2 = s * (az * p2 - ay * p4);
      T w0 = q3;
      if (mean) {
        u1v = L.comp ? (T)0 : q3;
        u2v = L.comp ? (T)0 : p3;
        d1 = L.comp ? (T)0 : q4;
        d2 = L.comp ? (T)0 : p4;
        w0 = (T)0;
      }
      const T w1 = s * kz * p0 - d2;
      const T w2 = d1 - s * ky * p0;
      const int x = x0 + MM::row(L.l, r);
      if (L.valid && x < g.n0) {
        const size_t o = (size_t)x * 2 * M + 2 * L.mcol + L.comp;
        Xr[o] = u0v;
        Xr[plane + o] = u1v;
        Xr[2 * plane + o] = u2v;
        Xr[3 * plane + o] = w0;
        Xr[4 * plane + o] = w1;
        Xr[5 * plane + o] = w2;
      }
    }
  }
}

// Magnetic field and current of the vector potential A at the nodes, with
// DX = mDX Ax, QX = mQX Ax, DY = mDY Ay, DZ = mDY Az, EY = mEY Ay, EZ = mEY Az,
// RY = mRY Ay, RZ = mRY Az (MhdMats):
//   bx = i ky DZ - i kz DY,  by = i kz DX - EZ,  bz = EY - i ky DX,
//   jx = k^2 DX + i ky EY + i kz EZ,
//   jy = kz^2 DY - ky kz DZ - RY + i ky QX,
//   jz = ky^2 DZ - ky kz DY - RZ + i kz QX.
// Conducting walls (A in TD^3): mQX = VD P_D D S_D^T and mRY = VD P_D D^2 S_D^T
// carry the Galerkin projections of d/dx B onto TD that J = curl B needs
// (update_B_from_A / update_J_from_B, examples/pcf_mhd_jax.py:269-311).
// Insulating walls (Ay, Az in TC, B and J in TC^3): mQX = VD1, mDY/mEY/mRY the
// Chebyshev values and derivatives (examples/pcf_mhd_mri_shearpy_jax.py:548-557).
// Fields 6..11 of X receive (bx, by, bz, jx, jy, jz).
template <typename T>
__global__ void __launch_bounds__(512) k_nodal_mhd_mfma(Geo g, MhdMats<T> mm, const double2 *__restrict__ A,
                                                        size_t cstride, typename VT<T>::V *__restrict__ X) {
  using MM = Mfma<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ct = blockIdx.x;
  const int M = g.m, kp = (mm.kyz + 3) / 4 * 4;
  const Panel<T> B{reinterpret_cast<T *>(smem), kp};
  {
    const int total = 3 * kp * 8;
    double2 z[STAGE_NODAL];
#pragma unroll
    for (int t = 0; t < STAGE_NODAL; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      const int f = idx / (kp * 8), rem = idx - f * kp * 8, k = rem >> 3, j = rem & 7;
      const int m = 8 * ct + j;
      z[t] = double2{0, 0};
      if (idx < total && m < M && k < (f ? mm.kyz : mm.kx)) z[t] = A[f * cstride + (size_t)k * M + m];
    }
#pragma unroll
    for (int t = 0; t < STAGE_NODAL; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      if (idx >= total) continue;
      const int f = idx / (kp * 8), rem = idx - f * kp * 8, k = rem >> 3, j = rem & 7;
      B.at(f, k, 2 * j) = (T)z[t].x;
      B.at(f, k, 2 * j + 1) = (T)z[t].y;
    }
  }
  __syncthreads();
  const Lane L = lane_of(g, ct);
  double kyd, kzd, kyo;
  mode_k(g, L.valid ? L.mcol : 0, kyd, kzd, kyo);
  const T ky = (T)kyo, kz = (T)kzd, k2 = (T)(kyo * kyo + kzd * kzd);
  const T kyy = ky * ky, kzz = kz * kz, kyz = ky * kz;
  const T s = L.comp ? (T)1 : (T)-1;
  const size_t plane = (size_t)g.n0 * 2 * M;
  T *Xr = reinterpret_cast<T *>(X) + 6 * plane;
  for (int x0 = (threadIdx.x >> 6) * 16; x0 < g.n0; x0 += (blockDim.x >> 6) * 16) {
    const int xa = x0 + (L.l & 15);
    typename MM::acc DX{}, DY{}, DZ{}, EY{}, EZ{}, QX{}, RY{}, RZ{};
    const int kq = L.l >> 4;
    // rows k >= kx of mDX, mQX and k >= kyz of the others are zero padding
    k_loop<T, 5, 6>(
        kp / 4,
        [&](int it, T(&a)[5]) {
          const size_t o = (size_t)(4 * it + kq) * mm.ldx + xa;
          a[0] = mm.mDX[o];
          a[1] = mm.mQX[o];
          a[2] = mm.mDY[o];
          a[3] = mm.mEY[o];
          a[4] = mm.mRY[o];
        },
        [&](int it, const T(&a)[5]) {
          const int k = 4 * it + kq;
          const T bx = B.at(0, k, L.c), by = B.at(1, k, L.c), bz = B.at(2, k, L.c);
          DX = MM::mma(a[0], bx, DX);
          QX = MM::mma(a[1], bx, QX);
          DY = MM::mma(a[2], by, DY);
          DZ = MM::mma(a[2], bz, DZ);
          EY = MM::mma(a[3], by, EY);
          EZ = MM::mma(a[3], bz, EZ);
          RY = MM::mma(a[4], by, RY);
          RZ = MM::mma(a[4], bz, RZ);
        });
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const T dx = DX[r], dy = DY[r], dz = DZ[r], ey = EY[r], ez = EZ[r], ry = RY[r], rz = RZ[r];
      const T pdx = partner(dx), pdy = partner(dy), pdz = partner(dz), pey = partner(ey), pez = partner(ez),
              pqx = partner(QX[r]);
      const T bxv = s * (ky * pdz - kz * pdy);
      const T byv = s * kz * pdx - ez;
      const T bzv = ey - s * ky * pdx;
      const T jxv = k2 * dx + s * (ky * pey + kz * pez);
      const T jyv = kzz * dy - kyz * dz - ry + s * ky * pqx;
      const T jzv = kyy * dz - kyz * dy - rz + s * kz * pqx;
      const int x = x0 + MM::row(L.l, r);
      if (L.valid && x < g.n0) {
        const size_t oo = (size_t)x * 2 * M + 2 * L.mcol + L.comp;
        Xr[oo] = bxv;
        Xr[plane + oo] = byv;
        Xr[2 * plane + oo] = bzv;
        Xr[3 * plane + oo] = jxv;
        Xr[4 * plane + oo] = jyv;
        Xr[5 * plane + oo] = jzv;
      }
    }
  }
}

// Induction rows of the nodal EMF E = u_tot x B (Hx fields 3..5):
// NA_x = s0 mA0 E_x and NA_{y,z} = s12 mA12 E_{y,z} — the mass rows of the
// forward transforms of E (_A_forward_emf + _A_mass_rhs,
// examples/pcf_mhd_jax.py:157-170; examples/pcf_mhd_mri_shearpy_jax.py:633-659).
template <typename T>
__global__ void __launch_bounds__(512) k_rhs_a_mfma(Geo g, MhdMats<T> mm,
                                                    const typename VT<T>::V *__restrict__ Hx,
                                                    double2 *__restrict__ NA, size_t cstride) {
  using MM = Mfma<T>;
  using V = typename VT<T>::V;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ct = blockIdx.x;
  const int M = g.m, kp = (g.n0 + 3) / 4 * 4;
  const size_t plane = (size_t)g.n0 * M;
  const Panel<T> B{reinterpret_cast<T *>(smem), kp};
  {
    const int total = kp * 8;
    V hv[STAGE_RHS][3];
#pragma unroll
    for (int t = 0; t < STAGE_RHS; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      const int x = idx >> 3, m = 8 * ct + (idx & 7);
      hv[t][0] = hv[t][1] = hv[t][2] = V{0, 0};
      if (idx < total && m < M && x < g.n0) {
        const size_t o = (size_t)x * M + m;
        hv[t][0] = Hx[3 * plane + o];
        hv[t][1] = Hx[4 * plane + o];
        hv[t][2] = Hx[5 * plane + o];
      }
    }
#pragma unroll
    for (int t = 0; t < STAGE_RHS; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      if (idx >= total) continue;
      const int x = idx >> 3, j = idx & 7;
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        B.at(f, x, 2 * j) = hv[t][f].x;
        B.at(f, x, 2 * j + 1) = hv[t][f].y;
      }
    }
  }
  __syncthreads();
  const Lane L = lane_of(g, ct);
  double *NAr = reinterpret_cast<double *>(NA);
  const int rows = mm.n0r > mm.n12r ? mm.n0r : mm.n12r;
  for (int i0 = (threadIdx.x >> 6) * 16; i0 < rows; i0 += (blockDim.x >> 6) * 16) {
    const int ia = i0 + (L.l & 15);
    typename MM::acc C0{}, C1{}, C2{};
    const int kq = L.l >> 4;
    // rows past a matrix's width are never stored: clamp their column
    const int ia0 = ia < mm.ld0 ? ia : mm.ld0 - 1, ia12 = ia < mm.ld12 ? ia : mm.ld12 - 1;
    k_loop<T, 2, 8>(
        kp / 4,
        [&](int it, T(&a)[2]) {
          const int x = 4 * it + kq;
          a[0] = mat_at(mm.mA0, mm.ld0, x, ia0);
          a[1] = mat_at(mm.mA12, mm.ld12, x, ia12);
        },
        [&](int it, const T(&a)[2]) {
          const int x = 4 * it + kq;
          C0 = MM::mma(a[0], B.at(0, x, L.c), C0);
          C1 = MM::mma(a[1], B.at(1, x, L.c), C1);
          C2 = MM::mma(a[1], B.at(2, x, L.c), C2);
        });
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + MM::row(L.l, r);
      if (!L.valid) continue;
      const size_t o = (size_t)i * 2 * M + 2 * L.mcol + L.comp;
      if (i < mm.n0r) NAr[o] = mm.s0 * (double)C0[r];
      if (i < mm.n12r) {
        NAr[2 * cstride + o] = mm.s12 * (double)C1[r];
        NAr[4 * cstride + o] = mm.s12 * (double)C2[r];
      }
    }
  }
}

// Nu = W (A1 c1 + A0 c0), Ng = W AG cg with c1 = i ky h1 + i kz h2, c0 = k^2 h0,
// cg = i kz h1 - i ky h2 (rows: Shen index i, K: nodes x); mean mode
// Nv00 = -AG Re h1, Nw00 = -AG Re h2 by the block that holds mode 0.
template <typename T>
__global__ void __launch_bounds__(512) k_rhs_mfma(Geo g, NodalMats<T> nm,
                                                  const typename VT<T>::V *__restrict__ Hx,
                                                  double2 *__restrict__ Nu, double2 *__restrict__ Ng,
                                                  double *__restrict__ Nv, double *__restrict__ Nw) {
  using MM = Mfma<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ct = blockIdx.x;
  const int M = g.m, kp = (g.n0 + 3) / 4 * 4;
  const size_t plane = (size_t)g.n0 * M;
  const Panel<T> B{reinterpret_cast<T *>(smem), kp};
  {
    using V = typename VT<T>::V;
    const int total = kp * 8;
    V hv[STAGE_RHS][3];
#pragma unroll
    for (int t = 0; t < STAGE_RHS; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      const int x = idx >> 3, m = 8 * ct + (idx & 7);
      hv[t][0] = hv[t][1] = hv[t][2] = V{0, 0};
      if (idx < total && m < M && x < g.n0) {
        const size_t o = (size_t)x * M + m;
        hv[t][0] = Hx[o];
        hv[t][1] = Hx[plane + o];
        hv[t][2] = Hx[2 * plane + o];
      }
    }
#pragma unroll
    for (int t = 0; t < STAGE_RHS; ++t) {
    const int idx = threadIdx.x + t * blockDim.x;
    if (idx >= total) continue;
    const int x = idx >> 3, j = idx & 7;
    const int m = 8 * ct + j;
    T c1r = 0, c1i = 0, c0r = 0, c0i = 0, cgr = 0, cgi = 0;
    if (m < M && x < g.n0) {
      double kyd, kzd, kyo;
      mode_k(g, m, kyd, kzd, kyo);
      const T ky = (T)kyo, kz = (T)kzd, k2 = (T)(kyd * kyd + kzd * kzd);
      const V h0 = hv[t][0], h1 = hv[t][1], h2 = hv[t][2];
      c1r = -ky * h1.y - kz * h2.y;
      c1i = ky * h1.x + kz * h2.x;
      cgr = -kz * h1.y + ky * h2.y;
      cgi = kz * h1.x - ky * h2.x;
      c0r = k2 * h0.x;
      c0i = k2 * h0.y;
      if (m == 0) {  // mode 0 has cg = 0: its columns carry (Re h1, Re h2) for the mean rows
        cgr = h1.x;
        cgi = h2.x;
      }
    }
    B.at(0, x, 2 * j) = c1r;
    B.at(0, x, 2 * j + 1) = c1i;
    B.at(1, x, 2 * j) = c0r;
    B.at(1, x, 2 * j + 1) = c0i;
    B.at(2, x, 2 * j) = cgr;
    B.at(2, x, 2 * j + 1) = cgi;
    }
  }
  __syncthreads();
  const Lane L = lane_of(g, ct);
  for (int i0 = (threadIdx.x >> 6) * 16; i0 < g.nd; i0 += (blockDim.x >> 6) * 16) {
    const int ia = i0 + (L.l & 15);
    typename MM::acc AU{}, AG{};
    const int kq = L.l >> 4;
    k_loop<T, 3, 8>(
        kp / 4,
        [&](int it, T(&a)[3]) {
          const int x = 4 * it + kq;
          a[0] = mat_at(nm.A1t, nm.ldi, x, ia);
          a[1] = mat_at(nm.A0t, nm.ldi, x, ia);
          a[2] = mat_at(nm.AGt, nm.ldi, x, ia);
        },
        [&](int it, const T(&a)[3]) {
          const int x = 4 * it + kq;
          AU = MM::mma(a[0], B.at(0, x, L.c), AU);
          AU = MM::mma(a[1], B.at(1, x, L.c), AU);
          AG = MM::mma(a[2], B.at(2, x, L.c), AG);
        });
    const double W = g.W;
    double *Nur = reinterpret_cast<double *>(Nu);
    double *Ngr = reinterpret_cast<double *>(Ng);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + MM::row(L.l, r);
      if (!L.valid) continue;
      const size_t o = (size_t)i * 2 * M + 2 * L.mcol + L.comp;
      if (i < g.nb) Nur[o] = W * (double)AU[r];
      if (i < g.nd) {
        if (L.mcol == 0) {
          Ngr[o] = 0.0;
          (L.comp ? Nw : Nv)[i] = -(double)AG[r];
        } else {
          Ngr[o] = W * (double)AG[r];
        }
      }
    }
  }
}

// ------------------------------------------------------------ adjoint ----
// mu_u0 = VB^T (X0 - i kz W1 + i ky W2), mu_u1 = VD^T (X1 + i kz W0) + VD1^T W2,
// mu_u2 = VD^T (X2 - i ky W0) - VD1^T W1 (rows: Shen index i, K: nodes x).
template <typename T>
__global__ void __launch_bounds__(512) k_nodal_adj_mfma(Geo g, AdjMats<T> am,
                                                        const typename VT<T>::V *__restrict__ X,
                                                        double2 *__restrict__ mu0,
                                                        double2 *__restrict__ mu1,
                                                        double2 *__restrict__ mu2) {
  using MM = Mfma<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ct = blockIdx.x;
  const int M = g.m, kp = (g.n0 + 3) / 4 * 4;
  const size_t plane = (size_t)g.n0 * M;
  const Panel<T> B{reinterpret_cast<T *>(smem), kp};
  using V = typename VT<T>::V;
  const int total = kp * 8;
  V xv[STAGE_NODAL_ADJ][6];
#pragma unroll
  for (int t = 0; t < STAGE_NODAL_ADJ; ++t) {
    const int idx = threadIdx.x + t * blockDim.x;
    const int x = idx >> 3, m = 8 * ct + (idx & 7);
#pragma unroll
    for (int f = 0; f < 6; ++f) xv[t][f] = V{0, 0};
    if (idx < total && m < M && x < g.n0 && !nyq_mode(g, m)) {
      const size_t o = (size_t)x * M + m;
#pragma unroll
      for (int f = 0; f < 6; ++f) xv[t][f] = X[f * plane + o];
    }
  }
#pragma unroll
  for (int t = 0; t < STAGE_NODAL_ADJ; ++t) {
    const int idx = threadIdx.x + t * blockDim.x;
    if (idx >= total) continue;
    const int x = idx >> 3, j = idx & 7;
    const int m = 8 * ct + j;
    T v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (m < M && x < g.n0 && !nyq_mode(g, m)) {
      double kyd, kzd, kyo;
      mode_k(g, m, kyd, kzd, kyo);
      const T ky = (T)kyo, kz = (T)kzd;
      const V X0 = xv[t][0], X1 = xv[t][1], X2 = xv[t][2];
      const V W0 = xv[t][3], W1 = xv[t][4], W2 = xv[t][5];
      v[0] = X0.x + kz * W1.y - ky * W2.y;  // t0
      v[1] = X0.y - kz * W1.x + ky * W2.x;
      v[2] = X1.x - kz * W0.y;  // t1
      v[3] = X1.y + kz * W0.x;
      v[4] = X2.x + ky * W0.y;  // t2
      v[5] = X2.y - ky * W0.x;
      v[6] = W2.x;  // w2
      v[7] = W2.y;
      v[8] = -W1.x;  // -w1
      v[9] = -W1.y;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      B.at(f, x, 2 * j) = v[2 * f];
      B.at(f, x, 2 * j + 1) = v[2 * f + 1];
    }
  }
  __syncthreads();
  const Lane L = lane_of(g, ct);
  double *o0 = reinterpret_cast<double *>(mu0);
  double *o1 = reinterpret_cast<double *>(mu1);
  double *o2 = reinterpret_cast<double *>(mu2);
  for (int i0 = (threadIdx.x >> 6) * 16; i0 < g.nd; i0 += (blockDim.x >> 6) * 16) {
  const int ia = i0 + (L.l & 15);
  typename MM::acc A0{}, A1{}, A2{};
  const int kq = L.l >> 4;
  k_loop<T, 3, 8>(
      kp / 4,
      [&](int it, T(&a)[3]) {
        const int x = 4 * it + kq;
        a[0] = mat_at(am.VB, am.ldi, x, ia);
        a[1] = mat_at(am.VD, am.ldi, x, ia);
        a[2] = mat_at(am.VD1, am.ldi, x, ia);
      },
      [&](int it, const T(&a)[3]) {
        const int x = 4 * it + kq;
        A0 = MM::mma(a[0], B.at(0, x, L.c), A0);
        A1 = MM::mma(a[1], B.at(1, x, L.c), A1);
        A1 = MM::mma(a[2], B.at(3, x, L.c), A1);
        A2 = MM::mma(a[1], B.at(2, x, L.c), A2);
        A2 = MM::mma(a[2], B.at(4, x, L.c), A2);
      });
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + MM::row(L.l, r);
    if (!L.valid) continue;
    const size_t o = (size_t)i * 2 * M + 2 * L.mcol + L.comp;
    if (i < g.nb) o0[o] = (double)A0[r];
    if (i < g.nd) {
      o1[o] = (double)A1[r];
      o2[o] = (double)A2[r];
    }
  }
  }
}

// The same contraction with the reconstruction adjoint and the steady SBDF3
// mass gather folded in, producing mu_x of a steady reverse step directly:
//   mu_x.u0 = VB^T t0 + (VD R)^T c1 + (VD1 R)^T c2 + M0^T sig.u0,
//   mu_x.g  = VD^T d1 + VD1^T d2 + M1^T sig.g,
// with t0..t2, w1, w2 the staged fields above, a = ky / k^2, b = kz / k^2 and
//   c1 = -i (a t1 + b t2), c2 = -i (a w2 - b w1),
//   d1 = -i (b t1 - a t2), d2 = -i (b w2 + a w1),
// since mu_u1 = VD^T t1 + VD1^T w2, mu_u2 = VD^T t2 - VD1^T w1 and the
// reconstruction u1 = i (a f + b g), u2 = i (b f - a g), f = R u0
// (examples/channelflow_kmm.py:855-870).  Mode 0 has no reconstruction: its
// u0 rows are VB^T t0 and its (real) mean cotangents ride in the g rows as
// d1 = t1 + i t2, d2 = w2 - i w1, so the g accumulator holds mu_v + i mu_w.
// sig = sum_j C_j rho_{k+j} (written by the transposed solve); M0 = W (C2_B -
// k^2 C0_B) and M1 = W C0_D are banded per parity block, the mean rows use
// C0_D without W.  Nyquist row: zero.
constexpr int STAGE_SIG = 4;
template <typename T>
__global__ void __launch_bounds__(512) k_nodal_adj_x_mfma(Geo g, AdjMats<T> am, StructMats sm,
                                                          const typename VT<T>::V *__restrict__ X,
                                                          PrimC sig, Prim out) {
  using MM = Mfma<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ct = blockIdx.x;
  const int M = g.m, kp = (g.n0 + 3) / 4 * 4, nb = g.nb, nd = g.nd;
  const size_t plane = (size_t)g.n0 * M;
  const Panel<T> B{reinterpret_cast<T *>(smem), kp};
  double2 *sg = reinterpret_cast<double2 *>(smem + (size_t)5 * kp * 16 * sizeof(T));  // [nb + nd][8]
  using V = typename VT<T>::V;
  const int total = kp * 8, stot = (nb + nd) * 8;
  V xv[STAGE_NODAL_ADJ][6];
  double2 sv[STAGE_SIG];
#pragma unroll
  for (int t = 0; t < STAGE_NODAL_ADJ; ++t) {
    const int idx = threadIdx.x + t * blockDim.x;
    const int x = idx >> 3, m = 8 * ct + (idx & 7);
#pragma unroll
    for (int f = 0; f < 6; ++f) xv[t][f] = V{0, 0};
    if (idx < total && m < M && x < g.n0 && !nyq_mode(g, m)) {
      const size_t o = (size_t)x * M + m;
#pragma unroll
      for (int f = 0; f < 6; ++f) xv[t][f] = X[f * plane + o];
    }
  }
#pragma unroll
  for (int t = 0; t < STAGE_SIG; ++t) {
    const int idx = threadIdx.x + t * blockDim.x;
    const int r = idx >> 3, m = 8 * ct + (idx & 7);
    sv[t] = double2{0, 0};
    if (idx < stot && m < M) sv[t] = (r < nb) ? sig.u0[(size_t)r * M + m] : sig.g[(size_t)(r - nb) * M + m];
  }
#pragma unroll
  for (int t = 0; t < STAGE_NODAL_ADJ; ++t) {
    const int idx = threadIdx.x + t * blockDim.x;
    if (idx >= total) continue;
    const int x = idx >> 3, j = idx & 7;
    const int m = 8 * ct + j;
    T v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (m < M && x < g.n0 && !nyq_mode(g, m)) {
      double kyd, kzd, kyo;
      mode_k(g, m, kyd, kzd, kyo);
      const T ky = (T)kyo, kz = (T)kzd;
      const V X0 = xv[t][0], X1 = xv[t][1], X2 = xv[t][2];
      const V W0 = xv[t][3], W1 = xv[t][4], W2 = xv[t][5];
      const T t0r = X0.x + kz * W1.y - ky * W2.y, t0i = X0.y - kz * W1.x + ky * W2.x;
      const T t1r = X1.x - kz * W0.y, t1i = X1.y + kz * W0.x;
      const T t2r = X2.x + ky * W0.y, t2i = X2.y - ky * W0.x;
      v[0] = t0r;
      v[1] = t0i;
      if (m == 0) {
        v[6] = t1r;  // d1 = t1 + i t2 (real parts)
        v[7] = t2r;
        v[8] = W2.x;  // d2 = w2 - i w1
        v[9] = -W1.x;
      } else {
        const double k2 = kyd * kyd + kzd * kzd;
        const T a = (T)(kyd / k2), b = (T)(kzd / k2);
        // -i z = (z.y, -z.x)
        const T c1r = a * t1r + b * t2r, c1i = a * t1i + b * t2i;
        const T c2r = a * W2.x - b * W1.x, c2i = a * W2.y - b * W1.y;
        const T d1r = b * t1r - a * t2r, d1i = b * t1i - a * t2i;
        const T d2r = b * W2.x + a * W1.x, d2i = b * W2.y + a * W1.y;
        v[2] = c1i;
        v[3] = -c1r;
        v[4] = c2i;
        v[5] = -c2r;
        v[6] = d1i;
        v[7] = -d1r;
        v[8] = d2i;
        v[9] = -d2r;
      }
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      B.at(f, x, 2 * j) = v[2 * f];
      B.at(f, x, 2 * j + 1) = v[2 * f + 1];
    }
  }
#pragma unroll
  for (int t = 0; t < STAGE_SIG; ++t) {
    const int idx = threadIdx.x + t * blockDim.x;
    if (idx < stot) sg[idx] = sv[t];
  }
  __syncthreads();
  const Lane L = lane_of(g, ct);
  const int jl = L.c >> 1;
  double kyd, kzd, kyo;
  mode_k(g, L.valid ? L.mcol : 0, kyd, kzd, kyo);
  const double k2 = kyd * kyd + kzd * kzd;
  const bool live = L.valid && !nyq_mode(g, L.mcol);
  auto sgc = [&](int r) { return L.comp ? sg[r * 8 + jl].y : sg[r * 8 + jl].x; };
  double *ou = reinterpret_cast<double *>(out.u0);
  double *og = reinterpret_cast<double *>(out.g);
  for (int i0 = (threadIdx.x >> 6) * 16; i0 < nd; i0 += (blockDim.x >> 6) * 16) {
    const int ia = i0 + (L.l & 15);
    typename MM::acc AU{}, AG{};
    const int kq = L.l >> 4;
    k_loop<T, 5, 6>(
        kp / 4,
        [&](int it, T(&a)[5]) {
          const size_t o = (size_t)(4 * it + kq) * am.ldi + ia;
          a[0] = am.VB[o];
          a[1] = am.VDR[o];
          a[2] = am.VD1R[o];
          a[3] = am.VD[o];
          a[4] = am.VD1[o];
        },
        [&](int it, const T(&a)[5]) {
          const int x = 4 * it + kq;
          AU = MM::mma(a[0], B.at(0, x, L.c), AU);
          AU = MM::mma(a[1], B.at(1, x, L.c), AU);
          AU = MM::mma(a[2], B.at(2, x, L.c), AU);
          AG = MM::mma(a[3], B.at(3, x, L.c), AG);
          AG = MM::mma(a[4], B.at(4, x, L.c), AG);
        });
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + MM::row(L.l, r);
      if (!L.valid || i >= nd) continue;
      const int par = i & 1, jp = i >> 1;
      const size_t o = (size_t)i * 2 * M + 2 * L.mcol + L.comp;
      if (i < nb) {
        double u = 0.0;
        if (live) {
          const StructSys &s0 = sm.sys[0][par];
          u = (double)AU[r];
          for (int d = -2; d <= 2; ++d) {
            const int ii = jp - d;
            if (ii < 0 || ii >= s0.n) continue;
            const size_t bo = (size_t)ii * 5 + (d + 2);
            u += g.W * (s0.b2[bo] - k2 * s0.b0[bo]) * sgc(2 * ii + par);
          }
        }
        ou[o] = u;
      }
      double gv = 0.0, mv = (double)AG[r];
      const StructSys &s1 = sm.sys[1][par];
      if (live) {
        for (int d = -1; d <= 1; ++d) {
          const int ii = jp - d;
          if (ii < 0 || ii >= s1.n) continue;
          const double c = s1.b0[(size_t)ii * 3 + (d + 1)];
          gv += g.W * c * sgc(nb + 2 * ii + par);
          if (L.mcol == 0) mv += c * (L.comp ? sig.w : sig.v)[2 * ii + par];
        }
      }
      og[o] = (L.mcol == 0) ? gv : (live ? gv + (double)AG[r] : 0.0);
      if (L.mcol == 0) (L.comp ? out.w : out.v)[i] = mv;
    }
  }
}

// mu_h0 = W k^2 A0^T nu_u, mu_h1 = -i W (ky A1^T nu_u + kz AG^T nu_g),
// mu_h2 = -i W (kz A1^T nu_u - ky AG^T nu_g) (rows: nodes x, K: Shen index i);
// mean mode: mu_h1,2 -= AG^T nu_v,w.  Nyquist row: zero.
template <typename T>
__global__ void __launch_bounds__(512) k_rhs_adj_mfma(Geo g, AdjMats<T> am, Prim nu,
                                                      typename VT<T>::V *__restrict__ Hx) {
  using MM = Mfma<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ct = blockIdx.x;
  const int M = g.m, kp = (g.nd + 3) / 4 * 4;
  const Panel<T> B{reinterpret_cast<T *>(smem), kp};
  {
    const int total = 2 * kp * 8;
    double2 z[STAGE_RHS_ADJ];
#pragma unroll
    for (int t = 0; t < STAGE_RHS_ADJ; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      const int f = idx / (kp * 8), rem = idx - f * kp * 8, i = rem >> 3, j = rem & 7;
      const int m = 8 * ct + j;
      z[t] = double2{0, 0};
      if (idx < total && m < M && !nyq_mode(g, m)) {
        if (f == 0 && i < g.nb) z[t] = nu.u0[(size_t)i * M + m];
        // mode 0 contributes nothing through its wavenumbers: its g columns
        // carry the mean cotangents (nu_v, nu_w) through the AG product
        if (f == 1 && i < g.nd) z[t] = (m == 0) ? double2{nu.v[i], nu.w[i]} : nu.g[(size_t)i * M + m];
      }
    }
#pragma unroll
    for (int t = 0; t < STAGE_RHS_ADJ; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      if (idx >= total) continue;
      const int f = idx / (kp * 8), rem = idx - f * kp * 8, i = rem >> 3, j = rem & 7;
      B.at(f, i, 2 * j) = (T)z[t].x;
      B.at(f, i, 2 * j + 1) = (T)z[t].y;
    }
  }
  __syncthreads();
  const Lane L = lane_of(g, ct);
  double kyd, kzd, kyo;
  mode_k(g, L.valid ? L.mcol : 0, kyd, kzd, kyo);
  const bool live = L.valid && !nyq_mode(g, L.mcol);
  const T W = (T)g.W, ky = (T)kyo, kz = (T)kzd, k2 = (T)(kyd * kyd + kzd * kzd);
  const T s = L.comp ? (T)1 : (T)-1;  // -i z: re lane +z.im, im lane -z.re
  const size_t plane = (size_t)g.n0 * 2 * M;
  T *Hr = reinterpret_cast<T *>(Hx);
  const bool mean_col = L.valid && L.mcol == 0;
  for (int x0 = (threadIdx.x >> 6) * 16; x0 < g.n0; x0 += (blockDim.x >> 6) * 16) {
  const int xa = x0 + (L.l & 15);
  typename MM::acc P0{}, P1{}, PG{};
  const int kq = L.l >> 4;
  k_loop<T, 3, 8>(
      kp / 4,
      [&](int it, T(&a)[3]) {
        const int i = 4 * it + kq;
        a[0] = mat_at(am.A0, am.ldx, i, xa);
        a[1] = mat_at(am.A1, am.ldx, i, xa);
        a[2] = mat_at(am.AG, am.ldx, i, xa);
      },
      [&](int it, const T(&a)[3]) {
        const int i = 4 * it + kq;
        const T bu = B.at(0, i, L.c), bg = B.at(1, i, L.c);
        P0 = MM::mma(a[0], bu, P0);
        P1 = MM::mma(a[1], bu, P1);
        PG = MM::mma(a[2], bg, PG);
      });
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const T p1 = partner(P1[r]), pg = partner(PG[r]);
    T h0 = W * k2 * P0[r];
    T h1 = -s * W * (ky * p1 + kz * pg);
    T h2 = -s * W * (kz * p1 - ky * pg);
    const int x = x0 + MM::row(L.l, r);
    if (mean_col) {  // the mean mode is real: - AG^T nu_v, - AG^T nu_w in its real lane
      h0 = 0;
      h1 = L.comp ? (T)0 : -PG[r];
      h2 = L.comp ? (T)0 : -pg;
    }
    if (L.valid && x < g.n0) {
      const size_t o = (size_t)x * 2 * M + 2 * L.mcol + L.comp;
      Hr[o] = live ? h0 : (T)0;
      Hr[plane + o] = live ? h1 : (T)0;
      Hr[2 * plane + o] = live ? h2 : (T)0;
    }
  }
  }
}

}  // namespace

// ------------------------------------------------------ reconstruction ----
// Velocity of a solved state (channelflow_kmm.py:855-870): f = R u0 with
// R = P_D d/dx (nd x nb) on the matrix cores, then per mode
// u1 = i (ky f + kz g) / k^2, u2 = i (kz f - ky g) / k^2 with the derivative
// wavenumbers (Nyquist row eliminated); the mean mode takes u1 = v00,
// u2 = w00.  FP64; u0 is read once per block instead of once per output row.
__global__ void __launch_bounds__(512) k_recon_mfma(Geo g, const double *__restrict__ Rt, int ldr,
                                                    const double2 *__restrict__ u0,
                                                    const double2 *__restrict__ gg,
                                                    const double *__restrict__ v,
                                                    const double *__restrict__ w, double2 *__restrict__ u1,
                                                    double2 *__restrict__ u2) {
  using MM = Mfma<double>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ct = blockIdx.x;
  const int M = g.m, kp = (g.nb + 3) / 4 * 4;
  const Panel<double> B{reinterpret_cast<double *>(smem), kp};
  {
    const int total = kp * 8;
    double2 z[STAGE_RECON];
#pragma unroll
    for (int t = 0; t < STAGE_RECON; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      const int k = idx >> 3, m = 8 * ct + (idx & 7);
      z[t] = double2{0, 0};
      if (idx < total && m < M && k < g.nb) z[t] = u0[(size_t)k * M + m];
    }
#pragma unroll
    for (int t = 0; t < STAGE_RECON; ++t) {
      const int idx = threadIdx.x + t * blockDim.x;
      if (idx >= total) continue;
      const int k = idx >> 3, j = idx & 7;
      B.at(0, k, 2 * j) = z[t].x;
      B.at(0, k, 2 * j + 1) = z[t].y;
    }
  }
  __syncthreads();
  const Lane L = lane_of(g, ct);
  double kyd, kzd, kyo;
  mode_k(g, L.valid ? L.mcol : 0, kyd, kzd, kyo);
  const double k2 = kyd * kyd + kzd * kzd, den = (k2 == 0.0) ? 1.0 : k2;
  const double ky = kyd / den, kz = kzd / den;
  const double s = L.comp ? 1.0 : -1.0;  // (i z) of this lane's component = s * z of the partner's
  const int kq = L.l >> 4;
  for (int x0 = (threadIdx.x >> 6) * 16; x0 < g.nd; x0 += (blockDim.x >> 6) * 16) {
    const int xa = x0 + (L.l & 15);
    // the epilogue's g values (mean mode: v00, w00) in flight during the K loop
    double gp[4], vm[4], wm[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int x = x0 + MM::row(L.l, r);
      gp[r] = vm[r] = wm[r] = 0.0;
      if (L.valid && x < g.nd) {
        if (L.mcol == 0) {
          vm[r] = v[x];
          wm[r] = w[x];
        } else {
          const double2 gv = gg[(size_t)x * M + L.mcol];
          gp[r] = L.comp ? gv.x : gv.y;
        }
      }
    }
    MM::acc F{};
    k_loop<double, 1, 8>(
        kp / 4, [&](int it, double(&a)[1]) { a[0] = mat_at(Rt, ldr, 4 * it + kq, xa); },
        [&](int it, const double(&a)[1]) { F = MM::mma(a[0], B.at(0, 4 * it + kq, L.c), F); });
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double fp = partner(F[r]);
      const int x = x0 + MM::row(L.l, r);
      if (L.valid && x < g.nd) {
        const size_t o = (size_t)x * M + L.mcol;
        double a1, a2;
        if (L.mcol == 0) {
          a1 = L.comp ? 0.0 : vm[r];
          a2 = L.comp ? 0.0 : wm[r];
        } else {
          a1 = s * (ky * fp + kz * gp[r]);
          a2 = s * (kz * fp - ky * gp[r]);
        }
        reinterpret_cast<double *>(u1)[2 * o + L.comp] = a1;
        reinterpret_cast<double *>(u2)[2 * o + L.comp] = a2;
      }
    }
  }
}

// One block per tile of 8 modes, one wave per 16 output rows; LDS holds the
// block's B panels: nf fields x roundup(K, 4) rows x 16 columns of T.
static dim3 mfma_grid(const Geo &g) { return dim3((g.m + 7) / 8); }
static dim3 mfma_block(int rows) { return dim3(64 * std::min(8, (rows + 15) / 16)); }
// true when blockDim * stage loads cover `items` staged items
static bool stage_covers(int rows, int stage, int items) { return (int)mfma_block(rows).x * stage >= items; }
template <typename T>
static size_t panel_bytes(int nf, int K) {
  return (size_t)nf * ((K + 3) / 4 * 4) * 16 * sizeof(T);
}

template <typename T>
hipError_t launch_nodal_mfma(const Geo &g, const NodalMats<T> &nm, const double2 *u0, const double2 *u1,
                             const double2 *u2, Workspace &ws, hipStream_t s) {
  using V = typename VT<T>::V;
  if (!stage_covers(g.n0, STAGE_NODAL, 3 * ((g.nd + 3) / 4 * 4) * 8)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_nodal_mfma<T>), mfma_grid(g), mfma_block(g.n0), panel_bytes<T>(3, g.nd), s, g, nm,
                     u0, u1, u2, reinterpret_cast<V *>(ws.X));
  return hipGetLastError();
}

template <typename T>
hipError_t launch_nodal_prim_mfma(const Geo &g, const NodalMats<T> &nm, const double2 *u0, const double2 *gf,
                                  const double *v00, const double *w00, Workspace &ws, hipStream_t s) {
  using V = typename VT<T>::V;
  if (!nm.VDRt || !nm.VD1Rt) return hipErrorInvalidValue;
  if (!stage_covers(g.n0, STAGE_NODAL_PRIM, 2 * ((g.nd + 3) / 4 * 4) * 8)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_nodal_prim_mfma<T>), mfma_grid(g), mfma_block(g.n0), panel_bytes<T>(2, g.nd), s, g, nm,
                     u0, gf, v00, w00, reinterpret_cast<V *>(ws.X));
  return hipGetLastError();
}

hipError_t launch_reconstruct(const Geo &g, const double *Rt, int ldr, const SolveArgs &a, hipStream_t s) {
  if (ldr < g.nd || ldr % 16 || !stage_covers(g.nd, STAGE_RECON, ((g.nb + 3) / 4 * 4) * 8))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_recon_mfma, mfma_grid(g), mfma_block(g.nd), panel_bytes<double>(1, g.nb), s, g, Rt, ldr,
                     a.u0_out, a.g_out, a.v_out, a.w_out, a.u1, a.u2);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rhs_mfma(const Geo &g, const NodalMats<T> &nm, Workspace &ws, double2 *Nu, double2 *Ng,
                           double *Nv, double *Nw, hipStream_t s) {
  using V = typename VT<T>::V;
  if (!stage_covers(g.nd, STAGE_RHS, ((g.n0 + 3) / 4 * 4) * 8)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_rhs_mfma<T>), mfma_grid(g), mfma_block(g.nd), panel_bytes<T>(3, g.n0), s, g, nm,
                     reinterpret_cast<const V *>(ws.Hx), Nu, Ng, Nv, Nw);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_nodal_mhd_mfma(const Geo &g, const MhdMats<T> &mm, const double2 *A, size_t cstride,
                                 Workspace &ws, hipStream_t s) {
  using V = typename VT<T>::V;
  if (!mm.mDX || !mm.mQX || !mm.mDY || !mm.mEY || !mm.mRY) return hipErrorInvalidValue;
  if (mm.kx > mm.kyz || mm.kyz > g.n0 || mm.ldx < g.n0 || mm.ldx % 16) return hipErrorInvalidValue;
  if (!stage_covers(g.n0, STAGE_NODAL, 3 * ((mm.kyz + 3) / 4 * 4) * 8)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_nodal_mhd_mfma<T>), mfma_grid(g), mfma_block(g.n0), panel_bytes<T>(3, mm.kyz), s, g,
                     mm, A, cstride, reinterpret_cast<V *>(ws.X));
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rhs_a_mfma(const Geo &g, const MhdMats<T> &mm, Workspace &ws, double2 *NA, size_t cstride,
                             hipStream_t s) {
  using V = typename VT<T>::V;
  const int rows = std::max(mm.n0r, mm.n12r);
  if (!mm.mA0 || !mm.mA12 || mm.n0r > mm.ld0 || mm.n12r > mm.ld12) return hipErrorInvalidValue;
  if ((size_t)rows * g.m > cstride) return hipErrorInvalidValue;
  if (!stage_covers(rows, STAGE_RHS, ((g.n0 + 3) / 4 * 4) * 8)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_rhs_a_mfma<T>), mfma_grid(g), mfma_block(rows), panel_bytes<T>(3, g.n0), s, g, mm,
                     reinterpret_cast<const V *>(ws.Hx), NA, cstride);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_nodal_adj_mfma(const Geo &g, const AdjMats<T> &am, Workspace &ws, double2 *mu0,
                                 double2 *mu1, double2 *mu2, hipStream_t s) {
  using V = typename VT<T>::V;
  if (!stage_covers(g.nd, STAGE_NODAL_ADJ, ((g.n0 + 3) / 4 * 4) * 8)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_nodal_adj_mfma<T>), mfma_grid(g), mfma_block(g.nd), panel_bytes<T>(5, g.n0), s, g,
                     am, reinterpret_cast<const V *>(ws.X), mu0, mu1, mu2);
  return hipGetLastError();
}

template <typename T>
static size_t nodal_adj_x_bytes(const Geo &g) {
  return panel_bytes<T>(5, g.n0) + (size_t)(g.nb + g.nd) * 8 * sizeof(double2);
}

template <typename T>
hipError_t launch_nodal_adj_x_mfma(const Geo &g, const AdjMats<T> &am, const StructMats &sm, PrimC sig,
                                   Workspace &ws, Prim out, hipStream_t s) {
  using V = typename VT<T>::V;
  if (!am.VDR || !am.VD1R || !sig.u0 || !sig.g || !sig.v || !sig.w) return hipErrorInvalidValue;
  if (!stage_covers(g.nd, STAGE_NODAL_ADJ, ((g.n0 + 3) / 4 * 4) * 8)) return hipErrorInvalidValue;
  if (!stage_covers(g.nd, STAGE_SIG, (g.nb + g.nd) * 8)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_nodal_adj_x_mfma<T>), mfma_grid(g), mfma_block(g.nd), nodal_adj_x_bytes<T>(g), s, g,
                     am, sm, reinterpret_cast<const V *>(ws.X), sig, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rhs_adj_mfma(const Geo &g, const AdjMats<T> &am, Prim nu, Workspace &ws, hipStream_t s) {
  using V = typename VT<T>::V;
  if (!stage_covers(g.n0, STAGE_RHS_ADJ, 2 * ((g.nd + 3) / 4 * 4) * 8)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_rhs_adj_mfma<T>), mfma_grid(g), mfma_block(g.n0),
                     panel_bytes<T>(2, g.nd), s, g, am, nu,
                     reinterpret_cast<V *>(ws.Hx));
  return hipGetLastError();
}

hipError_t mfma_init(const Geo &g, bool fp64) {
  // the largest panels (five fields of the transform adjoint, plus the gathered
  // rows of the fused variant) may exceed the default 64 KB of dynamic LDS
  auto raise = [](const void *fn, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
  };
  hipError_t e;
  if (fp64) {
    if ((e = raise(reinterpret_cast<const void *>(&k_nodal_adj_mfma<double>), panel_bytes<double>(5, g.n0))))
      return e;
    return raise(reinterpret_cast<const void *>(&k_nodal_adj_x_mfma<double>), nodal_adj_x_bytes<double>(g));
  }
  if ((e = raise(reinterpret_cast<const void *>(&k_nodal_adj_mfma<float>), panel_bytes<float>(5, g.n0))))
    return e;
  return raise(reinterpret_cast<const void *>(&k_nodal_adj_x_mfma<float>), nodal_adj_x_bytes<float>(g));
}

#define PCF_MFMA_INSTANTIATE(T)                                                                     \
  template hipError_t launch_nodal_mfma<T>(const Geo &, const NodalMats<T> &, const double2 *,      \
                                           const double2 *, const double2 *, Workspace &,           \
                                           hipStream_t);                                            \
  template hipError_t launch_nodal_prim_mfma<T>(const Geo &, const NodalMats<T> &, const double2 *, \
                                                const double2 *, const double *, const double *,      \
                                                Workspace &, hipStream_t);                           \
  template hipError_t launch_rhs_mfma<T>(const Geo &, const NodalMats<T> &, Workspace &, double2 *, \
                                         double2 *, double *, double *, hipStream_t);               \
  template hipError_t launch_nodal_adj_mfma<T>(const Geo &, const AdjMats<T> &, Workspace &,        \
                                               double2 *, double2 *, double2 *, hipStream_t);       \
  template hipError_t launch_rhs_adj_mfma<T>(const Geo &, const AdjMats<T> &, Prim, Workspace &,    \
                                             hipStream_t);                                          \
  template hipError_t launch_nodal_adj_x_mfma<T>(const Geo &, const AdjMats<T> &, const StructMats &, \
                                                 PrimC, Workspace &, Prim, hipStream_t);            \
  template hipError_t launch_nodal_mhd_mfma<T>(const Geo &, const MhdMats<T> &, const double2 *,    \
                                               size_t, Workspace &, hipStream_t);                   \
  template hipError_t launch_rhs_a_mfma<T>(const Geo &, const MhdMats<T> &, Workspace &, double2 *,   \
                                           size_t, hipStream_t);
PCF_MFMA_INSTANTIATE(double)
PCF_MFMA_INSTANTIATE(float)
#undef PCF_MFMA_INSTANTIATE

}  // namespace pcf
