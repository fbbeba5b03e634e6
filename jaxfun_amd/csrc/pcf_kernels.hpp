// Kernel interfaces for the plane-Couette KMM step (gfx950).
//
// Device data layout (all complex values interleaved re/im):
//   spectral field   [n][m]          m = iy * KZ + kz, kz < KZ = nz/2
//                                    (Hermitian half spectrum; the spanwise
//                                    Nyquist plane is identically zero and not
//                                    stored; the streamwise Nyquist row iy=ny/2
//                                    is stored and kept zero)
//   nodal pencils    [f][x][m]       after the wall-normal transform
//   padded y-lines   [f][x][y][kz]   y < MY = floor(1.5 ny)
// Mode index is fastest everywhere, so the Chebyshev contractions read and
// write whole 64-wide rows and the per-mode solve reads one 16-byte element
// per coefficient row.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace pcf {

struct Geo {
  int n0, nb, nd;  // Chebyshev points / clamped / Dirichlet dofs
  int ny, nz, kzs; // Fourier modes; kzs = nz/2 stored spanwise modes
  int my, mz;      // padded physical grid
  int m;           // modes per wall-normal index = ny * kzs
  double df1, df2; // 2 pi / L (scaled wavenumber factors)
  double uwall;    // base flow U = uwall * x
  double W;        // Fourier mass weight (2pi/df1)(2pi/df2) of every Galerkin row
  double nu;
  int npmax;       // largest parity block of the per-mode systems (<= 64)
  double b0[3];    // uniform background magnetic field (MHD)
  double omega2;   // twice the rotation rate of the Coriolis force (MHD)
};

// Radix plan of one FFT length (radices 2, 3, 4).
struct FFTPlan {
  int n;
  int nst;
  int radix[12];
};

// Contraction matrices in the precision T, transposed so that the 16 output
// rows of one matrix-core tile are contiguous for each contraction index.
// Every matrix carries zero rows up to a multiple of 4 >= n0 (whole K steps).
template <typename T>
struct NodalMats {
  const T *VBt;   // [nb][ldx]  VB^T  (zero padded to ldx)
  const T *VDt;   // [nd][ldx]
  const T *VD1t;  // [nd][ldx]
  int ldx;        // >= n0, multiple of 16
  const T *A1t;   // [n0][ldi]  A1^T
  const T *A0t;   // [n0][ldi]
  const T *AGt;   // [n0][ldi]
  int ldi;        // >= nd, multiple of 16
  const T *xnode; // [n0]
  const T *twy;   // [my] exp(-2 pi i t / my) as (cos, sin) pairs
  const T *twz;   // [mz]
  const T *VDRt;  // [nb][ldx]  (VD R)^T: nodal values of the reconstructed f = R u0
  const T *VD1Rt; // [nb][ldx]  (VD1 R)^T
};

// Wall-normal matrices of the MHD contractions, transposed ([K][ld]) like
// NodalMats.  Nodal b and j take DX = mDX Ax, QX = mQX Ax (K = kx rows of Ax)
// and DY/DZ = mDY A{y,z}, EY/EZ = mEY A{y,z}, RY/RZ = mRY A{y,z} (K = kyz rows);
// the induction rows are NA_x = s0 mA0 E_x (n0r rows) and
// NA_{y,z} = s12 mA12 E_{y,z} (n12r rows), K = nodes.
//   conducting walls: mDX = mDY = VD, mQX = VQ1, mEY = VD1, mRY = VQ2,
//                     mA0 = mA12 = AG, s0 = s12 = W;
//   insulating walls: mDX = VD, mQX = VD1, mDY = VC, mEY = VC1, mRY = VC2,
//                     mA0 = AG (s0 = W), mA12 = WQ (s12 = 1).
template <typename T>
struct MhdMats {
  const T *mDX, *mQX, *mDY, *mEY, *mRY;
  int ldx, kx, kyz;
  const T *mA0, *mA12;
  int ld0, ld12, n0r, n12r;
  double s0, s12;
};

// Structured per-mode operator data of one (basis, parity) block: the band
// of each Galerkin block (row-major [n][P+Q+1], offsets -P..Q) and the far
// factors U [n][R], V [n][R] with C[i][j] = sum_r U_r(i) V_r(j) for j > i+Q
// (C4 for the clamped basis, C2 for the Dirichlet basis; the others have no
// far part).
struct StructSys {
  int n;
  const double *b0, *b2, *b4;
  const double *U, *V;
};

struct StructMats {
  StructSys sys[2][2];  // [basis: 0 clamped, 1 Dirichlet][parity]
};

// One implicit update: rhs = M z + sum_l e_l N_l with z = sum c_l x_l; the
// factored solve of (M - gamma L) y = rhs writes x = y - sub x_0, then u1/u2
// are rebuilt from x.  A Crank-Nicolson-form update rhs = M x_0 + gamma L x_0
// + sum e N at the same gamma is (M - gamma L)(x + x_0) = 2 M x_0 + sum e N,
// i.e. z = 2 x_0, sub = 1.  d (zL = sum d_l x_l, an explicit L zL term) is read
// only where the update feeds the bordered solve.
struct SolveArgs {
  const double2 *u0[3];
  const double2 *g[3];
  const double *v[3];
  const double *w[3];
  double c[3], d[3];
  int nx;
  const double2 *Nu[3];
  const double2 *Ng[3];
  const double *Nv[3];
  const double *Nw[3];
  double e[3];
  int nn;
  double gamma;
  double sub;
  double2 *u0_out, *g_out;
  double *v_out, *w_out;
  double2 *u1, *u2;
};

// Per-mode LU factors of the steady systems at one gamma (computed once per
// handle): [row][value][mode] per (basis, parity) block, [row][value] for the
// two parities of the mean system.
struct FactorSet {
  double *blk[2][2];
  double *mean[2];
};

// Field counts double under MHD: X/Y hold (u, omega, b, j), Hz/Hx (H, E).
struct Workspace {
  void *X;   // [6|12][n0][m]   nodal u, omega (, b, j)
  void *Y;   // [6|12][n0][my][kzs]
  void *Hz;  // [3|6][n0][my][kzs]
  void *Hx;  // [3|6][n0][m]
};

// Launchers (return hipError_t of the launch).  One kernel each, in step order.
template <typename T>
hipError_t launch_fft_y(const Geo &g, const NodalMats<T> &nm, Workspace &ws, const FFTPlan &py,
                        bool inverse, bool mhd, hipStream_t s);
template <typename T>
hipError_t launch_rotational(const Geo &g, const NodalMats<T> &nm, Workspace &ws,
                             const FFTPlan &pz, bool mhd, hipStream_t s);
// MHD: b = curl A and j = curl b at the nodes (X fields 6..11) from the three
// vector-potential components cstride apart; induction rows NA = W AG E.
template <typename T>
hipError_t launch_nodal_mhd_mfma(const Geo &g, const MhdMats<T> &mm, const double2 *A, size_t cstride,
                                 Workspace &ws, hipStream_t s);
template <typename T>
hipError_t launch_rhs_a_mfma(const Geo &g, const MhdMats<T> &mm, Workspace &ws, double2 *NA, size_t cstride,
                             hipStream_t s);
// Insulating walls: the A_y/A_z update.  rhs_c = h z_c + eta [(T_i, T_j'') zL_c
// - k^2 h zL_c] + sum_l e_l N_l,c with z = sum c_l x_l, zL = sum d_l x_l, rotated
// to P = cy Ay + cz Az, Q = cz Ay - cy Az and solved with the bordered radial
// systems (last row of each parity replaced by the vacuum-matching row), then
// rotated back.  Components y, z sit at 1 and 2 cstrides from each pointer.
struct BorderedArgs {
  const double2 *x[3];
  double c[3], d[3];
  int nx;
  const double2 *n[3];
  double e[3];
  int nn;
  double gamma;
  double2 *out;
  size_t cstride;
};
hipError_t launch_solve_bordered(const Geo &g, const BorderedArgs &a, hipStream_t s);
// u1, u2 of the solved state (u0_out, g_out, v_out, w_out of a) into a.u1, a.u2;
// Rt = R^T [K][ldr] (K: nb rows zero padded past round4(nb); ldr >= nd, a
// multiple of 16).
hipError_t launch_reconstruct(const Geo &g, const double *Rt, int ldr, const SolveArgs &a, hipStream_t s);
// Per-row data of the reverse pass's on-the-fly transposed solves.
size_t solve_scratch_doubles(const Geo &g);
// Factored steady solves: factor storage (basis -1: the mean system), the
// one-time factorisation at gamma, and the per-step solve (SBDF3 update).
size_t factor_doubles(const Geo &g, const StructMats &sm, int basis, int par);
// dir_only / ncomp > 0: Dirichlet blocks alone (no mean), for ncomp fields
// whose g-side pointers are cstride apart.
hipError_t launch_factor(const Geo &g, const StructMats &sm, double gamma, FactorSet fs, bool dir_only,
                         hipStream_t s);
hipError_t launch_solve_factored(const Geo &g, const StructMats &sm, const FactorSet &fs, const SolveArgs &a,
                                 int ncomp, size_t cstride, hipStream_t s);

// Diagnostics partial sums: out[0] = sum wF |u|^2 (perturbation), out[1] = sum
// wF |div u|^2, both over the full spectrum (Hermitian weights applied), and
// out[2] = sum_x wF Re u1_00(x) x, out[3] = sum_x wF Re du1_00/dx(x).
// out[f] = max_b blockmax[b][f] over nblocks rows of nf maxima (one block).
hipError_t launch_reduce_blockmax(const double *bm, int nblocks, int nf, double *out, hipStream_t s);
// The Hermitian half spectrum [rows][ny][kzs] written out in the full
// [rows][ny][nz] layout (conjugate mirror above the zero Nyquist plane).
hipError_t launch_expand_half(const double2 *half, double2 *full, int rows, int ny, int nz, int kzs, hipStream_t s);
// diff |= 1 when the n complex values at a and b differ in any bit.
hipError_t launch_bits_differ(const double2 *a, const double2 *b, size_t n, unsigned *diff, hipStream_t s);
hipError_t launch_diag(const Geo &g, const double *VB, const double *VD, const double *VB1,
                       const double *VD1, const double *fejer, const double *xnode,
                       const double2 *u0, const double2 *u1, const double2 *u2, double *out,
                       hipStream_t s);

// MHD diagnostics partial sums: out[0] = sum wF |b|^2, out[1] = sum wF |div b|^2
// (b = curl A), out[2..4] = sum_x wF b_c(x) of the mean mode.  VDX [n0][kx],
// VDY and VEY [n0][kyz] are row-major nodal matrices (values, x-derivatives)
// of the Ax and Ay/Az bases.
hipError_t launch_diag_mhd(const Geo &g, const double *VDX, int kx, const double *VDY, const double *VEY, int kyz,
                           const double *fejer, const double2 *A, size_t cstride, double *out, hipStream_t s);
// Insulating walls: per-block maxima of the vacuum-matching wall residuals of
// b = curl A (blockmax holds insul_residual_blocks(g) doubles).
int insul_residual_blocks(const Geo &g);
hipError_t launch_insul_residual(const Geo &g, const double2 *A, size_t cstride, double *blockmax, hipStream_t s);

// Largest |value| of nf physical fields on the unpadded grid (g: the
// unpadded geometry; Y holds their y-transformed lines, [nf][n0][ny][kzs]):
// per-block maxima in blockmax[absmax_blocks][nf].  Fields 0..2 are the
// velocity (field 1 gets the base flow U x), 3..5 the magnetic field (+ B0).
// With phys set, the fluctuation fields themselves also go there, real FP64
// [nf][n0][ny][nz] on the unpadded grid.
template <typename T>
hipError_t launch_absmax(const Geo &g, const T *xnode, const T *twz, const FFTPlan &pz, int nf, const void *Y,
                         double *blockmax, hipStream_t s, double *phys = nullptr);
int absmax_blocks(const Geo &g, int nf, int elem_bytes);
bool absmax_fits(const Geo &g, const FFTPlan &pz, int nf, int elem_bytes);

// Largest butterfly count per thread any FFT launch needs (host-side check
// against the kernels' register budgets).
bool fft_fits(const Geo &g, const FFTPlan &py, const FFTPlan &pz, int elem_bytes, bool mhd);

}  // namespace pcf
