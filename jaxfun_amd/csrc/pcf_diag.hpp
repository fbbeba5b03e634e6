// Kernels of the rotating-shear MHD cadence diagnostics (pcf_diag.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "pcf_kernels.hpp"

namespace pcf {

// One term of a nodal-spectral evaluation: f(m) * M F with M [rows][K]
// row-major, F [K][m] and the mode factor f = c times 1, i ky, i kz, ky^2,
// kz^2, ky kz or ky^2 + kz^2 (kf 0..6; the streamwise Nyquist wavenumber
// counts as zero, as in the step's contractions).
struct EvalTerm {
  const double *M = nullptr;
  const double2 *F = nullptr;
  int K = 0;
  int kf = 0;
  double c = 1.0;
};

enum EvalFactor { KF_ONE = 0, KF_IKY = 1, KF_IKZ = 2, KF_KYY = 3, KF_KZZ = 4, KF_KYZ = 5, KF_K2 = 6 };

// out[p][m] = sum of up to twelve terms, p < rows.
constexpr int EVAL_TERMS = 12;
struct EvalOut {
  double2 *dst = nullptr;
  int rows = 0;
  int nt = 0;
  EvalTerm t[EVAL_TERMS];
};

// Outputs evaluated by one launch (gridDim.z).
constexpr int EVAL_BATCH = 4;
struct EvalBatch {
  int n = 0;
  EvalOut o[EVAL_BATCH];
};

// Sums over the stored modes at one node (k_parseval); wt = 1 for kz = 0,
// 2 for the other stored columns.
enum ParsevalKind {
  PV_DOT = 0,   // wt Re(a conj b), with add0a / add0b added to mode 0 of a / b
  PV_K2 = 1,    // wt (ky^2 + kz^2) |a|^2
  PV_MEAN = 2,  // Re a at mode 0 (+ add0a)
  PV_KZ1 = 3,   // |a|^2 at the mode (ky = 0, kz = 1)
  PV_SUM2 = 4,  // wt |a + b|^2
};

struct ParsevalJob {
  const double2 *a = nullptr;
  const double2 *b = nullptr;
  int kind = 0;
  int pad = 0;
  double add0a = 0.0, add0b = 0.0;
};

// Slots of k_emf's per-row results.
enum EmfSlot {
  EMF_IDEAL2 = 0,  // weighted sums over interior rows: |ideal|^2, |resistive|^2, |total|^2
  EMF_RES2 = 1,
  EMF_TOT2 = 2,
  EMF_MEAN = 3,    // 9 weighted sums: ideal x,y,z, resistive x,y,z, total x,y,z
  EMF_BMAX = 12,   // max |b| (fluctuation) and |b + B0| over interior rows
  EMF_BTMAX = 13,
  EMF_WIDEAL = 14, // wall rows: max tangential |ideal|, |resistive|, |total|
  EMF_WRES = 15,
  EMF_WTOT = 16,
  EMF_WSUM1 = 17,  // wall rows: sums of the total field's y and z components
  EMF_WSUM2 = 18,
  EMF_WDIVA = 19,  // wall rows: max |div A|
  EMF_Q = 20,
};

// The n outputs of o, EVAL_BATCH per launch.
hipError_t launch_eval_terms(const Geo &g, const EvalOut *o, int n, hipStream_t s);
hipError_t launch_curl_b(const Geo &g, int P, const double2 *b, const double2 *db, double2 *J, hipStream_t s);
hipError_t launch_dft_y(const Geo &g, bool inverse, int planes, const double2 *tw, const double2 *in, double2 *out,
                        hipStream_t s);
hipError_t launch_dft_z_c2r(const Geo &g, int planes, const double2 *tw, const double2 *in, double *out,
                            hipStream_t s);
hipError_t launch_dft_z_r2c(const Geo &g, int planes, const double2 *tw, const double *in, double2 *out,
                            hipStream_t s);
hipError_t launch_emf(const Geo &g, int P, double eta, const double *xp, const double *wq, const double *phys,
                      double *E, double *part, hipStream_t s);
// mu_b = scale fejer_x b (fields 0..2 of X = (b, j) [6][n0][m]), mu_j = 0.
hipError_t launch_seed_b(const Geo &g, const double *fejer, double scale, double2 *X, hipStream_t s);
hipError_t launch_parseval(const Geo &g, const ParsevalJob *jobs, int njobs, double *part, hipStream_t s);

}  // namespace pcf
