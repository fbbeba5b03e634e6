// C ABI of the Taylor-Couette step (include/tc_dns.h).
//
// Reference: AxisymmetricTCDNSJax (examples/taylor_couette_dns_jax.py:217-1078).
// The device state keeps the Hermitian half of every field ([component][k][i],
// k < Nz / 2); the solution and nonlinear histories are rings of device
// buffers whose roles rotate on the host after each step (no copies).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/tc_dns.h"
#include "tc_kernels.hpp"
#include "tc_setup.hpp"

namespace {

thread_local std::string g_tc_err;

int32_t fail(int32_t code, const std::string &msg) {
  g_tc_err = msg;
  return code;
}

#define TC_HIP(call)                                                             \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return fail(TC_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

// ARS(4,4,3) tableau (src/jaxfun/integrators/imex_rk.py:183-198)
constexpr double ARS_A[5][5] = {{0, 0, 0, 0, 0},
                                {0, 0.5, 0, 0, 0},
                                {0, 1.0 / 6.0, 0.5, 0, 0},
                                {0, -0.5, 0.5, 0.5, 0},
                                {0, 1.5, -1.5, 0.5, 0.5}};
constexpr double ARS_B[5][5] = {{0, 0, 0, 0, 0},
                                {0.5, 0, 0, 0, 0},
                                {11.0 / 18.0, 1.0 / 18.0, 0, 0, 0},
                                {5.0 / 6.0, -5.0 / 6.0, 0.5, 0, 0},
                                {0.25, 1.75, 0.75, -1.75, 0}};

}  // namespace

struct tc_handle {
  tc_config cfg{};
  tc::Params prm{};
  tc::Ops ops{};
  tc::TcGeo geo{};
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  double *dEv = nullptr, *dEd = nullptr, *dPrT = nullptr, *dinvr = nullptr, *dkz = nullptr, *dMassT = nullptr,
         *dmq = nullptr, *dmq_odd = nullptr, *dDtc = nullptr, *dEs = nullptr, *dEsd = nullptr, *drstd = nullptr, *dfejer = nullptr,
         *dDrc = nullptr, *dDzc = nullptr, *dout = nullptr, *dQbr = nullptr, *dQbt = nullptr, *dQbtd = nullptr,
         *dQbz = nullptr, *dQbrr = nullptr, *dQbzr = nullptr;
  double2 *dtwz = nullptr, *dtwn = nullptr, *dtwt = nullptr, *dtwnt = nullptr, *dAinvT = nullptr, *dAinv0T = nullptr;
  double2 *dLinT = nullptr, *dLexpT = nullptr, *ddv = nullptr;
  double2 *S[4]{}, *N[3]{}, *P = nullptr, *G = nullptr, *XS[3]{}, *NS[3]{};
  int su = 0, s1 = 1, s2 = 2, sf = 3, nn1 = 0, nn2 = 1, nf = 2;
  double have_old = 0.0, history_steps = 0.0;
  std::map<long, hipGraphExec_t> graphs;
  size_t fsize = 0;  // elements of one 3-component field set

  template <class T>
  int32_t alloc(T **p, size_t count) {
    void *q = nullptr;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return fail(TC_E_HIP, "hipMalloc failed");
    allocs.push_back(q);
    *p = reinterpret_cast<T *>(q);
    return TC_OK;
  }
  template <class T>
  int32_t upload(T **p, const std::vector<double> &src) {
    const size_t count = src.size() * sizeof(double) / sizeof(T);
    int32_t r = alloc(p, count ? count : 1);
    if (r) return r;
    if (!src.empty() && hipMemcpy(*p, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(TC_E_HIP, "hipMemcpy failed");
    return TC_OK;
  }
};

namespace {

tc::Params params_of(const tc_config &c) {
  tc::Params p;
  p.nr = c.nr;
  p.nz = c.nz;
  p.r1 = c.r1;
  p.r2 = c.r2;
  // CircularCouette (examples/taylor_couette_linear_jax.py:37-48)
  const double d2 = c.r2 * c.r2 - c.r1 * c.r1;
  p.a = (c.omega2 * c.r2 * c.r2 - c.omega1 * c.r1 * c.r1) / d2;
  p.b = (c.omega1 - c.omega2) * c.r1 * c.r1 * c.r2 * c.r2 / d2;
  p.nu = c.nu;
  p.lz = c.lz;
  p.dt = c.dt;
  p.dealias = c.dealias;
  p.integrator = c.time_integrator;
  p.ntheta = c.ntheta;
  p.mhd = c.mhd;
  p.b0 = c.b0;
  p.eta = c.eta_mag;
  return p;
}

// (Re)builds the operators and uploads the tables that depend on dt.
int32_t setup_ops(tc_handle *h) {
  std::string err = tc::build_ops(h->prm, h->ops);
  if (!err.empty()) return fail(TC_E_INVALID, err);
  const tc::Ops &o = h->ops;
  int32_t r;
  if ((r = h->upload(&h->dMassT, o.MassT)) || (r = h->upload(&h->dLinT, o.LinT)) ||
      (r = h->upload(&h->dLexpT, o.LexpT)) || (r = h->upload(&h->dAinvT, o.AinvT)) ||
      (r = h->upload(&h->dAinv0T, o.Ainv0T)))
    return r;  // (complex tables as double2)
  h->geo.MassT = h->dMassT;
  h->geo.LinT = h->dLinT;
  h->geo.LexpT = h->dLexpT;
  return TC_OK;
}

std::vector<double> twiddles(int n) {
  std::vector<double> t(2 * (size_t)n);
  for (int q = 0; q < n; ++q) {
    const long double a = 2.0L * 3.141592653589793238462643383279502884L * q / n;
    t[2 * q] = (double)cosl(a);
    t[2 * q + 1] = (double)sinl(a);
  }
  return t;
}

int32_t check_field(const tc_handle *h, int32_t field) {
  if (!h) return fail(TC_E_INVALID, "null handle");
  if (field < 0 || field >= TC_F_COUNT) return fail(TC_E_INVALID, "no such field");
  return TC_OK;
}

// Device pointer of one component of a group (nm x n complex).
double2 *component_ptr(tc_handle *h, int32_t group, int32_t comp) {
  const size_t cs = (size_t)h->ops.nm * h->ops.n;
  switch (group) {
    case TC_G_X:
      return h->S[h->su] + comp * cs;
    case TC_G_P:
      return h->P;
    case TC_G_NL_OLD:
      return h->N[h->nn1] + comp * cs;
    case TC_G_NL_OLDER:
      return h->N[h->nn2] + comp * cs;
    case TC_G_SOL_OLD:
      return h->S[h->s1] + comp * cs;
    default:
      return h->S[h->s2] + comp * cs;
  }
}

// Device pointer of a field id (one component, kh x n complex).
double2 *field_ptr(tc_handle *h, int32_t f) {
  const size_t cs = (size_t)h->ops.nm * h->ops.n;
  if (f <= TC_F_UZ) return h->S[h->su] + f * cs;
  if (f == TC_F_P) return h->P;
  if (f < TC_F_NL_OLDER) return h->N[h->nn1] + (f - TC_F_NL_OLD) * cs;
  if (f < TC_F_SOL_OLD) return h->N[h->nn2] + (f - TC_F_NL_OLDER) * cs;
  if (f < TC_F_SOL_OLDER) return h->S[h->s1] + (f - TC_F_SOL_OLD) * cs;
  return h->S[h->s2] + (f - TC_F_SOL_OLDER) * cs;
}

struct Term {
  int op;
  const double2 *x;
  double c;
};

hipError_t solve(tc_handle *h, const double2 *u_nl, std::initializer_list<Term> terms, double cN, double2 *nout,
                 const double2 *Ainv, double2 *uout) {
  hipError_t e = hipSuccess;
  if (u_nl) e = tc::launch_nonlinear(h->geo, u_nl, h->G, h->stream);
  if (e != hipSuccess) return e;
  tc::TcRhs a;
  for (const Term &t : terms) {
    a.op[a.nt] = t.op;
    a.x[a.nt] = t.x;
    a.c[a.nt] = t.c;
    ++a.nt;
  }
  a.G = u_nl ? h->G : nullptr;
  a.cN = cN;
  a.nout = nout;
  a.AinvT = Ainv;
  a.uout = uout;
  a.pout = h->P;
  return tc::launch_rhs_solve(h->geo, a, h->stream);
}

void rotate_history(tc_handle *h) {
  const int os2 = h->s2, onn2 = h->nn2;
  h->s2 = h->s1;
  h->s1 = h->su;
  h->su = h->sf;
  h->sf = os2;
  h->nn2 = h->nn1;
  h->nn1 = h->nf;
  h->nf = onn2;
}

// One ARS(4,4,3) step from S[su] into S[sf] (examples/taylor_couette_dns_jax.py:797-831
// with src/jaxfun/integrators/coupled.py:13-44); N(u_n) goes to n0.
hipError_t ars_step(tc_handle *h, const double2 *Ainv, double2 *n0) {
  using namespace tc;
  const double2 *un = h->S[h->su];
  const double2 *NS[4] = {n0, h->NS[0], h->NS[1], h->NS[2]};
  hipError_t e;
  if ((e = solve(h, un, {{TC_OP_MASS, un, 1.0}}, -ARS_B[1][0], n0, Ainv, h->XS[0]))) return e;
  if ((e = solve(h, h->XS[0], {{TC_OP_MASS, un, 1.0}, {TC_OP_ID, NS[0], -ARS_B[2][0]}, {TC_OP_LIN, h->XS[0], ARS_A[2][1]}},
                 -ARS_B[2][1], h->NS[0], Ainv, h->XS[1])))
    return e;
  if ((e = solve(h, h->XS[1],
                 {{TC_OP_MASS, un, 1.0}, {TC_OP_ID, NS[0], -ARS_B[3][0]}, {TC_OP_ID, NS[1], -ARS_B[3][1]},
                  {TC_OP_LIN, h->XS[0], ARS_A[3][1]}, {TC_OP_LIN, h->XS[1], ARS_A[3][2]}},
                 -ARS_B[3][2], h->NS[1], Ainv, h->XS[2])))
    return e;
  return solve(h, h->XS[2],
               {{TC_OP_MASS, un, 1.0}, {TC_OP_ID, NS[0], -ARS_B[4][0]}, {TC_OP_ID, NS[1], -ARS_B[4][1]},
                {TC_OP_ID, NS[2], -ARS_B[4][2]}, {TC_OP_LIN, h->XS[0], ARS_A[4][1]},
                {TC_OP_LIN, h->XS[1], ARS_A[4][2]}, {TC_OP_LIN, h->XS[2], ARS_A[4][3]}},
               -ARS_B[4][3], nullptr, Ainv, h->S[h->sf]);
}

// Steady SBDF3 step (:926-943 with src/jaxfun/integrators/sbdf3.py:16-43).
hipError_t sbdf3_steady(tc_handle *h) {
  using namespace tc;
  const double2 *u = h->S[h->su], *u1 = h->S[h->s1], *u2 = h->S[h->s2];
  hipError_t e = solve(h, u,
                       {{TC_OP_MASS, u, 18.0 / 11.0}, {TC_OP_MASS, u1, -9.0 / 11.0}, {TC_OP_MASS, u2, 2.0 / 11.0},
                        {TC_OP_ID, h->N[h->nn1], 18.0 / 11.0}, {TC_OP_ID, h->N[h->nn2], -6.0 / 11.0}},
                       -18.0 / 11.0, h->N[h->nf], h->dAinvT, h->S[h->sf]);
  if (e != hipSuccess) return e;
  rotate_history(h);
  return hipSuccess;
}

// CNAB2 step (:758-795): explicit L/2 and the AB2 (first step: Euler) nonlinear rows.
hipError_t cnab2_step(tc_handle *h) {
  using namespace tc;
  const double2 *u = h->S[h->su];
  hipError_t e;
  if (h->have_old != 0.0)
    e = solve(h, u, {{TC_OP_LEXP, u, 1.0}, {TC_OP_ID, h->N[h->nn1], 0.5}}, -1.5, h->N[h->nf], h->dAinvT,
              h->S[h->sf]);
  else
    e = solve(h, u, {{TC_OP_LEXP, u, 1.0}}, -1.0, h->N[h->nf], h->dAinvT, h->S[h->sf]);
  if (e != hipSuccess) return e;
  std::swap(h->su, h->sf);
  std::swap(h->nn1, h->nf);
  h->have_old = 1.0;
  return hipSuccess;
}

int32_t one_step(tc_handle *h) {
  using namespace tc;
  hipError_t e = hipSuccess;
  switch (h->cfg.time_integrator) {
    case TC_TI_CNAB2:
      e = cnab2_step(h);
      break;
    case TC_TI_IMEXRK3:  // :797-850: history untouched
      e = ars_step(h, h->dAinvT, h->NS[2]);
      if (e == hipSuccess) std::swap(h->su, h->sf);
      break;
    default:  // SBDF3 (:908-946)
      if (h->history_steps < 2.0) {
        e = ars_step(h, h->dAinv0T, h->N[h->nf]);
        if (e == hipSuccess) rotate_history(h);
      } else {
        e = sbdf3_steady(h);
      }
      if (e == hipSuccess) {
        h->history_steps = std::min(2.0, h->history_steps + 1.0);
        h->have_old = 1.0;
      }
  }
  if (e != hipSuccess) return fail(TC_E_HIP, std::string("step launch: ") + hipGetErrorString(e));
  return TC_OK;
}

int32_t check_component(const tc_handle *h, int32_t group, int32_t comp) {
  if (!h) return fail(TC_E_INVALID, "null handle");
  if (group < TC_G_X || group > TC_G_SOL_OLDER) return fail(TC_E_INVALID, "no such group");
  if (comp < 0 || comp >= (group == TC_G_P ? 1 : h->ops.nf)) return fail(TC_E_INVALID, "no such component");
  return TC_OK;
}

// The Hermitian half of a full ([Ntheta,] Nz, n) host array: rows k < kh of
// every azimuthal mode, in the library's mode order q = im * kh + k.
int32_t upload_half(tc_handle *h, double2 *dst, const double *host) {
  const int kh = h->ops.kh, n = h->ops.n, nth = h->ops.nth, nz = h->cfg.nz;
  TC_HIP(hipSetDevice(h->cfg.device));
  std::vector<double> half(2 * (size_t)h->ops.nm * n);
  for (int im = 0; im < nth; ++im)
    std::memcpy(&half[2 * (size_t)im * kh * n], host + 2 * (size_t)im * nz * n, 2 * (size_t)kh * n * sizeof(double));
  TC_HIP(hipMemcpyAsync(dst, half.data(), half.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  TC_HIP(hipStreamSynchronize(h->stream));
  return TC_OK;
}

// The full array from the stored half: f(-m, -k) = conj f(m, k); the axial
// Nyquist row stays zero.
int32_t download_full(tc_handle *h, const double2 *src_dev, double *host) {
  const int kh = h->ops.kh, n = h->ops.n, nth = h->ops.nth, nz = h->cfg.nz;
  TC_HIP(hipSetDevice(h->cfg.device));
  std::vector<double> half(2 * (size_t)h->ops.nm * n);
  TC_HIP(hipMemcpyAsync(half.data(), src_dev, half.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  TC_HIP(hipStreamSynchronize(h->stream));
  std::memset(host, 0, 2 * (size_t)nth * nz * n * sizeof(double));
  for (int im = 0; im < nth; ++im)
    for (int k = 0; k < kh; ++k) {
      const double *src = &half[2 * ((size_t)im * kh + k) * n];
      std::memcpy(host + 2 * ((size_t)im * nz + k) * n, src, 2 * (size_t)n * sizeof(double));
      if (k == 0) continue;
      double *dst = host + 2 * ((size_t)((nth - im) % nth) * nz + (nz - k)) * n;
      for (int i = 0; i < n; ++i) {
        dst[2 * i] = src[2 * i];
        dst[2 * i + 1] = -src[2 * i + 1];
      }
    }
  return TC_OK;
}

}  // namespace

extern "C" {

const char *tc_last_error(void) { return g_tc_err.c_str(); }

int32_t tc_create(const tc_config *cfg, tc_handle **out) {
  if (!cfg || !out) return fail(TC_E_INVALID, "null argument");
  *out = nullptr;
  if (cfg->abi_version != TC_ABI_VERSION) return fail(TC_E_INVALID, "ABI version mismatch");
  if (cfg->time_integrator < TC_TI_CNAB2 || cfg->time_integrator > TC_TI_IMEXRK3)
    return fail(TC_E_INVALID, "time_integrator must be one of {'CNAB2', 'SBDF3', 'IMEXRK3'}");
  if (!(cfg->dealias == 1.0 || cfg->dealias > 1.0)) return fail(TC_E_INVALID, "dealias must be >= 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TC_E_HIP, "no HIP device");
  if (cfg->device < 0 || cfg->device >= ndev) return fail(TC_E_INVALID, "no such device");
  if (hipSetDevice(cfg->device) != hipSuccess) return fail(TC_E_HIP, "hipSetDevice failed");
  auto *h = new tc_handle();
  h->cfg = *cfg;
  h->prm = params_of(*cfg);
  int32_t r = setup_ops(h);
  if (r) {
    tc_destroy(h);
    return r;
  }
  const tc::Ops &o = h->ops;
  tc::TcGeo sz;
  sz.nf = o.nf;
  sz.n = o.n;
  sz.kh = o.kh;
  sz.nth = o.nth;
  sz.nm = o.nm;
  sz.mr = o.mr;
  sz.mz = o.mz;
  sz.mth = o.mth;
  sz.nr = cfg->nr;
  int lds_max = 65536;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, cfg->device) != hipSuccess)
    lds_max = 65536;
  const size_t need = std::max({tc::nonlinear_lds(sz), tc::rhs_lds(sz), tc::diag_lds(sz)});
  if (need > (size_t)lds_max) {
    tc_destroy(h);
    return fail(TC_E_UNSUPPORTED, "grid too large for the per-block LDS tables");
  }
  if (need > 65536 && tc::allow_lds(need) != hipSuccess) {
    tc_destroy(h);
    return fail(TC_E_HIP, "could not raise the kernels' LDS limit");
  }
  std::vector<double> invr(o.mr), rstd(o.r_std);
  for (int j = 0; j < o.mr; ++j) invr[j] = 1.0 / o.r_pad[j];
  if ((r = h->upload(&h->dEv, o.Ev)) || (r = h->upload(&h->dEd, o.Ed)) || (r = h->upload(&h->dPrT, o.PrT)) ||
      (r = h->upload(&h->dinvr, invr)) || (r = h->upload(&h->dkz, o.kz)) || (r = h->upload(&h->dEs, o.Es)) ||
      (r = h->upload(&h->dEsd, o.Esd)) || (r = h->upload(&h->drstd, rstd)) || (r = h->upload(&h->dfejer, o.fejer)) ||
      (r = h->upload(&h->dDrc, o.Drc)) || (r = h->upload(&h->dDzc, o.Dzc)) ||
      (r = h->upload(&h->dtwz, twiddles(o.mz))) || (r = h->upload(&h->dtwn, twiddles(cfg->nz))) ||
      (r = h->upload(&h->dtwt, twiddles(o.mth))) || (r = h->upload(&h->dtwnt, twiddles(o.nth))) ||
      (r = h->upload(&h->dmq, o.mq)) || (r = h->upload(&h->dmq_odd, o.mq_odd)) || (r = h->upload(&h->dDtc, o.Dtc)) ||
      (r = h->upload(&h->dQbr, o.QbrT)) || (r = h->upload(&h->dQbt, o.QbtT)) || (r = h->upload(&h->dQbtd, o.QbtdT)) ||
      (r = h->upload(&h->dQbz, o.QbzT)) || (r = h->upload(&h->dQbrr, o.QbrrT)) || (r = h->upload(&h->dQbzr, o.QbzrT)) ||
      (r = h->alloc(&h->ddv, 2 * (size_t)o.nm * cfg->nr)) ||
      (r = h->alloc(&h->dout, 8))) {
    tc_destroy(h);
    return r;
  }
  h->fsize = (size_t)o.nf * o.nm * o.n;
  for (int i = 0; i < 4 && !r; ++i) r = h->alloc(&h->S[i], h->fsize);
  for (int i = 0; i < 3 && !r; ++i) r = h->alloc(&h->N[i], h->fsize);
  for (int i = 0; i < 3 && !r; ++i) r = h->alloc(&h->XS[i], h->fsize);
  for (int i = 0; i < 3 && !r; ++i) r = h->alloc(&h->NS[i], h->fsize);
  if (!r) r = h->alloc(&h->P, (size_t)o.nm * o.n);
  if (!r) r = h->alloc(&h->G, (o.nf == 6 ? 6 : 3) * (size_t)o.nm * o.mr);
  if (r) {
    tc_destroy(h);
    return r;
  }
  for (int i = 0; i < 4; ++i) (void)hipMemset(h->S[i], 0, h->fsize * sizeof(double2));
  for (int i = 0; i < 3; ++i) (void)hipMemset(h->N[i], 0, h->fsize * sizeof(double2));
  (void)hipMemset(h->P, 0, (size_t)o.nm * o.n * sizeof(double2));
  if (hipDeviceSynchronize() != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    tc_destroy(h);
    return fail(TC_E_HIP, "device setup failed");
  }
  tc::TcGeo &g = h->geo;
  g.nf = o.nf;
  g.QbrT = h->dQbr;
  g.QbtT = h->dQbt;
  g.QbtdT = h->dQbtd;
  g.QbzT = h->dQbz;
  g.QbrrT = h->dQbrr;
  g.QbzrT = h->dQbzr;
  g.n = o.n;
  g.kh = o.kh;
  g.nth = o.nth;
  g.nm = o.nm;
  g.mr = o.mr;
  g.mz = o.mz;
  g.mth = o.mth;
  g.fz = o.fz;
  g.mq = h->dmq;
  g.mq_odd = h->dmq_odd;
  g.Dtc = h->dDtc;
  g.twt = h->dtwt;
  g.twn_t = h->dtwnt;
  g.dv = h->ddv;
  g.nr = cfg->nr;
  g.nz = cfg->nz;
  g.lz = cfg->lz;
  g.Ev = h->dEv;
  g.Ed = h->dEd;
  g.PrT = h->dPrT;
  g.invr = h->dinvr;
  g.kz = h->dkz;
  g.twz = h->dtwz;
  g.twn = h->dtwn;
  g.Es = h->dEs;
  g.Esd = h->dEsd;
  g.rstd = h->drstd;
  g.fejer = h->dfejer;
  g.Drc = h->dDrc;
  g.Dzc = h->dDzc;
  *out = h;
  return TC_OK;
}

int32_t tc_destroy(tc_handle *h) {
  if (!h) return TC_OK;
  (void)hipSetDevice(h->cfg.device);
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  if (h->stream) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamDestroy(h->stream);
  }
  for (void *p : h->allocs) (void)hipFree(p);
  delete h;
  return TC_OK;
}

int32_t tc_set_field(tc_handle *h, int32_t field, const double *host) {
  if (int32_t r = check_field(h, field)) return r;
  if (!host) return fail(TC_E_INVALID, "null argument");
  return upload_half(h, field_ptr(h, field), host);
}

int32_t tc_get_field(tc_handle *h, int32_t field, double *host) {
  if (int32_t r = check_field(h, field)) return r;
  if (!host) return fail(TC_E_INVALID, "null argument");
  return download_full(h, field_ptr(h, field), host);
}

int32_t tc_set_component(tc_handle *h, int32_t group, int32_t comp, const double *host) {
  if (int32_t r = check_component(h, group, comp)) return r;
  if (!host) return fail(TC_E_INVALID, "null argument");
  return upload_half(h, component_ptr(h, group, comp), host);
}

int32_t tc_get_component(tc_handle *h, int32_t group, int32_t comp, double *host) {
  if (int32_t r = check_component(h, group, comp)) return r;
  if (!host) return fail(TC_E_INVALID, "null argument");
  return download_full(h, component_ptr(h, group, comp), host);
}

int32_t tc_set_scalars(tc_handle *h, double have_old, double history_steps) {
  if (!h) return fail(TC_E_INVALID, "null handle");
  h->have_old = have_old;
  h->history_steps = history_steps;
  return TC_OK;
}

int32_t tc_get_scalars(tc_handle *h, double *have_old, double *history_steps) {
  if (!h || !have_old || !history_steps) return fail(TC_E_INVALID, "null argument");
  *have_old = h->have_old;
  *history_steps = h->history_steps;
  return TC_OK;
}

int32_t tc_step(tc_handle *h, int32_t nsteps) {
  if (!h) return fail(TC_E_INVALID, "null handle");
  if (nsteps < 0) return fail(TC_E_INVALID, "steps must be non-negative");
  TC_HIP(hipSetDevice(h->cfg.device));
  for (int s = 0; s < nsteps; ++s)
    if (int32_t r = one_step(h)) return r;
  TC_HIP(hipStreamSynchronize(h->stream));
  return TC_OK;
}

int32_t tc_step_graph(tc_handle *h, int32_t nsteps) {
  if (!h) return fail(TC_E_INVALID, "null handle");
  if (nsteps < 0) return fail(TC_E_INVALID, "steps must be non-negative");
  TC_HIP(hipSetDevice(h->cfg.device));
  constexpr int BLOCK = 12;  // the rings (4 and 3 slots) return to their roles after 12 steps
  int done = 0;
  while (done < nsteps) {
    // steady SBDF3 steps, and CNAB2 steps once the extrapolation has its old rows
    const bool sbdf3 = h->cfg.time_integrator == TC_TI_SBDF3 && h->history_steps >= 2.0;
    const bool cnab2 = h->cfg.time_integrator == TC_TI_CNAB2 && h->have_old != 0.0;
    const bool steady = sbdf3 || cnab2;
    if (!steady || nsteps - done < BLOCK) {
      if (int32_t r = one_step(h)) return r;
      ++done;
      continue;
    }
    const long key =
        (((((((long)h->su * 4 + h->s1) * 4 + h->s2) * 4 + h->sf) * 3 + h->nn1) * 3 + h->nn2) * 3 + h->nf) * 2 + cnab2;
    auto it = h->graphs.find(key);
    if (it == h->graphs.end()) {
      hipGraph_t graph = nullptr;
      TC_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      hipError_t e = hipSuccess;
      for (int i = 0; i < BLOCK && e == hipSuccess; ++i) e = sbdf3 ? sbdf3_steady(h) : cnab2_step(h);
      hipError_t ec = hipStreamEndCapture(h->stream, &graph);
      if (e != hipSuccess || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(TC_E_HIP, "graph capture failed");
      }
      hipGraphExec_t exec = nullptr;
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(TC_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
      it = h->graphs.emplace(key, exec).first;
    } else {
      for (int i = 0; i < BLOCK; ++i) {  // the roles the graph leaves (its steps rotate them BLOCK times)
        if (sbdf3) {
          rotate_history(h);
        } else {
          std::swap(h->su, h->sf);
          std::swap(h->nn1, h->nf);
        }
      }
    }
    TC_HIP(hipGraphLaunch(it->second, h->stream));
    done += BLOCK;
  }
  TC_HIP(hipStreamSynchronize(h->stream));
  return TC_OK;
}

int32_t tc_diagnostics(tc_handle *h, double *out) {
  if (!h || !out) return fail(TC_E_INVALID, "null argument");
  TC_HIP(hipSetDevice(h->cfg.device));
  hipError_t e = tc::launch_diag(h->geo, h->S[h->su], h->dout, h->stream);
  if (e != hipSuccess) return fail(TC_E_HIP, std::string("diagnostics launch: ") + hipGetErrorString(e));
  const int count = h->ops.nf == 6 ? TC_MHD_DIAG_COUNT : TC_DIAG_COUNT;
  TC_HIP(hipMemcpyAsync(out, h->dout, count * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  TC_HIP(hipStreamSynchronize(h->stream));
  return TC_OK;
}

int32_t tc_linear_pencil(const tc_config *cfg, int32_t m, double kz, double *out_L, double *out_M) {
  if (!cfg || !out_L || !out_M) return fail(TC_E_INVALID, "null argument");
  if (cfg->abi_version != TC_ABI_VERSION) return fail(TC_E_INVALID, "ABI version mismatch");
  std::vector<double> L, M;
  std::string err = tc::linear_pencil(params_of(*cfg), m, kz, L, M);
  if (!err.empty()) return fail(TC_E_INVALID, err);
  std::memcpy(out_L, L.data(), L.size() * sizeof(double));
  std::memcpy(out_M, M.data(), M.size() * sizeof(double));
  return TC_OK;
}

int32_t tc_mri_pencil(const tc_config *cfg, int32_t m, double kz, double *out_L, double *out_M) {
  if (!cfg || !out_L || !out_M) return fail(TC_E_INVALID, "null argument");
  if (cfg->abi_version != TC_ABI_VERSION) return fail(TC_E_INVALID, "ABI version mismatch");
  std::vector<double> L, M;
  std::string err = tc::mri_pencil(params_of(*cfg), m, kz, L, M);
  if (!err.empty()) return fail(TC_E_INVALID, err);
  std::memcpy(out_L, L.data(), L.size() * sizeof(double));
  std::memcpy(out_M, M.data(), M.size() * sizeof(double));
  return TC_OK;
}

int32_t tc_set_dt(tc_handle *h, double dt) {
  if (!h) return fail(TC_E_INVALID, "null handle");
  if (!(dt > 0)) return fail(TC_E_INVALID, "dt must be positive");
  const double old = h->cfg.dt;
  if (std::fabs(dt - old) <= 8.0 * 2.220446049250313e-16 * std::fabs(dt)) return TC_OK;
  if (h->cfg.time_integrator == TC_TI_SBDF3)
    return fail(TC_E_FIXED_DT,
                "SBDF3 currently supports fixed dt only; start a fresh solver/history for a different timestep");
  TC_HIP(hipSetDevice(h->cfg.device));
  TC_HIP(hipStreamSynchronize(h->stream));
  h->cfg.dt = dt;
  h->prm.dt = dt;
  // captured graphs hold the old tables' addresses
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  return setup_ops(h);
}

}  // extern "C"
