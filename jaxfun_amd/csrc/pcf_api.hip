// C ABI of libpcfdns.so (declared in include/pcf_dns.h).
//
// Owns the device state of one plane-Couette run: the current spectral
// fields, a 3-slot ring of primary solutions (u0, g, v00, w00) and a 4-slot
// ring of nonlinear rows (Nu, Ng, Nv00, Nw00), so the multistep history shift
// of the reference (examples/channelflow_kmm.py:951-964,1051-1064) is a
// pointer rotation, never a copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/pcf_dns.h"
#include "pcf_kernels.hpp"
#include "pcf_setup.hpp"

namespace {

thread_local std::string g_err;

int32_t fail(int32_t code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define PCF_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(PCF_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

const double IMEX_A[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};   // imex_rk.py:103
const double IMEX_B[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};      // imex_rk.py:104

std::string validate(const pcf_config *c) {
  if (!c) return "config is NULL";
  if (c->abi_version != PCF_ABI_VERSION) return "abi_version mismatch";
  if (c->n0 < 9 || c->n0 > 130) return "n0 must be in [9, 130]";
  if (c->ny < 4 || c->nz < 4 || (c->ny % 2) || (c->nz % 2)) return "ny, nz must be even and >= 4";
  if (c->lx[0] != -1.0 || c->lx[1] != 1.0) return "wall-normal domain must be (-1, 1)";
  if (!(c->ly > 0) || !(c->lz > 0)) return "periodic lengths must be positive";
  if (!(c->dt > 0)) return "dt must be positive";
  if (!(c->nu > 0)) return "nu must be positive";
  if (c->precision != PCF_PREC_FP64 && c->precision != PCF_PREC_MIXED) return "unknown precision";
  if (c->time_integrator != PCF_TI_SBDF3) return "only SBDF3 is built";
  if (c->nonlinear_form != PCF_NL_ROTATIONAL) return "only the rotational nonlinear form is built";
  return "";
}

int padded(int n, double f) { return std::max(n, (int)std::floor(f * n)); }

bool make_plan(int n, pcf::FFTPlan &p) {
  p.n = n;
  p.nst = 0;
  int m = n;
  while (m % 4 == 0) p.radix[p.nst++] = 4, m /= 4;
  while (m % 2 == 0) p.radix[p.nst++] = 2, m /= 2;
  while (m % 3 == 0) p.radix[p.nst++] = 3, m /= 3;
  return m == 1 && p.nst <= 12;
}

double dfactor(double L) {
  const double tp = 2.0 * M_PI;
  return std::fabs(L - tp) > 1e-12 ? tp / L : 1.0;
}

struct StageTimer {
  hipEvent_t ev[7][2];
};

template <typename T>
struct DevMats {
  T *VBt = nullptr, *VDt = nullptr, *VD1t = nullptr, *A1t = nullptr, *A0t = nullptr,
    *AGt = nullptr, *xnode = nullptr, *twy = nullptr, *twz = nullptr;
  pcf::NodalMats<T> view{};
};

}  // namespace

struct pcf_handle {
  pcf_config cfg{};
  pcf::Geo geo{};
  pcf::Operators ops;
  pcf::FFTPlan py{}, pz{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;
  std::vector<void *> allocs;
  // state
  double2 *u0[3]{}, *g[3]{};
  double *v[3]{}, *w[3]{};
  double2 *u1 = nullptr, *u2 = nullptr;
  double2 *Nu[4]{}, *Ng[4]{};
  double *Nv[4]{}, *Nw[4]{};
  int so[3] = {0, 1, 2};     // solution ring: cur, old, older
  int no[4] = {0, 1, 2, 3};  // nonlinear ring: free0, old, older, free1
  double history_steps = 0, have_old = 0, previous_dt = 0;
  pcf::Workspace ws{};
  DevMats<double> md;
  DevMats<float> mf;
  pcf::StructMats sm{};
  double *Rt = nullptr;       // parity-compressed reconstruction rows, [nd][ldr]
  int ldr = 0;
  double *scratch = nullptr;  // per-row data of the structured solves
  double *dVB = nullptr, *dVD = nullptr, *dVB1 = nullptr, *dVD1 = nullptr, *dfejer = nullptr,
         *dxn = nullptr, *ddiag = nullptr;
  std::map<int, hipGraphExec_t> graphs;
  StageTimer *timer = nullptr;  // non-null while pcf_profile_steady runs

  template <typename T>
  int32_t alloc(T **p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return fail(PCF_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(q);
    *p = static_cast<T *>(q);
    e = hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return fail(PCF_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    return PCF_OK;
  }
  template <typename T>
  int32_t upload(T **p, const std::vector<T> &h) {
    int32_t r = alloc(p, h.size());
    if (r) return r;
    PCF_HIP(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return PCF_OK;
  }
};

namespace {

template <typename T>
int32_t build_nodal(pcf_handle *h, DevMats<T> &dm, int xc, int ic) {
  const pcf::Operators &o = h->ops;
  const int n0 = o.n0, nb = o.nb, nd = o.nd;
  const int ldx = (n0 + xc - 1) / xc * xc;
  const int ldi = (nd + ic - 1) / ic * ic;
  auto transpose_pad = [&](const pcf::Mat &M, int ld) {  // M (rows x cols) -> [cols][ld]
    std::vector<T> t((size_t)M.cols * ld, (T)0);
    for (int i = 0; i < M.rows; ++i)
      for (int j = 0; j < M.cols; ++j) t[(size_t)j * ld + i] = (T)M(i, j);
    return t;
  };
  int32_t r;
  if ((r = h->upload(&dm.VBt, transpose_pad(o.VB, ldx)))) return r;
  if ((r = h->upload(&dm.VDt, transpose_pad(o.VD, ldx)))) return r;
  if ((r = h->upload(&dm.VD1t, transpose_pad(o.VD1, ldx)))) return r;
  if ((r = h->upload(&dm.A1t, transpose_pad(o.A1, ldi)))) return r;
  if ((r = h->upload(&dm.A0t, transpose_pad(o.A0, ldi)))) return r;
  if ((r = h->upload(&dm.AGt, transpose_pad(o.AG, ldi)))) return r;
  std::vector<T> xn(n0);
  for (int j = 0; j < n0; ++j) xn[j] = (T)o.x[j];
  if ((r = h->upload(&dm.xnode, xn))) return r;
  auto twiddles = [](int n) {
    std::vector<T> t(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
      long double a = 2.0L * 3.141592653589793238462643383279502884L * k / n;
      t[2 * k] = (T)cosl(a);
      t[2 * k + 1] = (T)(-sinl(a));
    }
    return t;
  };
  if ((r = h->upload(&dm.twy, twiddles(h->geo.my)))) return r;
  if ((r = h->upload(&dm.twz, twiddles(h->geo.mz)))) return r;
  dm.view = pcf::NodalMats<T>{dm.VBt, dm.VDt, dm.VD1t, ldx, dm.A1t, dm.A0t, dm.AGt, ldi,
                              dm.xnode, dm.twy, dm.twz};
  (void)nb;
  return PCF_OK;
}

int32_t build_struct_mats(pcf_handle *h) {
  const pcf::Operators &o = h->ops;
  int32_t r;
  int npmax = 0;
  for (int b = 0; b < 2; ++b)
    for (int p = 0; p < 2; ++p) {
      const pcf::Operators::Struct &st = o.st[b][p];
      pcf::StructSys &ss = h->sm.sys[b][p];
      ss.n = st.n;
      npmax = std::max(npmax, st.n);
      double *d = nullptr;
      if ((r = h->upload(&d, st.b0))) return r;
      ss.b0 = d;
      if ((r = h->upload(&d, st.b2))) return r;
      ss.b2 = d;
      if ((r = h->upload(&d, st.b4))) return r;
      ss.b4 = d;
      if ((r = h->upload(&d, st.U))) return r;
      ss.U = d;
      if ((r = h->upload(&d, st.V))) return r;
      ss.V = d;
    }
  h->geo.npmax = npmax;
  // parity-compressed reconstruction rows: Rc[i][k] = R[i][2k + 1 - (i & 1)]
  h->ldr = (o.nb + 2) / 2;
  std::vector<double> Rc((size_t)o.nd * h->ldr, 0.0);
  for (int i = 0; i < o.nd; ++i) {
    const int j0 = 1 - (i & 1);
    for (int k = 0; j0 + 2 * k < o.nb; ++k) Rc[(size_t)i * h->ldr + k] = o.R(i, j0 + 2 * k);
  }
  if ((r = h->upload(&h->Rt, Rc))) return r;
  if ((r = h->alloc(&h->scratch, pcf::solve_scratch_doubles(h->geo)))) return r;
  return PCF_OK;
}

// Stage names, in launch order, for per-kernel timing.
constexpr int NSTAGE = 7;
const char *const STAGE_NAMES[NSTAGE] = {"nodal_backward", "fft_y_inverse", "rotational",
                                         "fft_y_forward", "forward_rhs", "solve", "reconstruct"};

// Record a timing event at a stage edge; accumulate the previous step's
// interval when a stage starts again.
inline hipError_t mark(pcf_handle *h, int stage, int edge) {
  StageTimer *t = h->timer;
  if (!t) return hipSuccess;
  return hipEventRecord(t->ev[stage][edge], h->stream);
}

template <typename T>
hipError_t nonlinear_stages(pcf_handle *h, const pcf::NodalMats<T> &nm, const double2 *u0, int slot) {
  const pcf::Geo &g = h->geo;
  hipError_t e;
  if ((e = mark(h, 0, 0))) return e;
  if ((e = pcf::launch_nodal<T>(g, nm, u0, h->u1, h->u2, h->ws, h->stream))) return e;
  if ((e = mark(h, 0, 1))) return e;
  if ((e = mark(h, 1, 0))) return e;
  if ((e = pcf::launch_fft_y<T>(g, nm, h->ws, h->py, true, h->stream))) return e;
  if ((e = mark(h, 1, 1))) return e;
  if ((e = mark(h, 2, 0))) return e;
  if ((e = pcf::launch_rotational<T>(g, nm, h->ws, h->pz, h->stream))) return e;
  if ((e = mark(h, 2, 1))) return e;
  if ((e = mark(h, 3, 0))) return e;
  if ((e = pcf::launch_fft_y<T>(g, nm, h->ws, h->py, false, h->stream))) return e;
  if ((e = mark(h, 3, 1))) return e;
  if ((e = mark(h, 4, 0))) return e;
  if ((e = pcf::launch_rhs<T>(g, nm, h->ws, h->Nu[slot], h->Ng[slot], h->Nv[slot], h->Nw[slot],
                              h->stream)))
    return e;
  return mark(h, 4, 1);
}

// Nonlinear rows of the state whose wall-normal velocity is u0 (with the
// handle's u1/u2) into nonlinear slot `slot`.
int32_t nonlinear_into(pcf_handle *h, const double2 *u0, int slot) {
  hipError_t e = (h->cfg.precision == PCF_PREC_FP64)
                     ? nonlinear_stages<double>(h, h->md.view, u0, slot)
                     : nonlinear_stages<float>(h, h->mf.view, u0, slot);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return PCF_OK;
}

// Structured solve + reconstruction of one implicit update.
hipError_t implicit_update(pcf_handle *h, const pcf::SolveArgs &a, bool steady) {
  hipError_t e;
  if ((e = mark(h, 5, 0))) return e;
  if ((e = pcf::launch_solve(h->geo, h->sm, a, h->scratch, steady, h->stream))) return e;
  if ((e = mark(h, 5, 1))) return e;
  if ((e = mark(h, 6, 0))) return e;
  if ((e = pcf::launch_reconstruct(h->geo, h->Rt, h->ldr, a, h->stream))) return e;
  return mark(h, 6, 1);
}

pcf::SolveArgs base_args(pcf_handle *h, int out_slot) {
  pcf::SolveArgs a{};
  a.u0_out = h->u0[out_slot];
  a.g_out = h->g[out_slot];
  a.v_out = h->v[out_slot];
  a.w_out = h->w[out_slot];
  a.u1 = h->u1;
  a.u2 = h->u2;
  return a;
}

void set_x(pcf::SolveArgs &a, int l, pcf_handle *h, int slot, double c, double d) {
  a.u0[l] = h->u0[slot];
  a.g[l] = h->g[slot];
  a.v[l] = h->v[slot];
  a.w[l] = h->w[slot];
  a.c[l] = c;
  a.d[l] = d;
}

void set_n(pcf::SolveArgs &a, int l, pcf_handle *h, int slot, double e) {
  a.Nu[l] = h->Nu[slot];
  a.Ng[l] = h->Ng[slot];
  a.Nv[l] = h->Nv[slot];
  a.Nw[l] = h->Nw[slot];
  a.e[l] = e;
}

void rotate(pcf_handle *h) {
  int s0 = h->so[0], s1 = h->so[1], s2 = h->so[2];
  h->so[0] = s2;
  h->so[1] = s0;
  h->so[2] = s1;
  int n0 = h->no[0], n1 = h->no[1], n2 = h->no[2], n3 = h->no[3];
  h->no[0] = n3;
  h->no[1] = n0;
  h->no[2] = n1;
  h->no[3] = n2;
}

// One steady SBDF3 step (sbdf3.py:44-62; channelflow_kmm.py:1016-1064).
int32_t steady_step(pcf_handle *h) {
  const double dt = h->cfg.dt;
  const int A = h->no[0];
  int32_t r = nonlinear_into(h, h->u0[h->so[0]], A);
  if (r) return r;
  pcf::SolveArgs a = base_args(h, h->so[2]);
  set_x(a, 0, h, h->so[0], 18.0 / 11.0, 0.0);
  set_x(a, 1, h, h->so[1], -9.0 / 11.0, 0.0);
  set_x(a, 2, h, h->so[2], 2.0 / 11.0, 0.0);
  a.nx = 3;
  set_n(a, 0, h, A, dt * 18.0 / 11.0);
  set_n(a, 1, h, h->no[1], -dt * 18.0 / 11.0);
  set_n(a, 2, h, h->no[2], dt * 6.0 / 11.0);
  a.nn = 3;
  a.gamma = dt * 6.0 / 11.0;
  hipError_t e = implicit_update(h, a, true);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("solve launch: ") + hipGetErrorString(e));
  rotate(h);
  h->history_steps = 2.0;
  h->have_old = 1.0;
  h->previous_dt = dt;
  return PCF_OK;
}

// One IMEXRK3 startup step (channelflow_kmm.py:872-935 inside the SBDF3
// wrapper :937-1014).  The stage-0 nonlinear rows are the step's current
// nonlinear rows, which the history shift keeps.
int32_t startup_step(pcf_handle *h) {
  const double dt = h->cfg.dt;
  const int A = h->no[0], B = h->no[3], C = h->no[2];
  const int cur = h->so[0], out = h->so[2];
  int32_t r = nonlinear_into(h, h->u0[cur], A);
  if (r) return r;
  const int nslot[3] = {A, B, C};
  for (int rk = 0; rk < 3; ++rk) {
    const double gamma = (IMEX_A[rk] + IMEX_B[rk]) * dt / 2.0;
    const int xs = (rk == 0) ? cur : out;
    if (rk > 0) {
      r = nonlinear_into(h, h->u0[out], nslot[rk]);
      if (r) return r;
    }
    pcf::SolveArgs a = base_args(h, out);
    set_x(a, 0, h, xs, 1.0, gamma);
    a.nx = 1;
    set_n(a, 0, h, nslot[rk], dt * IMEX_A[rk]);
    a.nn = 1;
    if (rk > 0) {
      set_n(a, 1, h, nslot[rk - 1], dt * IMEX_B[rk]);
      a.nn = 2;
    }
    a.gamma = gamma;
    hipError_t e = implicit_update(h, a, false);
    if (e != hipSuccess) return fail(PCF_E_HIP, std::string("solve launch: ") + hipGetErrorString(e));
  }
  rotate(h);
  h->history_steps = std::min(2.0, h->history_steps + 1.0);
  h->have_old = 1.0;
  h->previous_dt = dt;
  return PCF_OK;
}

// Field id -> (device pointer, is_complex, rows).
struct FieldRef {
  void *ptr;
  bool complex_;
  int rows;
};

bool field_ref(pcf_handle *h, int f, FieldRef &fr) {
  const int nb = h->geo.nb, nd = h->geo.nd;
  auto sol = [&](int lvl, int comp) -> FieldRef {
    int s = h->so[lvl];
    switch (comp) {
      case 0: return {h->u0[s], true, nb};
      case 1: return {h->g[s], true, nd};
      case 2: return {h->v[s], false, nd};
      default: return {h->w[s], false, nd};
    }
  };
  auto nl = [&](int lvl, int comp) -> FieldRef {
    int s = h->no[lvl];
    switch (comp) {
      case 0: return {h->Nu[s], true, nb};
      case 1: return {h->Ng[s], true, nd};
      case 2: return {h->Nv[s], false, nd};
      default: return {h->Nw[s], false, nd};
    }
  };
  switch (f) {
    case PCF_F_U0: fr = sol(0, 0); return true;
    case PCF_F_G: fr = sol(0, 1); return true;
    case PCF_F_U1: fr = {h->u1, true, nd}; return true;
    case PCF_F_U2: fr = {h->u2, true, nd}; return true;
    default: break;
  }
  if (f >= PCF_F_NL_OLD_U && f <= PCF_F_NL_OLD_W) { fr = nl(1, f - PCF_F_NL_OLD_U); return true; }
  if (f >= PCF_F_NL_OLDER_U && f <= PCF_F_NL_OLDER_W) { fr = nl(2, f - PCF_F_NL_OLDER_U); return true; }
  if (f >= PCF_F_SOL_OLD_U && f <= PCF_F_SOL_OLD_W) { fr = sol(1, f - PCF_F_SOL_OLD_U); return true; }
  if (f >= PCF_F_SOL_OLDER_U && f <= PCF_F_SOL_OLDER_W) { fr = sol(2, f - PCF_F_SOL_OLDER_U); return true; }
  return false;
}

int field_rows(const pcf_config *c, int f) {
  const int nb = c->n0 - 4, nd = c->n0 - 2;
  switch (f) {
    case PCF_F_U0:
    case PCF_F_NL_OLD_U:
    case PCF_F_NL_OLDER_U:
    case PCF_F_SOL_OLD_U:
    case PCF_F_SOL_OLDER_U: return nb;
    default: return nd;
  }
}

bool field_is_vector(int f) {
  return f == PCF_F_NL_OLD_V || f == PCF_F_NL_OLD_W || f == PCF_F_NL_OLDER_V ||
         f == PCF_F_NL_OLDER_W || f == PCF_F_SOL_OLD_V || f == PCF_F_SOL_OLD_W ||
         f == PCF_F_SOL_OLDER_V || f == PCF_F_SOL_OLDER_W;
}

}  // namespace

extern "C" {

int32_t pcf_abi_version(void) { return PCF_ABI_VERSION; }

const char *pcf_last_error(void) { return g_err.c_str(); }

int64_t pcf_field_size(const pcf_config *cfg, int32_t field) {
  std::string v = validate(cfg);
  if (!v.empty()) return fail(PCF_E_INVALID, v);
  if (field < 0 || field >= PCF_F_COUNT) return fail(PCF_E_INVALID, "unknown field");
  const int rows = field_rows(cfg, field);
  if (field_is_vector(field)) return rows;
  return (int64_t)rows * cfg->ny * cfg->nz;
}

int32_t pcf_setup_matrix(const pcf_config *cfg, int32_t which, double *out, int32_t *rows,
                         int32_t *cols) {
  std::string v = validate(cfg);
  if (!v.empty()) return fail(PCF_E_INVALID, v);
  pcf::Operators ops;
  std::string err = pcf::build_operators(cfg->n0, ops);
  if (!err.empty()) return fail(PCF_E_INVALID, err);
  const pcf::Mat *M = pcf::operator_by_index(ops, which);
  if (!M) return fail(PCF_E_INVALID, "unknown matrix index");
  if (rows) *rows = M->rows;
  if (cols) *cols = M->cols;
  if (out) std::memcpy(out, M->a.data(), M->a.size() * sizeof(double));
  return PCF_OK;
}

int32_t pcf_create(const pcf_config *cfg, pcf_handle **out) {
  if (!out) return fail(PCF_E_INVALID, "out is NULL");
  *out = nullptr;
  std::string v = validate(cfg);
  if (!v.empty()) return fail(PCF_E_INVALID, v);
  pcf_handle *h = new pcf_handle();
  h->cfg = *cfg;
  auto bail = [&](int32_t code) {
    pcf_destroy(h);
    return code;
  };
  std::string err = pcf::build_operators(cfg->n0, h->ops);
  if (!err.empty()) return bail(fail(PCF_E_INVALID, err));
  pcf::Geo &g = h->geo;
  g.n0 = cfg->n0;
  g.nb = cfg->n0 - 4;
  g.nd = cfg->n0 - 2;
  g.ny = cfg->ny;
  g.nz = cfg->nz;
  g.kzs = cfg->nz / 2;
  g.my = padded(cfg->ny, cfg->padding_y);
  g.mz = padded(cfg->nz, cfg->padding_z);
  g.m = g.ny * g.kzs;
  g.df1 = dfactor(cfg->ly);
  g.df2 = dfactor(cfg->lz);
  g.uwall = cfg->u_wall;
  g.W = (2.0 * M_PI / g.df1) * (2.0 * M_PI / g.df2);
  g.nu = cfg->nu;
  if (g.my % 2) return bail(fail(PCF_E_INVALID, "padded streamwise grid must be even"));
  if (g.mz < 2 * g.kzs) return bail(fail(PCF_E_INVALID, "spanwise padding must not truncate"));
  if (!make_plan(g.my, h->py) || !make_plan(g.mz, h->pz))
    return bail(fail(PCF_E_INVALID, "padded grid sizes must factor into 2, 3, 4"));
  if (!pcf::fft_fits(g, h->py, h->pz, cfg->precision == PCF_PREC_FP64 ? 16 : 8))
    return bail(fail(PCF_E_INVALID, "FFT sizes exceed the kernels' register budget"));

  h->device = cfg->device;
  if (hipSetDevice(cfg->device) != hipSuccess) return bail(fail(PCF_E_HIP, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(PCF_E_HIP, "hipStreamCreate failed"));
  h->own_stream = true;

  int32_t r;
  const size_t M = g.m;
  for (int s = 0; s < 3; ++s) {
    if ((r = h->alloc(&h->u0[s], (size_t)g.nb * M))) return bail(r);
    if ((r = h->alloc(&h->g[s], (size_t)g.nd * M))) return bail(r);
    if ((r = h->alloc(&h->v[s], (size_t)g.nd))) return bail(r);
    if ((r = h->alloc(&h->w[s], (size_t)g.nd))) return bail(r);
  }
  for (int s = 0; s < 4; ++s) {
    if ((r = h->alloc(&h->Nu[s], (size_t)g.nb * M))) return bail(r);
    if ((r = h->alloc(&h->Ng[s], (size_t)g.nd * M))) return bail(r);
    if ((r = h->alloc(&h->Nv[s], (size_t)g.nd))) return bail(r);
    if ((r = h->alloc(&h->Nw[s], (size_t)g.nd))) return bail(r);
  }
  if ((r = h->alloc(&h->u1, (size_t)g.nd * M))) return bail(r);
  if ((r = h->alloc(&h->u2, (size_t)g.nd * M))) return bail(r);
  const size_t ve = (cfg->precision == PCF_PREC_FP64) ? sizeof(double2) : sizeof(float2);
  unsigned char *wsp = nullptr;
  const size_t nX = 6 * (size_t)g.n0 * M, nY = 6 * (size_t)g.n0 * g.my * g.kzs;
  const size_t nHz = 3 * (size_t)g.n0 * g.my * g.kzs, nHx = 3 * (size_t)g.n0 * M;
  if ((r = h->alloc(&wsp, (nX + nY + nHz + nHx) * ve))) return bail(r);
  h->ws.X = wsp;
  h->ws.Y = wsp + nX * ve;
  h->ws.Hz = wsp + (nX + nY) * ve;
  h->ws.Hx = wsp + (nX + nY + nHz) * ve;

  if (cfg->precision == PCF_PREC_FP64) {
    if ((r = build_nodal<double>(h, h->md, 8, 8))) return bail(r);
  } else {
    if ((r = build_nodal<float>(h, h->mf, 8, 8))) return bail(r);
  }
  if ((r = build_struct_mats(h))) return bail(r);
  {
    hipError_t e = pcf::kernels_init(h->geo);
    if (e != hipSuccess)
      return bail(fail(PCF_E_UNSUPPORTED, std::string("kernel LDS limits: ") + hipGetErrorString(e)));
  }
  if ((r = h->upload(&h->dVB, h->ops.VB.a))) return bail(r);
  if ((r = h->upload(&h->dVD, h->ops.VD.a))) return bail(r);
  if ((r = h->upload(&h->dVB1, h->ops.VB1.a))) return bail(r);
  if ((r = h->upload(&h->dVD1, h->ops.VD1.a))) return bail(r);
  if ((r = h->upload(&h->dfejer, h->ops.fejer))) return bail(r);
  if ((r = h->upload(&h->dxn, h->ops.x))) return bail(r);
  if ((r = h->alloc(&h->ddiag, 4))) return bail(r);
  *out = h;
  return PCF_OK;
}

int32_t pcf_destroy(pcf_handle *h) {
  if (!h) return PCF_OK;
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PCF_OK;
}

int32_t pcf_set_stream(pcf_handle *h, void *stream) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  if (h->own_stream && h->stream) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamDestroy(h->stream);
  }
  h->own_stream = false;
  h->stream = static_cast<hipStream_t>(stream);
  if (!h->stream) {
    PCF_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->own_stream = true;
  }
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  return PCF_OK;
}

int32_t pcf_synchronize(pcf_handle *h) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  PCF_HIP(hipStreamSynchronize(h->stream));
  return PCF_OK;
}

int32_t pcf_set_field(pcf_handle *h, int32_t field, const double *host) {
  if (!h || !host) return fail(PCF_E_INVALID, "null argument");
  FieldRef fr;
  if (!field_ref(h, field, fr)) return fail(PCF_E_INVALID, "unknown field");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipStreamSynchronize(h->stream));
  if (!fr.complex_) {
    PCF_HIP(hipMemcpy(fr.ptr, host, (size_t)fr.rows * sizeof(double), hipMemcpyHostToDevice));
    return PCF_OK;
  }
  // full (rows, ny, nz) -> half [rows][ny][kzs]: keep the first nz/2 spanwise modes
  PCF_HIP(hipMemcpy2D(fr.ptr, (size_t)g.kzs * sizeof(double2), host,
                      (size_t)g.nz * sizeof(double2), (size_t)g.kzs * sizeof(double2),
                      (size_t)fr.rows * g.ny, hipMemcpyHostToDevice));
  if (field == PCF_F_U1 || field == PCF_F_U2) {
    // the mean-mode vector of the current level is Re u[:, 0, 0]
    std::vector<double> vec(g.nd);
    const size_t plane = (size_t)g.ny * g.nz * 2;
    for (int i = 0; i < g.nd; ++i) vec[i] = host[i * plane];
    double *dst = (field == PCF_F_U1) ? h->v[h->so[0]] : h->w[h->so[0]];
    PCF_HIP(hipMemcpy(dst, vec.data(), g.nd * sizeof(double), hipMemcpyHostToDevice));
  }
  return PCF_OK;
}

int32_t pcf_get_field(pcf_handle *h, int32_t field, double *host) {
  if (!h || !host) return fail(PCF_E_INVALID, "null argument");
  FieldRef fr;
  if (!field_ref(h, field, fr)) return fail(PCF_E_INVALID, "unknown field");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipStreamSynchronize(h->stream));
  if (!fr.complex_) {
    PCF_HIP(hipMemcpy(host, fr.ptr, (size_t)fr.rows * sizeof(double), hipMemcpyDeviceToHost));
    return PCF_OK;
  }
  std::vector<double2> half((size_t)fr.rows * g.m);
  PCF_HIP(hipMemcpy(half.data(), fr.ptr, half.size() * sizeof(double2), hipMemcpyDeviceToHost));
  // expand the Hermitian half spectrum: c(-ky, -kz) = conj(c(ky, kz)); Nyquist plane = 0
  for (int n = 0; n < fr.rows; ++n)
    for (int iy = 0; iy < g.ny; ++iy) {
      double *row = host + ((size_t)n * g.ny + iy) * g.nz * 2;
      const double2 *hr = half.data() + ((size_t)n * g.ny + iy) * g.kzs;
      const double2 *mr = half.data() + ((size_t)n * g.ny + (g.ny - iy) % g.ny) * g.kzs;
      for (int kz = 0; kz < g.nz; ++kz) {
        double re = 0, im = 0;
        if (kz < g.kzs) {
          re = hr[kz].x;
          im = hr[kz].y;
        } else if (kz > g.kzs) {
          re = mr[g.nz - kz].x;
          im = -mr[g.nz - kz].y;
        }
        row[2 * kz] = re;
        row[2 * kz + 1] = im;
      }
    }
  return PCF_OK;
}

int32_t pcf_set_history(pcf_handle *h, double history_steps, double have_old, double previous_dt) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  h->history_steps = history_steps;
  h->have_old = have_old;
  h->previous_dt = previous_dt;
  return PCF_OK;
}

int32_t pcf_get_history(pcf_handle *h, double *history_steps, double *have_old, double *previous_dt) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  if (history_steps) *history_steps = h->history_steps;
  if (have_old) *have_old = h->have_old;
  if (previous_dt) *previous_dt = h->previous_dt;
  return PCF_OK;
}

int32_t pcf_step(pcf_handle *h, int32_t nsteps) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  if (nsteps < 0) return fail(PCF_E_INVALID, "nsteps must be >= 0");
  PCF_HIP(hipSetDevice(h->device));
  for (int s = 0; s < nsteps; ++s) {
    int32_t r = (h->history_steps < 2.0) ? startup_step(h) : steady_step(h);
    if (r) return r;
  }
  return PCF_OK;
}

int32_t pcf_step_steady(pcf_handle *h, int32_t nsteps) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  if (h->history_steps < 2.0) return fail(PCF_E_INVALID, "history not bootstrapped (history_steps < 2)");
  PCF_HIP(hipSetDevice(h->device));
  for (int s = 0; s < nsteps; ++s) {
    int32_t r = steady_step(h);
    if (r) return r;
  }
  return PCF_OK;
}

int32_t pcf_step_steady_graph(pcf_handle *h, int32_t nsteps, int32_t replays) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  if (h->history_steps < 2.0) return fail(PCF_E_INVALID, "history not bootstrapped (history_steps < 2)");
  if (nsteps <= 0 || nsteps % 12) return fail(PCF_E_INVALID, "graph length must be a positive multiple of 12 steps");
  PCF_HIP(hipSetDevice(h->device));
  auto it = h->graphs.find(nsteps);
  if (it == h->graphs.end()) {
    hipGraph_t graph = nullptr;
    PCF_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int32_t r = PCF_OK;
    for (int s = 0; s < nsteps && !r; ++s) r = steady_step(h);
    hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (r) return r;
    if (e != hipSuccess) return fail(PCF_E_HIP, std::string("capture: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(PCF_E_HIP, std::string("instantiate: ") + hipGetErrorString(e));
    it = h->graphs.emplace(nsteps, exec).first;
    // capture recorded the launches without running them and advanced the
    // ring nsteps times (a whole number of ring periods: back at the start)
  }
  for (int rp = 0; rp < replays; ++rp) PCF_HIP(hipGraphLaunch(it->second, h->stream));
  return PCF_OK;
}

int32_t pcf_diagnostics(pcf_handle *h, double *out) {
  if (!h || !out) return fail(PCF_E_INVALID, "null argument");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipSetDevice(h->device));
  PCF_HIP(hipMemsetAsync(h->ddiag, 0, 4 * sizeof(double), h->stream));
  hipError_t e = pcf::launch_diag(g, h->dVB, h->dVD, h->dVB1, h->dVD1, h->dfejer, h->dxn,
                                  h->u0[h->so[0]], h->u1, h->u2, h->ddiag, h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("diag launch: ") + hipGetErrorString(e));
  double s[4];
  std::vector<double2> col(g.nd);
  PCF_HIP(hipMemcpyAsync(s, h->ddiag, sizeof(s), hipMemcpyDeviceToHost, h->stream));
  PCF_HIP(hipMemcpy2DAsync(col.data(), sizeof(double2), h->u1, (size_t)g.m * sizeof(double2),
                           sizeof(double2), g.nd, hipMemcpyDeviceToHost, h->stream));
  PCF_HIP(hipStreamSynchronize(h->stream));
  const double Ly = h->cfg.ly, Lz = h->cfg.lz, U = h->cfg.u_wall;
  double sx2 = 0, sw = 0;
  for (int j = 0; j < g.n0; ++j) {
    sx2 += h->ops.fejer[j] * h->ops.x[j] * h->ops.x[j];
    sw += h->ops.fejer[j];
  }
  const double epert = Ly * Lz * s[0];
  double top = U, bot = U;
  for (int i = 0; i < g.nd; ++i) {
    top += h->ops.dphi_top[i] * col[i].x;
    bot += h->ops.dphi_bot[i] * col[i].x;
  }
  out[0] = epert;
  out[1] = epert + Ly * Lz * (2.0 * U * s[2] + U * U * sx2);
  out[2] = std::sqrt(Ly * Lz * s[1]);
  out[3] = U;
  out[4] = -U;
  out[5] = top;
  out[6] = bot;
  out[7] = (s[3] + U * sw) / 2.0;
  return PCF_OK;
}

int32_t pcf_stage_count(void) { return NSTAGE; }

const char *pcf_stage_name(int32_t i) { return (i >= 0 && i < NSTAGE) ? STAGE_NAMES[i] : ""; }

int32_t pcf_profile_steady(pcf_handle *h, int32_t nsteps, double *stage_ms) {
  if (!h || !stage_ms) return fail(PCF_E_INVALID, "null argument");
  if (h->history_steps < 2.0) return fail(PCF_E_INVALID, "history not bootstrapped (history_steps < 2)");
  if (nsteps <= 0) return fail(PCF_E_INVALID, "nsteps must be positive");
  PCF_HIP(hipSetDevice(h->device));
  StageTimer t{};
  for (int i = 0; i < NSTAGE; ++i)
    for (int j = 0; j < 2; ++j) PCF_HIP(hipEventCreate(&t.ev[i][j]));
  double acc[NSTAGE] = {0, 0, 0, 0, 0, 0, 0};
  int32_t r = PCF_OK;
  h->timer = &t;
  for (int s = 0; s < nsteps && !r; ++s) {
    r = steady_step(h);
    if (r) break;
    if (hipEventSynchronize(t.ev[NSTAGE - 1][1]) != hipSuccess) {
      r = fail(PCF_E_HIP, "event synchronize failed");
      break;
    }
    for (int i = 0; i < NSTAGE; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, t.ev[i][0], t.ev[i][1]) != hipSuccess) {
        r = fail(PCF_E_HIP, "event elapsed time failed");
        break;
      }
      acc[i] += ms;
    }
  }
  h->timer = nullptr;
  for (int i = 0; i < NSTAGE; ++i)
    for (int j = 0; j < 2; ++j) (void)hipEventDestroy(t.ev[i][j]);
  if (r) return r;
  for (int i = 0; i < NSTAGE; ++i) stage_ms[i] = acc[i] / nsteps;
  return PCF_OK;
}

int32_t pcf_kernels_per_step(pcf_handle *h) {
  (void)h;
  return NSTAGE;
}

}  // extern "C"
