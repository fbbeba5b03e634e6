// C ABI of libpcfdns.so (declared in include/pcf_dns.h).
//
// Owns the device state of one plane-Couette run: the current spectral
// fields, a 3-slot ring of primary solutions (u0, g, v00, w00) and a 4-slot
// ring of nonlinear rows (Nu, Ng, Nv00, Nw00), so the multistep history shift
// of the reference (examples/channelflow_kmm.py:951-964,1051-1064) is a
// pointer rotation, never a copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pcf_dns.h"
#include "pcf_adjoint.hpp"
#include "pcf_device.hpp"
#include "pcf_diag.hpp"
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
const double ARS_G = (2.0 - std::sqrt(2.0)) / 2.0;                // imex_rk.py:176
// gamma of IMEXRK3 stage rk: the implicit weight (a + b) dt / 2 (channelflow_kmm.py:872-935)
double stage_gamma(int rk, double dt) { return (IMEX_A[rk] + IMEX_B[rk]) * dt / 2.0; }

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
  if (c->time_integrator != PCF_TI_SBDF3 && c->time_integrator != PCF_TI_IMEXRK3 &&
      c->time_integrator != PCF_TI_IMEXRK222 && c->time_integrator != PCF_TI_CNAB2)
    return "unknown time integrator";
  if (c->mhd && c->time_integrator == PCF_TI_IMEXRK3) return "the MHD step is built for SBDF3, IMEXRK222 and CNAB2";
  if (c->nonlinear_form != PCF_NL_ROTATIONAL) return "only the rotational nonlinear form is built";
  if (c->mhd != 0 && c->mhd != 1) return "mhd must be 0 or 1";
  if (c->mhd && !(c->eta > 0)) return "eta must be positive";
  if (c->mhd && c->walls != PCF_WALLS_CONDUCTING && c->walls != PCF_WALLS_INSULATING) return "unknown wall type";
  if (c->mhd && !std::isfinite(c->omega)) return "omega must be finite";
  return "";
}

int padded(int n, double f) { return std::max(n, (int)std::floor(f * n)); }

bool make_plan(int n, pcf::FFTPlan &p) {
  // the device plans' radix choice (pcf::cplan): radix 8 only where it saves a pass
  const pcf::CPlan c = pcf::cplan(n);
  p.n = n;
  p.nst = c.nst;
  int m = n;
  for (int s = 0; s < c.nst; ++s) p.radix[s] = c.r[s], m /= c.r[s];
  return m == 1 && p.nst <= 12;
}

double dfactor(double L) {
  const double tp = 2.0 * M_PI;
  return std::fabs(L - tp) > 1e-12 ? tp / L : 1.0;
}

struct StageTimer {
  hipEvent_t ev[6][2];
};

template <typename T>
struct DevMats {
  T *VBt = nullptr, *VDt = nullptr, *VD1t = nullptr, *A1t = nullptr, *A0t = nullptr,
    *AGt = nullptr, *xnode = nullptr, *twy = nullptr, *twz = nullptr;
  T *VQ1t = nullptr, *VQ2t = nullptr, *VCt = nullptr, *VC1t = nullptr, *VC2t = nullptr, *WQt = nullptr;
  T *twy0 = nullptr, *twz0 = nullptr;  // twiddles of the unpadded grid (ny, nz)
  T *VDRt = nullptr, *VD1Rt = nullptr;
  pcf::NodalMats<T> view{};
  pcf::MhdMats<T> mview{};  // MHD contractions (conducting or insulating walls)
};

}  // namespace

// Block maxima of pcf_field_maxima_noalloc: a file-scope device array, one
// copy per GPU the library runs on, held until the process ends.  Host code
// never uses the array's name as an address: each device's address comes from
// hipGetSymbolAddress with that device current (scratch_on_device).
constexpr size_t NOALLOC_SCRATCH_DOUBLES = 65536;  // 512 KiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
double *g_scratch_ptr[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// The address of `scratch` on device dev (current), looked up once per device.
int32_t scratch_on_device(int dev, double **out) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(PCF_E_INVALID, "device ordinal out of range for the no-alloc scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(PCF_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return PCF_OK;
}
}  // namespace

struct pcf_handle {
  pcf_config cfg{};
  pcf::Geo geo{};
  pcf::Operators ops;
  pcf::FFTPlan py{}, pz{};
  pcf::Geo geo0{};            // the unpadded grid (my = ny, mz = nz) of the field maxima
  pcf::FFTPlan py0{}, pz0{};
  bool maxima_ok = false;     // the unpadded sizes have plans that fit the kernels
  double *dblockmax = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;
  std::vector<void *> allocs;
  // state
  double2 *u0[3]{}, *g[3]{};
  double *v[3]{}, *w[3]{};
  double2 *u1 = nullptr, *u2 = nullptr;
  unsigned char *state_block = nullptr;  // all of the above, contiguous
  size_t state_bytes = 0;
  double2 *Nu[4]{}, *Ng[4]{};
  double *Nv[4]{}, *Nw[4]{};
  // MHD: vector potential levels (three components [c][row][m], cstride =
  // nd * m apart) on the solution ring and their nonlinear rows on the
  // nonlinear ring; they rotate with the flow's rings (one coupled timeline,
  // examples/pcf_mhd_jax.py:217-248).
  bool mhd = false;
  bool insul = false;         // insulating walls: Ay, Az in TC (n0 rows)
  int arows = 0;              // rows of Ay, Az (nd or n0); Ax uses the first nd
  double2 *A[3]{}, *NA[4]{};
  size_t cstride = 0;         // arows * m
  double *dVC = nullptr, *dVC1 = nullptr, *dresid = nullptr;
  pcf::Geo geo_eta{};         // geo with nu = eta: the A systems are the g systems at eta
  pcf::FactorSet fA{};        // steady LU factors of the A systems (Dirichlet blocks)
  pcf::FactorSet fA_ars{};    // ARS(2,2,2) LU factors of the A systems
  pcf::FactorSet fA_stage[3]{};  // IMEXRK3 stage factors of the A systems
  pcf::FactorSet fA_cn{};     // CNAB2 factors of the A systems
  hipStream_t side = nullptr;  // insulating MHD: the vector-potential kernels beside the flow ones
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  double *ddiagm = nullptr;
  int so[3] = {0, 1, 2};     // solution ring: cur, old, older
  int no[4] = {0, 1, 2, 3};  // nonlinear ring: free0, old, older, free1
  double history_steps = 0, have_old = 0, previous_dt = 0;
  // u1/u2 of the current state.  A state loaded through pcf_set_field carries
  // its own u1, u2 (u12_given: the next nonlinear evaluation reads them, as the
  // reference's convection reads state.u).  A solve writes only (u0, g, v00,
  // w00); the next evaluation then forms u1, u2 inside its nodal contraction
  // (channelflow_kmm.py:855-870 folded into it), and the buffers are filled by
  // the reconstruction kernel only when something reads them (u12_valid).
  bool u12_given = true, u12_valid = true;
  // A loaded state whose u1, u2 equal, bit for bit, what the reconstruction
  // makes of its (u0, g, v00, w00) -- a state this library produced and
  // handed out -- is stepped as if it had never left the device, so that a
  // resume from it is exact; u12_checked records that the test was made.
  bool u12_checked = true;
  unsigned *dflag = nullptr;
  pcf::Workspace ws{};
  DevMats<double> md;
  DevMats<float> mf;
  pcf::StructMats sm{};
  // Device data of the rotating-shear cadence diagnostics
  // (pcf_mhd_extended_diagnostics), allocated by its first call.
  struct XDiag {
    bool ready = false;
    int P = 0, njobs = 0;
    pcf::DiagOps ops;
    double *EB0 = nullptr, *EB1 = nullptr, *ED0 = nullptr, *ED1 = nullptr, *ED2 = nullptr, *EC0 = nullptr,
           *EC1 = nullptr, *EC2 = nullptr, *Dn = nullptr, *I = nullptr, *xp = nullptr, *wq = nullptr;
    double2 *twy = nullptr, *twz = nullptr;
    double2 *NS = nullptr;   // [XD_COUNT][P][m] nodal-spectral fields
    double2 *G = nullptr;    // [10][P][ny][kzs] streamwise-inverse lines of the first ten
    double *phys = nullptr;  // [10][P][ny][nz] physical u, b, J, div A
    double *E = nullptr;     // [6][n0][ny][nz] ideal and resistive electric field
    double2 *EH = nullptr;   // [6][n0][ny][kzs] their spanwise forward transforms
    double2 *EC = nullptr;   // [6][n0][m] their TC coefficients at the nodes
    double2 *DIV = nullptr;  // [2][n0][m] divergence of the ideal and resistive parts
    pcf::ParsevalJob *jobs = nullptr;
    double *part = nullptr;  // [njobs][n0]
    double *emf = nullptr;   // [P][EMF_Q]
  } xd;
  pcf::FactorSet fsteady{};   // LU factors of the steady SBDF3 systems
  pcf::FactorSet fars{};      // LU factors of M - g dt L for the ARS(2,2,2) stages
  pcf::FactorSet fstage[3]{}; // LU factors at the three IMEXRK3 stage gammas
  pcf::FactorSet fcn{};       // LU factors of M - (dt/2) L (CNAB2)
  double *Rt = nullptr;       // R^T of the reconstruction, [round4(nb)][ldr]
  int ldr = 0;
  double *dVB = nullptr, *dVD = nullptr, *dVB1 = nullptr, *dVD1 = nullptr, *dfejer = nullptr,
         *dxn = nullptr, *ddiag = nullptr;
  std::map<uint64_t, hipGraphExec_t> graphs;  // (ring phase, steps) -> steady-step graph
  std::map<uint64_t, hipGraphExec_t> seg_graphs;  // steady runs of the reverse pass's forward sweeps
  // graph cache counters: lookups that found / captured a graph, graphs
  // dropped, and the number of times the caches were dropped (stream change)
  int64_t graph_hits = 0, graph_misses = 0, graph_evictions = 0, graph_generation = 0;
  StageTimer *timer = nullptr;  // non-null while pcf_profile_steady runs

  // Reverse mode, set up by the first pcf_vjp.
  struct Snap {
    char *dev = nullptr;
    char *devA = nullptr;  // MHD: the A and NA slots
    int so[3], no[4];
    double hs, ho, pd;
    bool given, valid;
  };
  struct Adj {
    bool ready = false;
    pcf::StructMats smt{};        // transposed, index-reversed parity blocks
    pcf::FactorSet fsteady_t{};   // their steady LU factors
    double *RcT = nullptr;        // parity-compressed R^T rows, [nb][ldrT]
    int ldrT = 0;
    double *QB = nullptr, *QD = nullptr;  // Fejer-weighted Gram matrices of the bases
    pcf::AdjMats<double> ad{};
    pcf::AdjMats<float> af{};
    pcf::Prim muX{}, rho[3]{}, rs[3]{}, gh[4]{}, nu{};
    pcf::Prim sig{};              // sum_j C_j rho_{k+j} of the current steady reverse step
    double2 *mu_u[3]{};
    double2 *cot_u[3]{}, *cot_g = nullptr;
    bool cot_set[4] = {false, false, false, false};
    int rr[3] = {0, 1, 2};        // rho ring: rho_k, rho_k+1, rho_k+2
    // The reverse sweep runs on its own stream with its own workspace and
    // solve scratch, so that the recompute of segment s-1 (handle stream,
    // tape set (s-1) & 1) overlaps the reverse of segment s (tape set s & 1).
    hipStream_t stream = nullptr;
    hipEvent_t ev_seed = nullptr, ev_rc[2] = {nullptr, nullptr}, ev_rev[2] = {nullptr, nullptr};
    pcf::Workspace ws{};
    double *scratch = nullptr;
    int per_set = 0;              // tape slots per set
    std::vector<void *> tape;     // padded y-line fields of each nonlinear evaluation
    std::vector<Snap> snaps;      // state every `every` steps, then the final state
    size_t snap_bytes = 0;
    bool have_grad = false;
    // MHD (conducting walls, FP64): cotangents of the vector potential, three
    // components cstride apart, run through the flow-shaped Dirichlet
    // kernels at nu = eta with zero / sink buffers in the other slots.
    bool mhd_ready = false;
    pcf::FactorSet fA_t{};        // steady factors of the transposed A systems (geo_eta)
    double2 *muA = nullptr, *nuA = nullptr, *sigA = nullptr, *cotA = nullptr;
    double2 *rhoA[3]{}, *rsA[3]{}, *ghA[4]{};
    bool cotA_set = false;
    pcf::Prim zero{}, sink{};     // zero (never written) and write-sink primary sets
    double *scratchA = nullptr;
    double *mEWAG = nullptr;      // W AG^T (n0 x nd): cotangent of the nodal EMF from mu_NA
    double *mVDT = nullptr, *mVD1T = nullptr, *mVQ1T = nullptr, *mVQ2T = nullptr;  // (nd x n0) transposes
    double *mMAT = nullptr, *mC2T = nullptr;  // W C0D^T, W C2D^T (nd x nd): A mass and stiffness transposed
  } adj;
  std::vector<void *> adj_allocs;  // freed on destroy or when the tape grows

  template <typename T>
  int32_t alloc(T **p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return fail(PCF_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(q);
    *p = static_cast<T *>(q);
    e = hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T));
    // hipMemset runs on the null stream and may still be in flight when it
    // returns; the handle's stream does not wait for the null stream, so wait
    // here before any kernel on it writes this memory (the setup factors).
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
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
  // M (rows x cols) -> [cols][ld], zero rows appended up to a multiple of 4 >= n0:
  // every matrix-core K loop (4 K rows per step, K <= n0) stays inside the
  // allocation without bounds checks
  const int kr = (n0 + 3) / 4 * 4;
  auto transpose_pad = [&](const pcf::Mat &M, int ld) {
    std::vector<T> t((size_t)std::max(M.cols, kr) * ld, (T)0);
    for (int i = 0; i < M.rows; ++i)
      for (int j = 0; j < M.cols; ++j) t[(size_t)j * ld + i] = (T)M(i, j);
    return t;
  };
  int32_t r;
  if ((r = h->upload(&dm.VBt, transpose_pad(o.VB, ldx)))) return r;
  if ((r = h->upload(&dm.VDt, transpose_pad(o.VD, ldx)))) return r;
  if ((r = h->upload(&dm.VD1t, transpose_pad(o.VD1, ldx)))) return r;
  if ((r = h->upload(&dm.VDRt, transpose_pad(o.VDR, ldx)))) return r;
  if ((r = h->upload(&dm.VD1Rt, transpose_pad(o.VD1R, ldx)))) return r;
  if ((r = h->upload(&dm.A1t, transpose_pad(o.A1, ldi)))) return r;
  if ((r = h->upload(&dm.A0t, transpose_pad(o.A0, ldi)))) return r;
  if ((r = h->upload(&dm.AGt, transpose_pad(o.AG, ldi)))) return r;
  if (h->mhd && !h->insul) {
    if ((r = h->upload(&dm.VQ1t, transpose_pad(o.VQ1, ldx)))) return r;
    if ((r = h->upload(&dm.VQ2t, transpose_pad(o.VQ2, ldx)))) return r;
  }
  const int ldc = (n0 + ic - 1) / ic * ic;
  if (h->mhd && h->insul) {
    if ((r = h->upload(&dm.VCt, transpose_pad(o.VC, ldx)))) return r;
    if ((r = h->upload(&dm.VC1t, transpose_pad(o.VC1, ldx)))) return r;
    if ((r = h->upload(&dm.VC2t, transpose_pad(o.VC2, ldx)))) return r;
    if ((r = h->upload(&dm.WQt, transpose_pad(o.WQ, ldc)))) return r;
  }
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
  if (h->maxima_ok) {
    if ((r = h->upload(&dm.twy0, twiddles(h->geo.ny)))) return r;
    if ((r = h->upload(&dm.twz0, twiddles(h->geo.nz)))) return r;
  }
  dm.view = pcf::NodalMats<T>{dm.VBt, dm.VDt, dm.VD1t, ldx, dm.A1t, dm.A0t, dm.AGt, ldi,
                              dm.xnode, dm.twy, dm.twz, dm.VDRt, dm.VD1Rt};
  if (h->mhd) {
    pcf::MhdMats<T> &mm = dm.mview;
    const double W = h->geo.W;
    if (h->insul)
      mm = pcf::MhdMats<T>{dm.VDt, dm.VD1t, dm.VCt, dm.VC1t, dm.VC2t, ldx, nd, n0,
                           dm.AGt, dm.WQt, ldi, ldc, nd, n0, W, 1.0};
    else
      mm = pcf::MhdMats<T>{dm.VDt, dm.VQ1t, dm.VDt, dm.VD1t, dm.VQ2t, ldx, nd, nd,
                           dm.AGt, dm.AGt, ldi, ldi, nd, nd, W, W};
  }
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
  // R^T for the matrix-core reconstruction: [K][ldr], K = nb rows padded with
  // zeros past round4(nb), ldr = nd rounded up to a multiple of 16
  h->ldr = (o.nd + 15) / 16 * 16;
  std::vector<double> Rt((size_t)((o.nb + 3) / 4 * 4) * h->ldr, 0.0);
  for (int i = 0; i < o.nd; ++i)
    for (int j = 0; j < o.nb; ++j) Rt[(size_t)j * h->ldr + i] = o.R(i, j);
  if ((r = h->upload(&h->Rt, Rt))) return r;
  return PCF_OK;
}

// Storage for one FactorSet and its factorisation at gamma.
int32_t build_factors(pcf_handle *h, const pcf::StructMats &sm, double gamma, pcf::FactorSet &fs) {
  int32_t r;
  for (int b = 0; b < 2; ++b)
    for (int p = 0; p < 2; ++p)
      if ((r = h->alloc(&fs.blk[b][p], pcf::factor_doubles(h->geo, sm, b, p)))) return r;
  for (int p = 0; p < 2; ++p)
    if ((r = h->alloc(&fs.mean[p], pcf::factor_doubles(h->geo, sm, -1, p)))) return r;
  hipError_t e = pcf::launch_factor(h->geo, sm, gamma, fs, false, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("factorisation: ") + hipGetErrorString(e));
  return PCF_OK;
}

// Steady factors of the vector-potential systems S_A = M_A - gamma L_A, the
// Dirichlet systems at nu = eta (examples/pcf_mhd_jax.py:116-139).
int32_t build_factors_A(pcf_handle *h, double gamma, pcf::FactorSet &fs) {
  int32_t r;
  for (int p = 0; p < 2; ++p)
    if ((r = h->alloc(&fs.blk[1][p], pcf::factor_doubles(h->geo, h->sm, 1, p)))) return r;
  hipError_t e = pcf::launch_factor(h->geo_eta, h->sm, gamma, fs, true, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("factorisation: ") + hipGetErrorString(e));
  return PCF_OK;
}

// Stage names, in launch order, for per-kernel timing.
constexpr int NSTAGE = 6;
const char *const STAGE_NAMES[NSTAGE] = {"nodal_backward", "fft_y_inverse", "rotational",
                                         "fft_y_forward", "forward_rhs", "solve"};

// Record a timing event at a stage edge; accumulate the previous step's
// interval when a stage starts again.
inline hipError_t mark(pcf_handle *h, int stage, int edge) {
  StageTimer *t = h->timer;
  if (!t) return hipSuccess;
  return hipEventRecord(t->ev[stage][edge], h->stream);
}

// With insulating walls the vector-potential kernels run on a side stream
// beside the flow kernels they do not depend on: fork_side makes the side stream wait for the
// work queued so far on the handle stream, join_side makes the handle stream
// wait for the side stream (inside a graph capture both become graph edges).
hipError_t fork_side(pcf_handle *h) {
  hipError_t e = hipEventRecord(h->ev_fork, h->stream);
  return e != hipSuccess ? e : hipStreamWaitEvent(h->side, h->ev_fork, 0);
}

hipError_t join_side(pcf_handle *h) {
  hipError_t e = hipEventRecord(h->ev_join, h->side);
  return e != hipSuccess ? e : hipStreamWaitEvent(h->stream, h->ev_join, 0);
}

// Under MHD the same stages carry (u, w, b, j) forward and (H, E) back, and
// the A fields of the state (A, three components) give b and j and receive
// the induction rows in NA[slot]; b, j and the induction rows are formed on
// the side stream when there is one.
// The nodal contraction of the state in solution slot xs: from the given
// u1, u2 of a loaded state, else from (u0, g, v00, w00) with the
// reconstruction folded in.
template <typename T>
hipError_t launch_nodal_state(pcf_handle *h, const DevMats<T> &dm, int xs, hipStream_t st) {
  if (h->u12_given) return pcf::launch_nodal_mfma<T>(h->geo, dm.view, h->u0[xs], h->u1, h->u2, h->ws, st);
  return pcf::launch_nodal_prim_mfma<T>(h->geo, dm.view, h->u0[xs], h->g[xs], h->v[xs], h->w[xs], h->ws, st);
}

template <typename T>
hipError_t nonlinear_stages(pcf_handle *h, const DevMats<T> &dm, int xs, int slot, const double2 *A) {
  const pcf::Geo &g = h->geo;
  const pcf::NodalMats<T> &nm = dm.view;
  const bool mhd = h->mhd;
  hipError_t e;
  if ((e = mark(h, 0, 0))) return e;
  if (mhd) {
    if (h->side && (e = fork_side(h))) return e;
    if ((e = pcf::launch_nodal_mhd_mfma<T>(g, dm.mview, A, h->cstride, h->ws, h->side ? h->side : h->stream)))
      return e;
  }
  if ((e = launch_nodal_state<T>(h, dm, xs, h->stream))) return e;
  if (h->side && (e = join_side(h))) return e;
  if ((e = mark(h, 0, 1))) return e;
  if ((e = mark(h, 1, 0))) return e;
  if ((e = pcf::launch_fft_y<T>(g, nm, h->ws, h->py, true, mhd, h->stream))) return e;
  if ((e = mark(h, 1, 1))) return e;
  if ((e = mark(h, 2, 0))) return e;
  if ((e = pcf::launch_rotational<T>(g, nm, h->ws, h->pz, mhd, h->stream))) return e;
  if ((e = mark(h, 2, 1))) return e;
  if ((e = mark(h, 3, 0))) return e;
  if ((e = pcf::launch_fft_y<T>(g, nm, h->ws, h->py, false, mhd, h->stream))) return e;
  if ((e = mark(h, 3, 1))) return e;
  if ((e = mark(h, 4, 0))) return e;
  if (mhd) {
    if (h->side && (e = fork_side(h))) return e;
    if ((e = pcf::launch_rhs_a_mfma<T>(g, dm.mview, h->ws, h->NA[slot], h->cstride, h->side ? h->side : h->stream)))
      return e;
  }
  if ((e = pcf::launch_rhs_mfma<T>(g, nm, h->ws, h->Nu[slot], h->Ng[slot], h->Nv[slot], h->Nw[slot],
                              h->stream)))
    return e;
  if (h->side && (e = join_side(h))) return e;
  return mark(h, 4, 1);
}

// Nonlinear rows of the state in solution slot xs (under MHD with vector
// potential A) into nonlinear slot `slot`.  With ytape set, the padded y-line
// fields of u and curl u are written there (and kept) instead of the
// workspace: the reverse pass reads them back.
int32_t nonlinear_into(pcf_handle *h, int xs, int slot, void *ytape = nullptr, const double2 *A = nullptr) {
  void *saved = h->ws.Y;
  if (ytape) h->ws.Y = ytape;
  hipError_t e = (h->cfg.precision == PCF_PREC_FP64)
                     ? nonlinear_stages<double>(h, h->md, xs, slot, A)
                     : nonlinear_stages<float>(h, h->mf, xs, slot, A);
  h->ws.Y = saved;
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return PCF_OK;
}

// Factored solve + reconstruction of one implicit update on the factor sets
// fs (flow) and fsA (A), by default the steady SBDF3 ones; under MHD also the
// vector-potential solve described by aA (g side, component 0).  aBord: the
// update of the bordered A_y/A_z solve (insulating walls, factorised per
// launch) when it differs from aA.
hipError_t implicit_update(pcf_handle *h, const pcf::SolveArgs &a, const pcf::SolveArgs *aA = nullptr,
                           const pcf::FactorSet *fs = nullptr, const pcf::FactorSet *fsA = nullptr,
                           const pcf::SolveArgs *aBord = nullptr) {
  hipError_t e;
  if (!fs) fs = &h->fsteady;
  if (!fsA) fsA = &h->fA;
  if ((e = mark(h, 5, 0))) return e;
  if (aA) {
    // the A solves on the side stream: components with the Dirichlet (TD)
    // systems (all three for conducting walls, Ax alone for insulating walls)
    // and the bordered A_y/A_z solve
    hipStream_t sa = h->side ? h->side : h->stream;
    if (h->side && (e = fork_side(h))) return e;
    const int ntd = h->insul ? 1 : 3;
    if ((e = pcf::launch_solve_factored(h->geo_eta, h->sm, *fsA, *aA, ntd, h->cstride, sa))) return e;
    if (h->insul) {
      const pcf::SolveArgs &ab = aBord ? *aBord : *aA;
      pcf::BorderedArgs b{};
      for (int l = 0; l < ab.nx; ++l) {
        b.x[l] = ab.g[l];
        b.c[l] = ab.c[l];
        b.d[l] = ab.d[l];
      }
      b.nx = ab.nx;
      for (int l = 0; l < ab.nn; ++l) {
        b.n[l] = ab.Ng[l];
        b.e[l] = ab.e[l];
      }
      b.nn = ab.nn;
      b.gamma = ab.gamma;
      b.out = ab.g_out;
      b.cstride = h->cstride;
      if ((e = pcf::launch_solve_bordered(h->geo_eta, b, sa))) return e;
    }
  }
  if ((e = pcf::launch_solve_factored(h->geo, h->sm, *fs, a, 0, 0, h->stream))) return e;
  if (aA && h->side && (e = join_side(h))) return e;
  // the new state's u1, u2 are formed by the next nodal contraction
  h->u12_given = false;
  h->u12_valid = false;
  return mark(h, 5, 1);
}

// Fill the u1, u2 buffers of the current state when a solve left them unset
// (channelflow_kmm.py:855-870), before anything reads them.
hipError_t ensure_u12(pcf_handle *h);

// Before the first step from a state loaded with given u1, u2: when they are
// the reconstruction of the state's own unknowns (bitwise), drop the flag.
// Synchronous; never called under graph capture.
int32_t resolve_given(pcf_handle *h) {
  if (!h->u12_given || h->u12_checked) return PCF_OK;
  h->u12_checked = true;
  const pcf::Geo &g = h->geo;
  const size_t n = (size_t)g.nd * g.m;
  double2 *r1 = reinterpret_cast<double2 *>(h->ws.X), *r2 = r1 + n;  // workspace as scratch
  pcf::SolveArgs a{};
  const int c = h->so[0];
  a.u0_out = h->u0[c];
  a.g_out = h->g[c];
  a.v_out = h->v[c];
  a.w_out = h->w[c];
  a.u1 = r1;
  a.u2 = r2;
  unsigned diff = 1;
  hipError_t e = pcf::launch_reconstruct(g, h->Rt, h->ldr, a, h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(h->dflag, 0, sizeof(unsigned), h->stream);
  if (e == hipSuccess) e = pcf::launch_bits_differ(h->u1, r1, n, h->dflag, h->stream);
  if (e == hipSuccess) e = pcf::launch_bits_differ(h->u2, r2, n, h->dflag, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&diff, h->dflag, sizeof(unsigned), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("state check: ") + hipGetErrorString(e));
  if (!diff) {
    h->u12_given = false;
    h->u12_valid = true;
  }
  return PCF_OK;
}

hipError_t ensure_u12(pcf_handle *h) {
  if (h->u12_valid) return hipSuccess;
  pcf::SolveArgs a{};
  const int c = h->so[0];
  a.u0_out = h->u0[c];
  a.g_out = h->g[c];
  a.v_out = h->v[c];
  a.w_out = h->w[c];
  a.u1 = h->u1;
  a.u2 = h->u2;
  hipError_t e = pcf::launch_reconstruct(h->geo, h->Rt, h->ldr, a, h->stream);
  if (e == hipSuccess) h->u12_valid = true;
  return e;
}

// A Crank-Nicolson-form update, rhs = M x + gamma L x + sum e N solved with
// M - gamma L at the same gamma (the IMEXRK3 stages, CNAB2), on the stored
// factors fs / fsA at gamma: (M - gamma L)(x_new + x) = 2 M x + sum e N, so the
// factored kernel takes z = 2 x and subtracts x from its output.  The
// bordered A_y/A_z solve of insulating walls keeps the general form.
hipError_t cn_update(pcf_handle *h, const pcf::SolveArgs &a, const pcf::SolveArgs *aA, const pcf::FactorSet &fs,
                     const pcf::FactorSet &fsA) {
  if (a.nx != 1 || a.c[0] != 1.0 || a.d[0] != a.gamma) return hipErrorInvalidValue;
  pcf::SolveArgs af = a, aAf{};
  af.c[0] = 2.0;
  af.d[0] = 0.0;
  af.sub = 1.0;
  if (aA) {
    aAf = *aA;
    aAf.c[0] = 2.0;
    aAf.d[0] = 0.0;
    aAf.sub = 1.0;
  }
  return implicit_update(h, af, aA ? &aAf : nullptr, &fs, &fsA, aA);
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

// The vector-potential part of an implicit update: the g side of SolveArgs
// carries component 0 of A (the solver steps the other two cstride apart).
pcf::SolveArgs a_args_from(pcf_handle *h, const pcf::SolveArgs &a, const int xs[3], int nx, const int ns[3],
                           int nn, int out_slot) {
  pcf::SolveArgs b{};
  for (int l = 0; l < nx; ++l) {
    b.g[l] = h->A[xs[l]];
    b.c[l] = a.c[l];
    b.d[l] = a.d[l];
  }
  b.nx = nx;
  for (int l = 0; l < nn; ++l) {
    b.Ng[l] = h->NA[ns[l]];
    b.e[l] = a.e[l];
  }
  b.nn = nn;
  b.gamma = a.gamma;
  b.g_out = h->A[out_slot];
  return b;
}

// One steady SBDF3 step (sbdf3.py:44-62; channelflow_kmm.py:1016-1064; under
// MHD with the vector-potential update, examples/pcf_mhd_jax.py:643-715).
int32_t steady_step(pcf_handle *h, void *const *tape = nullptr) {
  const double dt = h->cfg.dt;
  const int A = h->no[0];
  int32_t r = nonlinear_into(h, h->so[0], A, tape ? tape[0] : nullptr, h->A[h->so[0]]);
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
  pcf::SolveArgs aA{};
  if (h->mhd) {
    const int xs[3] = {h->so[0], h->so[1], h->so[2]}, ns[3] = {A, h->no[1], h->no[2]};
    aA = a_args_from(h, a, xs, 3, ns, 3, h->so[2]);
  }
  hipError_t e = implicit_update(h, a, h->mhd ? &aA : nullptr);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("solve launch: ") + hipGetErrorString(e));
  rotate(h);
  h->history_steps = 2.0;
  h->have_old = 1.0;
  h->previous_dt = dt;
  return PCF_OK;
}

// One IMEXRK3 startup step (channelflow_kmm.py:872-935 inside the SBDF3
// wrapper :937-1014; under MHD with the vector-potential stages,
// examples/pcf_mhd_jax.py:488-631).  The stage-0 nonlinear rows are the
// step's current nonlinear rows, which the history shift keeps.
int32_t startup_step(pcf_handle *h, void *const *tape = nullptr) {
  const double dt = h->cfg.dt;
  const int A = h->no[0], B = h->no[3], C = h->no[2];
  const int cur = h->so[0], out = h->so[2];
  int32_t r = nonlinear_into(h, cur, A, tape ? tape[0] : nullptr, h->A[cur]);
  if (r) return r;
  const int nslot[3] = {A, B, C};
  for (int rk = 0; rk < 3; ++rk) {
    const double gamma = stage_gamma(rk, dt);
    const int xs = (rk == 0) ? cur : out;
    if (rk > 0) {
      r = nonlinear_into(h, out, nslot[rk], tape ? tape[rk] : nullptr, h->A[out]);
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
    pcf::SolveArgs aA{};
    if (h->mhd) {
      const int xl[3] = {xs, 0, 0}, ns[3] = {nslot[rk], rk > 0 ? nslot[rk - 1] : 0, 0};
      aA = a_args_from(h, a, xl, 1, ns, a.nn, out);
    }
    hipError_t e = cn_update(h, a, h->mhd ? &aA : nullptr, h->fstage[rk], h->fA_stage[rk]);
    if (e != hipSuccess) return fail(PCF_E_HIP, std::string("solve launch: ") + hipGetErrorString(e));
  }
  rotate(h);
  h->history_steps = std::min(2.0, h->history_steps + 1.0);
  h->have_old = 1.0;
  h->previous_dt = dt;
  return PCF_OK;
}

// One ARS(2,2,2) step, the reference's IMEXRK222 KMM step
// (channelflow_kmm.py:1123-1201; tableau src/jaxfun/integrators/imex_rk.py:175-181):
// both stages solve (M - g dt L) x = rhs with
//   rhs_1 = M x_n + g dt N_0,
//   rhs_2 = M x_n + dt (d N_0 + (1 - d) N_1) + (1 - g) dt L x_1.
// Since (M - g dt L) x_1 = rhs_1, the stage-2 linear term is
// (1 - g) dt L x_1 = c1 (M x_1 - rhs_1) with c1 = (1 - g) / g, so
//   rhs_2 = M ((1 - c1) x_n + c1 x_1) + dt (d - 1 + g) N_0 + dt (1 - d) N_1,
// the M z + sum e N form of the factored steady solve (one factor set, fars).
// Under MHD the vector potential takes the same two stages (its Dirichlet
// systems with fA_ars; for insulating walls A_y, A_z by the bordered solve,
// whose replaced wall rows ignore the right-hand side, so the identity holds
// on every row it uses; examples/pcf_mhd_jax.py:389-486).
// The solution ring's other slots and two nonlinear slots are scratch.
int32_t ars222_step(pcf_handle *h) {
  const double dt = h->cfg.dt, gm = ARS_G, dl = 1.0 - 1.0 / (2.0 * ARS_G), c1 = (1.0 - gm) / gm;
  const int cur = h->so[0], x1 = h->so[1], x2 = h->so[2];
  const int NA = h->no[0], NB = h->no[3];
  int32_t r = nonlinear_into(h, cur, NA, nullptr, h->A[cur]);
  if (r) return r;
  hipError_t e;
  {
    pcf::SolveArgs a = base_args(h, x1);
    set_x(a, 0, h, cur, 1.0, 0.0);
    a.nx = 1;
    set_n(a, 0, h, NA, dt * gm);
    a.nn = 1;
    a.gamma = dt * gm;
    pcf::SolveArgs aA{};
    if (h->mhd) {
      const int xs[3] = {cur, 0, 0}, ns[3] = {NA, 0, 0};
      aA = a_args_from(h, a, xs, 1, ns, 1, x1);
    }
    if ((e = implicit_update(h, a, h->mhd ? &aA : nullptr, &h->fars, &h->fA_ars)))
      return fail(PCF_E_HIP, std::string("solve launch: ") + hipGetErrorString(e));
  }
  if ((r = nonlinear_into(h, x1, NB, nullptr, h->A[x1]))) return r;
  {
    pcf::SolveArgs a = base_args(h, x2);
    set_x(a, 0, h, cur, 1.0 - c1, 0.0);
    set_x(a, 1, h, x1, c1, 0.0);
    a.nx = 2;
    set_n(a, 0, h, NA, dt * (dl - 1.0 + gm));
    set_n(a, 1, h, NB, dt * (1.0 - dl));
    a.nn = 2;
    a.gamma = dt * gm;
    pcf::SolveArgs aA{};
    if (h->mhd) {
      const int xs[3] = {cur, x1, 0}, ns[3] = {NA, NB, 0};
      aA = a_args_from(h, a, xs, 2, ns, 2, x2);
    }
    if ((e = implicit_update(h, a, h->mhd ? &aA : nullptr, &h->fars, &h->fA_ars)))
      return fail(PCF_E_HIP, std::string("solve launch: ") + hipGetErrorString(e));
  }
  rotate(h);  // x2 becomes the current slot
  return PCF_OK;
}

// One CNAB2 step (channelflow_kmm.py:1071-1121; under MHD with the
// vector-potential update, examples/pcf_mhd_jax.py:359-387):
//   rhs = M x + (dt/2) L x + dt E,  (M - (dt/2) L) x_new = rhs,
// E = (1 + r/2) N - (r/2) N_old with r = dt / previous_dt when have_old != 0
// (src/jaxfun/integrators/cnab2.py:211-239), else E = N, on the factors at
// dt/2 (fcn, fA_cn; pcf_set_dt refactors them).  N lands in the current
// nonlinear slot, which the ring shift makes NL_OLD.
int32_t cnab2_step(pcf_handle *h) {
  const double dt = h->cfg.dt;
  const int cur = h->so[0], out = h->so[2], NA = h->no[0], NO = h->no[1];
  const bool have = h->have_old != 0.0;
  if (have && !(h->previous_dt > 0.0))
    return fail(PCF_E_INVALID, "CNAB2 history: have_old set with previous_dt <= 0");
  int32_t r = nonlinear_into(h, cur, NA, nullptr, h->A[cur]);
  if (r) return r;
  pcf::SolveArgs a = base_args(h, out);
  set_x(a, 0, h, cur, 1.0, 0.5 * dt);
  a.nx = 1;
  if (have) {
    const double q = dt / h->previous_dt;
    set_n(a, 0, h, NA, dt * (1.0 + 0.5 * q));
    set_n(a, 1, h, NO, -dt * 0.5 * q);
    a.nn = 2;
  } else {
    set_n(a, 0, h, NA, dt);
    a.nn = 1;
  }
  a.gamma = 0.5 * dt;
  pcf::SolveArgs aA{};
  if (h->mhd) {
    const int xs[3] = {cur, 0, 0}, ns[3] = {NA, NO, 0};
    aA = a_args_from(h, a, xs, 1, ns, a.nn, out);
  }
  hipError_t e = cn_update(h, a, h->mhd ? &aA : nullptr, h->fcn, h->fA_cn);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("solve launch: ") + hipGetErrorString(e));
  rotate(h);
  h->have_old = 1.0;
  h->previous_dt = dt;
  return PCF_OK;
}

// Nodal fields of the current state, their y-lines on the unpadded grid and
// the per-block maxima (velocity; under MHD also b = curl A).
template <typename T>
hipError_t field_maxima_stages(pcf_handle *h, const DevMats<T> &dm, double *blockmax, double *phys = nullptr) {
  const pcf::Geo &g = h->geo, &g0 = h->geo0;
  pcf::NodalMats<T> nm0 = dm.view;
  nm0.twy = dm.twy0;
  nm0.twz = dm.twz0;
  hipError_t e;
  if ((e = launch_nodal_state<T>(h, dm, h->so[0], h->stream))) return e;
  if (h->mhd && (e = pcf::launch_nodal_mhd_mfma<T>(g, dm.mview, h->A[h->so[0]], h->cstride, h->ws, h->stream)))
    return e;
  T *X = reinterpret_cast<T *>(h->ws.X), *Y = reinterpret_cast<T *>(h->ws.Y);  // complex: 2 T each
  if ((e = pcf::launch_fft_y_fields<T>(g0, nm0, h->py0, true, 3, X, Y, h->stream))) return e;
  // b (nodal fields 6..8) after the three velocity fields of the unpadded lines
  if (h->mhd && (e = pcf::launch_fft_y_fields<T>(g0, nm0, h->py0, true, 3, X + (size_t)2 * 6 * g.n0 * g.m,
                                                 Y + (size_t)2 * 3 * g.n0 * g0.my * g0.kzs, h->stream)))
    return e;
  return pcf::launch_absmax<T>(g0, dm.xnode, dm.twz0, h->pz0, h->mhd ? 6 : 3, Y, blockmax, h->stream, phys);
}

// The kernel(s) of nonlinear stage `stage` (0 nodal_backward .. 4 forward_rhs)
// for the current state, as nonlinear_stages launches them.
template <typename T>
hipError_t stage_launch(pcf_handle *h, const DevMats<T> &dm, int stage) {
  const pcf::Geo &g = h->geo;
  const bool mhd = h->mhd;
  const int slot = h->no[0], cur = h->so[0];
  hipError_t e = hipSuccess;
  switch (stage) {
    case 0:
      e = launch_nodal_state<T>(h, dm, cur, h->stream);
      if (e == hipSuccess && mhd) e = pcf::launch_nodal_mhd_mfma<T>(g, dm.mview, h->A[cur], h->cstride, h->ws, h->stream);
      return e;
    case 1: return pcf::launch_fft_y<T>(g, dm.view, h->ws, h->py, true, mhd, h->stream);
    case 2: return pcf::launch_rotational<T>(g, dm.view, h->ws, h->pz, mhd, h->stream);
    case 3: return pcf::launch_fft_y<T>(g, dm.view, h->ws, h->py, false, mhd, h->stream);
    default:
      e = pcf::launch_rhs_mfma<T>(g, dm.view, h->ws, h->Nu[slot], h->Ng[slot], h->Nv[slot], h->Nw[slot], h->stream);
      if (e == hipSuccess && mhd) e = pcf::launch_rhs_a_mfma<T>(g, dm.mview, h->ws, h->NA[slot], h->cstride, h->stream);
      return e;
  }
}

template <typename T>
hipError_t stage_repeat(pcf_handle *h, const DevMats<T> &dm, int stage, int reps, hipEvent_t e0, hipEvent_t e1) {
  hipError_t e;
  for (int s = 0; s < stage; ++s)  // the stage's inputs for the current state
    if ((e = stage_launch<T>(h, dm, s))) return e;
  if ((e = hipEventRecord(e0, h->stream))) return e;
  for (int r = 0; r < reps; ++r)
    if ((e = stage_launch<T>(h, dm, stage))) return e;
  return hipEventRecord(e1, h->stream);
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
  if (f >= PCF_F_AX && f < PCF_F_COUNT) {
    if (!h->mhd) return false;
    const int k = f - PCF_F_AX, grp = k / 3, comp = k % 3;
    // groups: A, NL_A_OLD, NL_A_OLDER, A_OLD, A_OLDER
    double2 *base = (grp == 0)   ? h->A[h->so[0]]
                    : (grp == 1) ? h->NA[h->no[1]]
                    : (grp == 2) ? h->NA[h->no[2]]
                    : (grp == 3) ? h->A[h->so[1]]
                                 : h->A[h->so[2]];
    fr = {base + comp * h->cstride, true, comp ? h->arows : nd};
    return true;
  }
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
  if (f >= PCF_F_AX && f < PCF_F_COUNT) {
    const bool tc = c->mhd && c->walls == PCF_WALLS_INSULATING && (f - PCF_F_AX) % 3 != 0;
    return tc ? c->n0 : nd;
  }
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


// ----------------------------------------------------------------------------
// Reverse mode.  Cotangents mu are gradients for Re sum conj(a) b in the
// Hermitian half layout; the exported gradient is jax's convention conj(mu)
// expanded to the full spectrum (examples/pcf_minimal_seed_jax.py:78-94).
//
// Levels: x_k = (u0, g, v00, w00) of state k, N_k its nonlinear rows.  A steady
// step k solves S x_{k+1} = sum_j c_j M x_{k-j} + sum_j e_j N_{k-j}; its
// transpose gives rho_k = S^{-T} mu_{x_{k+1}}, and then
//   mu_{N_k} = e_0 rho_k + e_1 rho_{k+1} + e_2 rho_{k+2},
//   mu_{x_k} = M^T (c_0 rho_k + c_1 rho_{k+1} + c_2 rho_{k+2}) + R^T F'(u_k)^T mu_{N_k},
// with rho_k = 0 for the IMEXRK3 start-up steps, whose three stages are
// reversed one by one.  The forward is recomputed segment by segment from
// state snapshots taken every `every` steps; the recompute keeps the padded
// y-line fields of every nonlinear evaluation (the tape), so the reverse
// never repeats the wall-normal transform.  The recompute of segment s - 1
// (handle stream, graph-replayed steady runs) runs concurrently with the
// reverse sweep of segment s (reverse stream, own workspace), with two tape
// sets ordered by events: at this size every kernel leaves most of the
// device idle, so the two sweeps share it.

const double C_SBDF[3] = {18.0 / 11.0, -9.0 / 11.0, 2.0 / 11.0};

std::vector<std::pair<void *, size_t>> state_pieces(pcf_handle *h) {
  const pcf::Geo &g = h->geo;
  const size_t M = g.m, c = sizeof(double2), r = sizeof(double);
  std::vector<std::pair<void *, size_t>> p;
  for (int s = 0; s < 3; ++s) {
    p.push_back({h->u0[s], g.nb * M * c});
    p.push_back({h->g[s], g.nd * M * c});
    p.push_back({h->v[s], g.nd * r});
    p.push_back({h->w[s], g.nd * r});
  }
  for (int s = 0; s < 4; ++s) {
    p.push_back({h->Nu[s], g.nb * M * c});
    p.push_back({h->Ng[s], g.nd * M * c});
    p.push_back({h->Nv[s], g.nd * r});
    p.push_back({h->Nw[s], g.nd * r});
  }
  p.push_back({h->u1, g.nd * M * c});
  p.push_back({h->u2, g.nd * M * c});
  return p;
}

// n steady steps, each keeping its nonlinear evaluation in tape slot
// tape_base + i (no tape when tape_base < 0), replayed from a graph cached per
// (ring phase, n, tape_base): the captured launches and pointers are the same
// whenever the rings stand at the same phase, so one graph serves every
// segment of the sweeps that starts there.
// The ring phase: which slot holds each level of the solution and nonlinear
// rings (captured launches hold absolute slot pointers).
uint64_t ring_phase(const pcf_handle *h) {
  uint64_t key = 0;
  for (int i = 0; i < 3; ++i) key = key * 4 + (uint64_t)h->so[i];
  for (int i = 0; i < 4; ++i) key = key * 4 + (uint64_t)h->no[i];
  return key;  // < 2^14
}

// The host bookkeeping of n steady steps: ring rotation and the history
// scalars a steady step leaves (channelflow_kmm.py:1056-1064).
void advance_host(pcf_handle *h, int n) {
  for (int i = 0; i < n; ++i) rotate(h);
  h->history_steps = 2.0;
  h->have_old = 1.0;
  h->previous_dt = h->cfg.dt;
  h->u12_given = false;
  h->u12_valid = false;
}

// Graph cache key: ring phase (< 2^14), whether the first step reads given
// u1/u2 (bit 14), the length and the tape base.
uint64_t graph_key(const pcf_handle *h, int n, int tape_base) {
  return ring_phase(h) | (uint64_t)(h->u12_given ? 1 : 0) << 14 | (uint64_t)n << 16 |
         (uint64_t)(tape_base + 1) << 40;
}

// Capture n steady steps (tape slots tape_base + i, none when tape_base < 0)
// at the rings' current phase into an executable graph.  Capture records the
// launches without running them; the host ring indices and history scalars,
// which steady_step advances as it enqueues, are put back afterwards, also
// when the capture or the instantiation fails (the partial graph is then
// destroyed and nothing is cached).
int32_t capture_steady(pcf_handle *h, int n, int tape_base, hipGraphExec_t *out) {
  int so[3], no[4];
  std::copy(h->so, h->so + 3, so);
  std::copy(h->no, h->no + 4, no);
  const double hs = h->history_steps, ho = h->have_old, pd = h->previous_dt;
  const bool given = h->u12_given, valid = h->u12_valid;
  auto restore = [&]() {
    std::copy(so, so + 3, h->so);
    std::copy(no, no + 4, h->no);
    h->history_steps = hs;
    h->have_old = ho;
    h->previous_dt = pd;
    h->u12_given = given;
    h->u12_valid = valid;
  };
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("begin capture: ") + hipGetErrorString(e));
  int32_t r = PCF_OK;
  for (int i = 0; i < n && !r; ++i) r = steady_step(h, tape_base >= 0 ? &h->adj.tape[tape_base + i] : nullptr);
  e = hipStreamEndCapture(h->stream, &graph);
  restore();
  if (r || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return r ? r : fail(PCF_E_HIP, std::string("capture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("instantiate: ") + hipGetErrorString(e));
  *out = exec;
  return PCF_OK;
}

// The graph of n steady steps at the current ring phase from cache (captured
// on a miss).  Captured launches hold absolute slot pointers, so one graph
// serves every run that starts with the rings at the same phase.
int32_t cached_steady(pcf_handle *h, std::map<uint64_t, hipGraphExec_t> &cache, uint64_t key,
                      int n, int tape_base, hipGraphExec_t *out) {
  auto it = cache.find(key);
  if (it == cache.end()) {
    hipGraphExec_t exec = nullptr;
    int32_t r = capture_steady(h, n, tape_base, &exec);
    if (r) return r;
    it = cache.emplace(key, exec).first;
    ++h->graph_misses;
  } else {
    ++h->graph_hits;
  }
  *out = it->second;
  return PCF_OK;
}

// n steady steps, each keeping its nonlinear evaluation in tape slot
// tape_base + i (no tape when tape_base < 0), replayed from a graph cached per
// (ring phase, n, tape_base).
int32_t steady_run(pcf_handle *h, int n, int tape_base) {
  const uint64_t key = graph_key(h, n, tape_base);
  hipGraphExec_t exec = nullptr;
  int32_t r = cached_steady(h, h->seg_graphs, key, n, tape_base, &exec);
  if (r) return r;
  PCF_HIP(hipGraphLaunch(exec, h->stream));
  advance_host(h, n);
  return PCF_OK;
}

// MHD: the seven vector-potential slots (3 solution, 4 nonlinear) beside the
// state block in a snapshot.
hipError_t snap_mhd(pcf_handle *h, char *dev, bool save) {
  const size_t bytes = 3 * h->cstride * sizeof(double2);
  for (int s = 0; s < 7; ++s) {
    double2 *slot = (s < 3) ? h->A[s] : h->NA[s - 3];
    char *snap = dev + (size_t)s * bytes;
    hipError_t e = save ? hipMemcpyAsync(snap, slot, bytes, hipMemcpyDeviceToDevice, h->stream)
                        : hipMemcpyAsync(slot, snap, bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t snap_save(pcf_handle *h, pcf_handle::Snap &sn) {
  // the state block holds the pieces at the snapshot's offsets: one copy
  hipError_t e = hipMemcpyAsync(sn.dev, h->state_block, h->state_bytes, hipMemcpyDeviceToDevice, h->stream);
  if (e != hipSuccess) return e;
  if (sn.devA && (e = snap_mhd(h, sn.devA, true))) return e;
  std::copy(h->so, h->so + 3, sn.so);
  std::copy(h->no, h->no + 4, sn.no);
  sn.hs = h->history_steps;
  sn.ho = h->have_old;
  sn.pd = h->previous_dt;
  sn.given = h->u12_given;
  sn.valid = h->u12_valid;
  return hipSuccess;
}

hipError_t snap_load(pcf_handle *h, const pcf_handle::Snap &sn) {
  hipError_t e = hipMemcpyAsync(h->state_block, sn.dev, h->state_bytes, hipMemcpyDeviceToDevice, h->stream);
  if (e != hipSuccess) return e;
  if (sn.devA && (e = snap_mhd(h, sn.devA, false))) return e;
  std::copy(sn.so, sn.so + 3, h->so);
  std::copy(sn.no, sn.no + 4, h->no);
  h->history_steps = sn.hs;
  h->have_old = sn.ho;
  h->previous_dt = sn.pd;
  h->u12_given = sn.given;
  h->u12_valid = sn.valid;
  return hipSuccess;
}

template <typename T>
int32_t adj_alloc(pcf_handle *h, T **p, size_t count) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("hipMalloc (adjoint): ") + hipGetErrorString(e));
  h->adj_allocs.push_back(q);
  *p = static_cast<T *>(q);
  return PCF_OK;
}

int32_t alloc_prim(pcf_handle *h, pcf::Prim &p) {
  const pcf::Geo &g = h->geo;
  int32_t r;
  if ((r = h->alloc(&p.u0, (size_t)g.nb * g.m))) return r;
  if ((r = h->alloc(&p.g, (size_t)g.nd * g.m))) return r;
  if ((r = h->alloc(&p.v, (size_t)g.nd))) return r;
  return h->alloc(&p.w, (size_t)g.nd);
}

pcf::PrimC cprim(const pcf::Prim &p) { return pcf::PrimC{p.u0, p.g, p.v, p.w}; }

hipError_t zero_prim(pcf_handle *h, const pcf::Prim &p) {
  const pcf::Geo &g = h->geo;
  hipError_t e;
  if ((e = hipMemsetAsync(p.u0, 0, (size_t)g.nb * g.m * sizeof(double2), h->adj.stream))) return e;
  if ((e = hipMemsetAsync(p.g, 0, (size_t)g.nd * g.m * sizeof(double2), h->adj.stream))) return e;
  if ((e = hipMemsetAsync(p.v, 0, (size_t)g.nd * sizeof(double), h->adj.stream))) return e;
  return hipMemsetAsync(p.w, 0, (size_t)g.nd * sizeof(double), h->adj.stream);
}

template <typename T>
int32_t build_adj_mats(pcf_handle *h, pcf::AdjMats<T> &am) {
  const pcf::Operators &o = h->ops;
  const int ldx = (o.n0 + 15) / 16 * 16, ldi = (o.nd + 15) / 16 * 16;  // whole 16-row MFMA tiles
  const int kr = (o.n0 + 3) / 4 * 4;  // zero rows up to a multiple of 4 >= n0 (K loops)
  auto pad = [&](const pcf::Mat &Mx, int ld) {  // row-major [rows][ld]
    std::vector<T> t((size_t)std::max(Mx.rows, kr) * ld, (T)0);
    for (int i = 0; i < Mx.rows; ++i)
      for (int j = 0; j < Mx.cols; ++j) t[(size_t)i * ld + j] = (T)Mx(i, j);
    return t;
  };
  T *d = nullptr;
  int32_t r;
  if ((r = h->upload(&d, pad(o.A1, ldx)))) return r;
  am.A1 = d;
  if ((r = h->upload(&d, pad(o.A0, ldx)))) return r;
  am.A0 = d;
  if ((r = h->upload(&d, pad(o.AG, ldx)))) return r;
  am.AG = d;
  am.ldx = ldx;
  if ((r = h->upload(&d, pad(o.VB, ldi)))) return r;
  am.VB = d;
  if ((r = h->upload(&d, pad(o.VD, ldi)))) return r;
  am.VD = d;
  if ((r = h->upload(&d, pad(o.VD1, ldi)))) return r;
  am.VD1 = d;
  // the reconstruction f = R u0 composed into the nodal values of f
  auto times_R = [&](const pcf::Mat &V) {
    pcf::Mat P;
    P.rows = V.rows;
    P.cols = o.nb;
    P.a.assign((size_t)P.rows * P.cols, 0.0);
    for (int x = 0; x < V.rows; ++x)
      for (int i = 0; i < o.nd; ++i)
        for (int j = 0; j < o.nb; ++j) P(x, j) += V(x, i) * o.R(i, j);
    return P;
  };
  if ((r = h->upload(&d, pad(times_R(o.VD), ldi)))) return r;
  am.VDR = d;
  if ((r = h->upload(&d, pad(times_R(o.VD1), ldi)))) return r;
  am.VD1R = d;
  am.ldi = ldi;
  return PCF_OK;
}

// One-time reverse-mode set-up: transposed structures, R^T, energy Gram
// matrices, contraction matrices and cotangent buffers.
int32_t adj_setup(pcf_handle *h) {
  if (h->adj.ready) return PCF_OK;
  const pcf::Operators &o = h->ops;
  const pcf::Geo &g = h->geo;
  auto &A = h->adj;
  int32_t r;
  if (!pcf::rot_adj_fits(g, h->pz, h->cfg.precision == PCF_PREC_FP64 ? 16 : 8))
    return fail(PCF_E_UNSUPPORTED, "padded spanwise grid too large for the cross-product adjoint");
  if (h->mhd && (size_t)19 * g.mz * sizeof(double2) > 65536)
    return fail(PCF_E_UNSUPPORTED, "padded spanwise grid too large for the MHD cross-product adjoint");
  // A' = J A^T J per parity block: same band offsets, b'[i][d+P] = b[n-1-i-d][d+P],
  // far factors U'[i] = V[n-1-i], V'[j] = U[n-1-j].
  for (int b = 0; b < 2; ++b)
    for (int p = 0; p < 2; ++p) {
      const pcf::Operators::Struct &st = o.st[b][p];
      const int n = st.n, BW = st.P + st.Q + 1, R = st.R;
      std::vector<double> bt[3];
      const std::vector<double> *src[3] = {&st.b0, &st.b2, &st.b4};
      for (int t = 0; t < 3; ++t) {
        bt[t].assign((size_t)n * BW, 0.0);
        for (int i = 0; i < n; ++i)
          for (int d = -st.P; d <= st.Q; ++d) {
            const int row = n - 1 - i - d;
            if (row < 0 || row >= n) continue;
            bt[t][(size_t)i * BW + d + st.P] = (*src[t])[(size_t)row * BW + d + st.P];
          }
      }
      std::vector<double> Ut((size_t)n * R), Vt((size_t)n * R);
      for (int i = 0; i < n; ++i)
        for (int q = 0; q < R; ++q) {
          Ut[(size_t)i * R + q] = st.V[(size_t)(n - 1 - i) * R + q];
          Vt[(size_t)i * R + q] = st.U[(size_t)(n - 1 - i) * R + q];
        }
      pcf::StructSys &ss = A.smt.sys[b][p];
      ss.n = n;
      double *d = nullptr;
      if ((r = h->upload(&d, bt[0]))) return r;
      ss.b0 = d;
      if ((r = h->upload(&d, bt[1]))) return r;
      ss.b2 = d;
      if ((r = h->upload(&d, bt[2]))) return r;
      ss.b4 = d;
      if ((r = h->upload(&d, Ut))) return r;
      ss.U = d;
      if ((r = h->upload(&d, Vt))) return r;
      ss.V = d;
    }
  if ((r = build_factors(h, A.smt, h->cfg.dt * 6.0 / 11.0, A.fsteady_t))) return r;
  // RcT[j][k] = R[2k + 1 - (j & 1)][j]
  A.ldrT = (o.nd + 2) / 2;
  std::vector<double> RcT((size_t)o.nb * A.ldrT, 0.0);
  for (int j = 0; j < o.nb; ++j) {
    const int i0 = 1 - (j & 1);
    for (int k = 0; i0 + 2 * k < o.nd; ++k) RcT[(size_t)j * A.ldrT + k] = o.R(i0 + 2 * k, j);
  }
  if ((r = h->upload(&A.RcT, RcT))) return r;
  // Q = V^T diag(fejer) V for the clamped and Dirichlet bases
  auto gram = [&](const pcf::Mat &Vm) {
    std::vector<double> Q((size_t)Vm.cols * Vm.cols, 0.0);
    for (int i = 0; i < Vm.cols; ++i)
      for (int j = 0; j < Vm.cols; ++j) {
        double s = 0;
        for (int x = 0; x < Vm.rows; ++x) s += Vm(x, i) * o.fejer[x] * Vm(x, j);
        Q[(size_t)i * Vm.cols + j] = s;
      }
    return Q;
  };
  if ((r = h->upload(&A.QB, gram(o.VB)))) return r;
  if ((r = h->upload(&A.QD, gram(o.VD)))) return r;
  if (h->cfg.precision == PCF_PREC_FP64) {
    if ((r = build_adj_mats<double>(h, A.ad))) return r;
  } else {
    if ((r = build_adj_mats<float>(h, A.af))) return r;
  }
  if ((r = alloc_prim(h, A.muX)) || (r = alloc_prim(h, A.nu)) || (r = alloc_prim(h, A.sig))) return r;
  {
    const size_t ve = (h->cfg.precision == PCF_PREC_FP64) ? sizeof(double2) : sizeof(float2);
    const size_t nf = h->mhd ? 12 : 6, nh = h->mhd ? 6 : 3;
    const size_t nX = nf * (size_t)g.n0 * g.m, nY = nf * (size_t)g.n0 * g.my * g.kzs;
    const size_t nHz = nh * (size_t)g.n0 * g.my * g.kzs, nHx = nh * (size_t)g.n0 * g.m;
    unsigned char *wsp = nullptr;
    if ((r = h->alloc(&wsp, (nX + nY + nHz + nHx) * ve))) return r;
    A.ws.X = wsp;
    A.ws.Y = wsp + nX * ve;
    A.ws.Hz = wsp + (nX + nY) * ve;
    A.ws.Hx = wsp + (nX + nY + nHz) * ve;
    if ((r = h->alloc(&A.scratch, pcf::solve_scratch_doubles(g)))) return r;
    PCF_HIP(hipStreamCreateWithFlags(&A.stream, hipStreamNonBlocking));
    for (hipEvent_t *ev : {&A.ev_seed, &A.ev_rc[0], &A.ev_rc[1], &A.ev_rev[0], &A.ev_rev[1]})
      PCF_HIP(hipEventCreateWithFlags(ev, hipEventDisableTiming));
  }
  for (int i = 0; i < 3; ++i)
    if ((r = alloc_prim(h, A.rho[i])) || (r = alloc_prim(h, A.rs[i]))) return r;
  for (int i = 0; i < 4; ++i)
    if ((r = alloc_prim(h, A.gh[i]))) return r;
  for (int i = 0; i < 3; ++i)
    if ((r = h->alloc(&A.mu_u[i], (size_t)(i == 0 ? g.nb : g.nd) * g.m))) return r;
  for (int i = 0; i < 3; ++i)
    if ((r = h->alloc(&A.cot_u[i], (size_t)(i == 0 ? g.nb : g.nd) * g.m))) return r;
  if ((r = h->alloc(&A.cot_g, (size_t)g.nd * g.m))) return r;
  A.snap_bytes = 0;
  for (auto &pc : state_pieces(h)) A.snap_bytes += (pc.second + 255) / 256 * 256;
  A.ready = true;
  return PCF_OK;
}

// Tape slots for `every` steps and snapshots for `steps` steps.
int32_t adj_reserve(pcf_handle *h, int steps, int every) {
  auto &A = h->adj;
  const pcf::Geo &g = h->geo;
  const size_t ve = (h->cfg.precision == PCF_PREC_FP64) ? sizeof(double2) : sizeof(float2);
  const size_t slot_bytes = (h->mhd ? 12 : 6) * (size_t)g.n0 * g.my * g.kzs * ve;
  const int per_set = std::min(every, steps) + 4;  // start-up steps evaluate 3 times
  const size_t need_snap = (size_t)((steps + every - 1) / every) + 1;
  int32_t r;
  // two sets (double buffering between the recompute and the reverse sweep);
  // slot pointers never change, so cached graphs stay valid
  A.per_set = per_set;
  while (A.tape.size() < 2 * (size_t)per_set) {
    unsigned char *p = nullptr;
    if ((r = adj_alloc(h, &p, slot_bytes))) return r;
    A.tape.push_back(p);
  }
  while (A.snaps.size() < need_snap) {
    pcf_handle::Snap sn;
    if ((r = adj_alloc(h, &sn.dev, A.snap_bytes))) return r;
    if (h->mhd && (r = adj_alloc(h, &sn.devA, 7 * 3 * h->cstride * sizeof(double2)))) return r;
    A.snaps.push_back(sn);
  }
  return PCF_OK;
}

template <typename T>
hipError_t nonlinear_adjoint(pcf_handle *h, const pcf::NodalMats<T> &nm, const pcf::AdjMats<T> &am,
                             const pcf::Prim &nu, const void *ytape) {
  const pcf::Geo &g = h->geo;
  auto &A = h->adj;
  hipError_t e;
  if ((e = pcf::launch_rhs_adj_mfma<T>(g, am, nu, A.ws, A.stream))) return e;
  if ((e = pcf::launch_fft_y_fields<T>(g, nm, h->py, true, 3, A.ws.Hx, A.ws.Hz, A.stream))) return e;
  if ((e = pcf::launch_rot_adj<T>(g, nm, ytape, A.ws, h->pz, A.stream))) return e;
  if ((e = pcf::launch_fft_y_fields<T>(g, nm, h->py, false, 6, A.ws.Y, A.ws.X, A.stream))) return e;
  return pcf::launch_nodal_adj_mfma<T>(g, am, A.ws, A.mu_u[0], A.mu_u[1], A.mu_u[2], A.stream);
}

// mu_u = F'(u)^T nu for the nonlinear evaluation whose forward fields are ytape;
// a weighted sum of cotangents is first folded into adj.nu.
hipError_t nl_adj(pcf_handle *h, const pcf::PrimSum &nu, const void *ytape) {
  auto &A = h->adj;
  if (!(nu.n == 1 && nu.wgt[0] == 1.0 && nu.t[0].u0 == A.nu.u0)) {
    hipError_t e = pcf::launch_prim_sum(h->geo, nu, A.nu, A.stream);
    if (e != hipSuccess) return e;
  }
  return (h->cfg.precision == PCF_PREC_FP64) ? nonlinear_adjoint<double>(h, h->md.view, A.ad, A.nu, ytape)
                                             : nonlinear_adjoint<float>(h, h->mf.view, A.af, A.nu, ytape);
}

// Steady reverse step with k > 0: the nonlinear adjoint of adj.nu with the
// reconstruction adjoint and the mass gather of adj.sig folded into its last
// contraction, writing mu_{x_k} to adj.muX.
template <typename T>
hipError_t nl_adj_x(pcf_handle *h, const pcf::NodalMats<T> &nm, const pcf::AdjMats<T> &am, const void *ytape) {
  const pcf::Geo &g = h->geo;
  auto &A = h->adj;
  hipError_t e;
  if ((e = pcf::launch_rhs_adj_mfma<T>(g, am, A.nu, A.ws, A.stream))) return e;
  if ((e = pcf::launch_fft_y_fields<T>(g, nm, h->py, true, 3, A.ws.Hx, A.ws.Hz, A.stream))) return e;
  if ((e = pcf::launch_rot_adj<T>(g, nm, ytape, A.ws, h->pz, A.stream))) return e;
  if ((e = pcf::launch_fft_y_fields<T>(g, nm, h->py, false, 6, A.ws.Y, A.ws.X, A.stream))) return e;
  return pcf::launch_nodal_adj_x_mfma<T>(g, am, h->sm, cprim(A.sig), A.ws, A.muX, A.stream);
}

void ps_add(pcf::PrimSum &s, const pcf::Prim &p, double w) {
  s.t[s.n] = cprim(p);
  s.wgt[s.n] = w;
  ++s.n;
}

// mu_{x_k} from the nonlinear-adjoint output: R^T mu_u, or zero when state k
// is the initial state (its velocity leaves keep mu_u).
hipError_t recon_or_zero(pcf_handle *h, bool leaves) {
  auto &A = h->adj;
  if (leaves) return zero_prim(h, A.muX);
  return pcf::launch_recon_adj(h->geo, A.RcT, A.ldrT, A.mu_u[0], A.mu_u[1], A.mu_u[2], nullptr, A.muX,
                               A.stream);
}

// Reverse of step k (state k -> k+1); mu_{x_{k+1}} in muX on entry, mu_{x_k}
// on exit.  tape[0..2] are the step's nonlinear evaluations.
hipError_t reverse_step(pcf_handle *h, int k, bool steady, void *const *tape) {
  auto &A = h->adj;
  const pcf::Geo &g = h->geo;
  const double dt = h->cfg.dt;
  const double e[3] = {dt * 18.0 / 11.0, -dt * 18.0 / 11.0, dt * 6.0 / 11.0};
  int *rr = A.rr;
  {
    const int t = rr[2];
    rr[2] = rr[1];
    rr[1] = rr[0];
    rr[0] = t;  // rr[0] now receives rho_k
  }
  hipError_t er;
  if (steady) {
    // rho_k and, in the same pass, mu_{N_k} = e0 rho_k + e1 rho_k+1 + e2 rho_k+2
    // and sig = sum_j C_j rho_{k+j}
    if ((er = pcf::launch_adj_solve_factored(g, A.smt, A.fsteady_t, cprim(A.muX), A.rho[rr[0]], A.nu,
                                             cprim(A.rho[rr[1]]), cprim(A.rho[rr[2]]), e[0], e[1], e[2],
                                             A.stream, A.sig, C_SBDF[0], C_SBDF[1], C_SBDF[2])))
      return er;
    if (k == 0) {  // initial state: its velocity leaves keep mu_u
      pcf::PrimSum nu{};
      ps_add(nu, A.nu, 1.0);
      if ((er = nl_adj(h, nu, tape[0]))) return er;
      pcf::PrimSum ga{};
      for (int j = 0; j < 3; ++j) ps_add(ga, A.rho[rr[j]], C_SBDF[j]);
      if ((er = zero_prim(h, A.muX))) return er;
      return pcf::launch_adj_mass(g, h->sm, 0.0, pcf::PrimC{}, ga, A.muX, A.stream);
    }
    // mu_{x_k} = R^T F'(u_k)^T nu + M^T sig in one pass after the transforms
    return (h->cfg.precision == PCF_PREC_FP64) ? nl_adj_x<double>(h, h->md.view, A.ad, tape[0])
                                               : nl_adj_x<float>(h, h->mf.view, A.af, tape[0]);
  }
  if ((er = zero_prim(h, A.rho[rr[0]]))) return er;
  for (int s = 2; s >= 0; --s) {
    const double gamma = (IMEX_A[s] + IMEX_B[s]) * dt / 2.0;
    if ((er = pcf::launch_adj_solve(g, A.smt, gamma, cprim(A.muX), A.rs[s], A.scratch, A.stream)))
      return er;
    pcf::PrimSum nu{};
    ps_add(nu, A.rs[s], dt * IMEX_A[s]);
    if (s < 2) ps_add(nu, A.rs[s + 1], dt * IMEX_B[s + 1]);
    if (s == 0) {
      ps_add(nu, A.rho[rr[1]], e[1]);
      ps_add(nu, A.rho[rr[2]], e[2]);
    }
    if ((er = nl_adj(h, nu, tape[s]))) return er;
    if ((er = recon_or_zero(h, s == 0 && k == 0))) return er;
    pcf::PrimSum ga{};
    if (s == 0) {
      ps_add(ga, A.rho[rr[1]], C_SBDF[1]);
      ps_add(ga, A.rho[rr[2]], C_SBDF[2]);
    }
    if ((er = pcf::launch_adj_mass(g, h->sm, gamma, cprim(A.rs[s]), ga, A.muX, A.stream))) return er;
  }
  return hipSuccess;
}

// ------------------------------------------------------------- MHD reverse
// The vector-potential side of the reverse pass, conducting walls (A in TD^3),
// FP64.  Each A component is a Dirichlet system at nu = eta, so the flow's
// transposed solves, mass and sum kernels run on it through the g slot of a
// primary set whose other slots are a zero set (read) and a sink set (written).

pcf::Prim a_out(const pcf_handle *h, double2 *comp) {
  const auto &A = h->adj;
  return pcf::Prim{A.sink.u0, comp, A.sink.v, A.sink.w};
}
pcf::PrimC a_in(const pcf_handle *h, const double2 *comp) {
  const auto &A = h->adj;
  return pcf::PrimC{A.zero.u0, comp, A.zero.v, A.zero.w};
}

int32_t adj_setup_mhd(pcf_handle *h) {
  auto &A = h->adj;
  if (A.mhd_ready) return PCF_OK;
  const pcf::Geo &g = h->geo;
  const pcf::Operators &o = h->ops;
  const size_t n3 = 3 * h->cstride;
  int32_t r;
  if ((r = h->alloc(&A.muA, n3)) || (r = h->alloc(&A.nuA, n3)) || (r = h->alloc(&A.sigA, n3)) ||
      (r = h->alloc(&A.cotA, n3)))
    return r;
  for (int i = 0; i < 3; ++i)
    if ((r = h->alloc(&A.rhoA[i], n3)) || (r = h->alloc(&A.rsA[i], n3))) return r;
  for (int i = 0; i < 4; ++i)
    if ((r = h->alloc(&A.ghA[i], n3))) return r;
  if ((r = alloc_prim(h, A.zero)) || (r = alloc_prim(h, A.sink))) return r;
  // the start-up transposed solves of the three components run in one launch
  if ((r = h->alloc(&A.scratchA, pcf::adj_solve_dir_scratch_doubles(g, 3)))) return r;
  // steady factors of the transposed blocks at nu = eta (the clamped and mean
  // blocks of these sets only ever see the zero set)
  for (int b = 0; b < 2; ++b)
    for (int p = 0; p < 2; ++p)
      if ((r = h->alloc(&A.fA_t.blk[b][p], pcf::factor_doubles(g, A.smt, b, p)))) return r;
  for (int p = 0; p < 2; ++p)
    if ((r = h->alloc(&A.fA_t.mean[p], pcf::factor_doubles(g, A.smt, -1, p)))) return r;
  hipError_t e = pcf::launch_factor(h->geo_eta, A.smt, h->cfg.dt * 6.0 / 11.0, A.fA_t, false, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("factorisation: ") + hipGetErrorString(e));
  const int n0 = g.n0, nd = g.nd;
  std::vector<double> ewag((size_t)n0 * nd), vdt((size_t)nd * n0), vd1t((size_t)nd * n0), vq1t((size_t)nd * n0),
      vq2t((size_t)nd * n0);
  for (int x = 0; x < n0; ++x)
    for (int i = 0; i < nd; ++i) {
      ewag[(size_t)x * nd + i] = g.W * o.AG(i, x);
      vdt[(size_t)i * n0 + x] = o.VD(x, i);
      vd1t[(size_t)i * n0 + x] = o.VD1(x, i);
      vq1t[(size_t)i * n0 + x] = o.VQ1(x, i);
      vq2t[(size_t)i * n0 + x] = o.VQ2(x, i);
    }
  std::vector<double> mat((size_t)nd * nd), c2t((size_t)nd * nd);
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j) {
      mat[(size_t)i * nd + j] = g.W * o.C0D(j, i);
      c2t[(size_t)i * nd + j] = g.W * o.C2D(j, i);
    }
  if ((r = h->upload(&A.mEWAG, ewag)) || (r = h->upload(&A.mVDT, vdt)) || (r = h->upload(&A.mVD1T, vd1t)) ||
      (r = h->upload(&A.mVQ1T, vq1t)) || (r = h->upload(&A.mVQ2T, vq2t)) || (r = h->upload(&A.mMAT, mat)) ||
      (r = h->upload(&A.mC2T, c2t)))
    return r;
  A.mhd_ready = true;
  return PCF_OK;
}

pcf::EvalTerm eterm(const double *M, const double2 *F, int K, int kf, double c) {
  pcf::EvalTerm t;
  t.M = M;
  t.F = F;
  t.K = K;
  t.kf = kf;
  t.c = c;
  return t;
}

struct ATerm {
  const double2 *p;
  double w;
};

// Mass terms of an A cotangent: (M_A + gamma L_A)^T rstage (when rstage is
// set) + M_A^T sum w_t p_t, with M_A = W C0D and L_A = eta W (C2D - k^2 C0D)
// the Dirichlet blocks at nu = eta.
struct AMass {
  double gamma = 0.0;
  const double2 *rstage = nullptr;
  int n = 0;
  ATerm t[3];
};

AMass amass(double gamma, const double2 *rstage, std::initializer_list<ATerm> gather) {
  AMass m;
  m.gamma = gamma;
  m.rstage = rstage;
  for (const ATerm &t : gather) m.t[m.n++] = t;
  return m;
}

// Appends the mass terms of component c to an evaluation.
void add_mass_terms(const pcf_handle *h, const AMass &ms, int c, pcf::EvalOut &o) {
  const auto &A = h->adj;
  const int nd = h->geo.nd;
  const size_t off = c * h->cstride;
  const double gn = ms.gamma * h->geo_eta.nu;
  if (ms.rstage) {
    o.t[o.nt++] = eterm(A.mMAT, ms.rstage + off, nd, pcf::KF_ONE, 1.0);
    if (gn != 0.0) {
      o.t[o.nt++] = eterm(A.mMAT, ms.rstage + off, nd, pcf::KF_K2, -gn);
      o.t[o.nt++] = eterm(A.mC2T, ms.rstage + off, nd, pcf::KF_ONE, gn);
    }
  }
  for (int t = 0; t < ms.n; ++t) o.t[o.nt++] = eterm(A.mMAT, ms.t[t].p + off, nd, pcf::KF_ONE, ms.t[t].w);
}

// out = the mass terms alone (three components cstride apart).
hipError_t a_mass(pcf_handle *h, const AMass &ms, double2 *out) {
  pcf::EvalOut o[3];
  for (int c = 0; c < 3; ++c) {
    o[c].dst = out + c * h->cstride;
    o[c].rows = h->geo.nd;
    o[c].nt = 0;
    add_mass_terms(h, ms, c, o[c]);
  }
  return pcf::launch_eval_terms(h->geo, o, 3, h->adj.stream);
}

// mu_A = (b, j map)^H (mu_b, mu_j): the transpose of the nodal b and j of
// k_nodal_mhd_mfma (conducting walls), from nodal cotangents bj = [6][n0][m]
// (bx, by, bz, jx, jy, jz) into out (three components cstride apart), plus
// the mass terms ms when given (the history gather, and the stage term of a
// start-up stage).
hipError_t a_nodal_adj(pcf_handle *h, const double2 *bj, double2 *out, const AMass *ms = nullptr) {
  const auto &A = h->adj;
  const pcf::Geo &g = h->geo;
  const size_t pl = (size_t)g.n0 * g.m, cs = h->cstride;
  const double2 *bx = bj, *by = bj + pl, *bz = bj + 2 * pl, *jx = bj + 3 * pl, *jy = bj + 4 * pl, *jz = bj + 5 * pl;
  const int K = g.n0;
  using namespace pcf;
  EvalOut o[3];
  const EvalTerm tx[5] = {eterm(A.mVDT, by, K, KF_IKZ, -1.0), eterm(A.mVDT, bz, K, KF_IKY, 1.0),
                          eterm(A.mVDT, jx, K, KF_K2, 1.0), eterm(A.mVQ1T, jy, K, KF_IKY, -1.0),
                          eterm(A.mVQ1T, jz, K, KF_IKZ, -1.0)};
  const EvalTerm ty[6] = {eterm(A.mVDT, bx, K, KF_IKZ, 1.0), eterm(A.mVDT, jy, K, KF_KZZ, 1.0),
                          eterm(A.mVDT, jz, K, KF_KYZ, -1.0), eterm(A.mVD1T, bz, K, KF_ONE, 1.0),
                          eterm(A.mVD1T, jx, K, KF_IKY, -1.0), eterm(A.mVQ2T, jy, K, KF_ONE, -1.0)};
  const EvalTerm tz[6] = {eterm(A.mVDT, bx, K, KF_IKY, -1.0), eterm(A.mVDT, jy, K, KF_KYZ, -1.0),
                          eterm(A.mVDT, jz, K, KF_KYY, 1.0), eterm(A.mVD1T, by, K, KF_ONE, -1.0),
                          eterm(A.mVD1T, jx, K, KF_IKZ, -1.0), eterm(A.mVQ2T, jz, K, KF_ONE, -1.0)};
  const EvalTerm *ts[3] = {tx, ty, tz};
  const int nt[3] = {5, 6, 6};
  for (int c = 0; c < 3; ++c) {
    o[c].dst = out + c * cs;
    o[c].rows = g.nd;
    o[c].nt = nt[c];
    for (int t = 0; t < nt[c]; ++t) o[c].t[t] = ts[c][t];
    if (ms) add_mass_terms(h, *ms, c, o[c]);
  }
  return launch_eval_terms(g, o, 3, A.stream);
}

// Cotangents of u (and curl u), b and j of the coupled nonlinear rows from
// adj.nu (flow rows) and adj.nuA (induction rows) at the evaluation whose
// padded y-line fields (u, omega, b, j) are ytape: the flow part into
// adj.muX (steady_x: with the reconstruction adjoint and M^T sig folded in)
// or adj.mu_u, the vector-potential part into adj.muA (overwritten).
hipError_t nl_adj_mhd(pcf_handle *h, const void *ytape, bool steady_x, const AMass *ms) {
  auto &A = h->adj;
  const pcf::Geo &g = h->geo;
  const size_t pl = (size_t)g.n0 * g.m;
  hipError_t e;
  if ((e = pcf::launch_rhs_adj_mfma<double>(g, A.ad, A.nu, A.ws, A.stream))) return e;
  double2 *Hx = reinterpret_cast<double2 *>(A.ws.Hx);
  pcf::EvalOut o[3];
  for (int c = 0; c < 3; ++c) {
    o[c].dst = Hx + (3 + c) * pl;
    o[c].rows = g.n0;
    o[c].nt = 1;
    o[c].t[0] = eterm(A.mEWAG, A.nuA + c * h->cstride, g.nd, pcf::KF_ONE, 1.0);
  }
  if ((e = pcf::launch_eval_terms(g, o, 3, A.stream))) return e;
  if ((e = pcf::launch_fft_y_fields<double>(g, h->md.view, h->py, true, 6, A.ws.Hx, A.ws.Hz, A.stream))) return e;
  if ((e = pcf::launch_rot_adj_mhd<double>(g, h->md.view, ytape, A.ws, h->pz, A.stream))) return e;
  if ((e = pcf::launch_fft_y_fields<double>(g, h->md.view, h->py, false, 12, A.ws.Y, A.ws.X, A.stream))) return e;
  if (steady_x)
    e = pcf::launch_nodal_adj_x_mfma<double>(g, A.ad, h->sm, cprim(A.sig), A.ws, A.muX, A.stream);
  else
    e = pcf::launch_nodal_adj_mfma<double>(g, A.ad, A.ws, A.mu_u[0], A.mu_u[1], A.mu_u[2], A.stream);
  if (e != hipSuccess) return e;
  return a_nodal_adj(h, reinterpret_cast<const double2 *>(A.ws.X) + 6 * pl, A.muA, ms);
}

// Transposed A solves of the steady systems: out = S_A^{-T} in with
// nuA = e0 out + e1 r1 + e2 r2 and sigA = C_SBDF . (out, r1, r2).
hipError_t a_solve_factored(pcf_handle *h, const double2 *in, double2 *out, const double2 *r1, const double2 *r2,
                            const double e[3]) {
  auto &A = h->adj;
  return pcf::launch_adj_solve_factored_dir(h->geo_eta, A.smt, A.fA_t, a_in(h, in), a_out(h, out), a_out(h, A.nuA),
                                            a_in(h, r1), a_in(h, r2), e[0], e[1], e[2], A.stream,
                                            a_out(h, A.sigA), C_SBDF[0], C_SBDF[1], C_SBDF[2], 3, h->cstride);
}

// out = (M_A - gamma L_A)^{-T} in (start-up stages, factorised per launch),
// the three components in one launch.
hipError_t a_solve(pcf_handle *h, double gamma, const double2 *in, double2 *out) {
  auto &A = h->adj;
  return pcf::launch_adj_solve_dir(h->geo_eta, A.smt, gamma, in, out, A.scratchA, 3, h->cstride, A.stream);
}

// out = sum w_t p_t.
hipError_t a_sum(pcf_handle *h, std::initializer_list<ATerm> terms, double2 *out) {
  auto &A = h->adj;
  for (int c = 0; c < 3; ++c) {
    const size_t o = c * h->cstride;
    pcf::PrimSum s{};
    for (const ATerm &t : terms) {
      s.t[s.n] = a_in(h, t.p + o);
      s.wgt[s.n] = t.w;
      ++s.n;
    }
    hipError_t er = pcf::launch_prim_sum(h->geo, s, a_out(h, out + o), A.stream);
    if (er != hipSuccess) return er;
  }
  return hipSuccess;
}

hipError_t a_zero(pcf_handle *h, double2 *p) {
  return hipMemsetAsync(p, 0, 3 * h->cstride * sizeof(double2), h->adj.stream);
}

// Reverse of MHD step k (conducting walls): the flow as reverse_step, the
// vector potential beside it (examples/pcf_mhd_jax.py:488-715 transposed).
// mu of the step's output in adj.muX / adj.muA on entry, of its input on exit.
hipError_t reverse_step_mhd(pcf_handle *h, int k, bool steady, void *const *tape) {
  auto &A = h->adj;
  const pcf::Geo &g = h->geo;
  const double dt = h->cfg.dt;
  const double e[3] = {dt * 18.0 / 11.0, -dt * 18.0 / 11.0, dt * 6.0 / 11.0};
  int *rr = A.rr;
  {
    const int t = rr[2];
    rr[2] = rr[1];
    rr[1] = rr[0];
    rr[0] = t;
  }
  hipError_t er;
  if (steady) {
    if ((er = pcf::launch_adj_solve_factored(g, A.smt, A.fsteady_t, cprim(A.muX), A.rho[rr[0]], A.nu,
                                             cprim(A.rho[rr[1]]), cprim(A.rho[rr[2]]), e[0], e[1], e[2],
                                             A.stream, A.sig, C_SBDF[0], C_SBDF[1], C_SBDF[2])))
      return er;
    if ((er = a_solve_factored(h, A.muA, A.rhoA[rr[0]], A.rhoA[rr[1]], A.rhoA[rr[2]], e))) return er;
    // mu_A = (b, j map)^H mu_bj + M_A^T sig_A in one evaluation
    const AMass ms = amass(0.0, nullptr, {{A.sigA, 1.0}});
    if (k == 0) {  // initial state: its velocity leaves keep mu_u
      if ((er = nl_adj_mhd(h, tape[0], false, &ms))) return er;
      pcf::PrimSum ga{};
      for (int j = 0; j < 3; ++j) ps_add(ga, A.rho[rr[j]], C_SBDF[j]);
      if ((er = zero_prim(h, A.muX))) return er;
      return pcf::launch_adj_mass(g, h->sm, 0.0, pcf::PrimC{}, ga, A.muX, A.stream);
    }
    return nl_adj_mhd(h, tape[0], true, &ms);
  }
  if ((er = zero_prim(h, A.rho[rr[0]])) || (er = a_zero(h, A.rhoA[rr[0]]))) return er;
  for (int s = 2; s >= 0; --s) {
    const double gamma = (IMEX_A[s] + IMEX_B[s]) * dt / 2.0;
    if ((er = pcf::launch_adj_solve(g, A.smt, gamma, cprim(A.muX), A.rs[s], A.scratch, A.stream))) return er;
    if ((er = a_solve(h, gamma, A.muA, A.rsA[s]))) return er;
    pcf::PrimSum nu{};
    ps_add(nu, A.rs[s], dt * IMEX_A[s]);
    if (s < 2) ps_add(nu, A.rs[s + 1], dt * IMEX_B[s + 1]);
    if (s == 0) {
      ps_add(nu, A.rho[rr[1]], e[1]);
      ps_add(nu, A.rho[rr[2]], e[2]);
    }
    if ((er = pcf::launch_prim_sum(g, nu, A.nu, A.stream))) return er;
    if (s == 2)
      er = a_sum(h, {{A.rsA[2], dt * IMEX_A[2]}}, A.nuA);
    else if (s == 1)
      er = a_sum(h, {{A.rsA[1], dt * IMEX_A[1]}, {A.rsA[2], dt * IMEX_B[2]}}, A.nuA);
    else
      er = a_sum(h, {{A.rsA[0], dt * IMEX_A[0]}, {A.rsA[1], dt * IMEX_B[1]}, {A.rhoA[rr[1]], e[1]},
                     {A.rhoA[rr[2]], e[2]}}, A.nuA);
    if (er) return er;
    // mu_A = (b, j map)^H mu_bj + (M_A + gamma L_A)^T rs_A (+ the history gather at s = 0)
    const AMass ms = (s == 0) ? amass(gamma, A.rsA[s], {{A.rhoA[rr[1]], C_SBDF[1]}, {A.rhoA[rr[2]], C_SBDF[2]}})
                              : amass(gamma, A.rsA[s], {});
    if ((er = nl_adj_mhd(h, tape[s], false, &ms))) return er;
    if ((er = recon_or_zero(h, s == 0 && k == 0))) return er;
    pcf::PrimSum ga{};
    if (s == 0) {
      ps_add(ga, A.rho[rr[1]], C_SBDF[1]);
      ps_add(ga, A.rho[rr[2]], C_SBDF[2]);
    }
    if ((er = pcf::launch_adj_mass(g, h->sm, gamma, cprim(A.rs[s]), ga, A.muX, A.stream))) return er;
  }
  return hipSuccess;
}

// pcf_vjp for the conducting-wall MHD step: the same checkpointing as the
// hydrodynamic reverse pass, with the vector-potential cotangents beside the
// flow's.  The recompute of segment seg - 1 (eager steps on the handle
// stream) overlaps the reverse of segment seg (reverse stream) through the
// two tape sets, as in pcf_vjp.  Seed: Epert + Emag of the final state, or
// the cotangents set.
int32_t vjp_mhd(pcf_handle *h, int steps, int every, int seed) {
  auto &A = h->adj;
  const pcf::Geo &g = h->geo;
  const int nseg = (steps + every - 1) / every;
  const int h0 = (int)std::floor(std::min(2.0, std::max(0.0, h->history_steps)));
  const double dt = h->cfg.dt;
  const size_t pl = (size_t)g.n0 * g.m;
  A.have_grad = false;
  auto hip_fail = [&](hipError_t e, const char *what) {
    return fail(PCF_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e;
  // segment seg's steps on the handle stream, keeping the tape in set seg & 1
  // when keep is set; first[i] receives the tape slot of step ka + i
  auto run_seg = [&](int seg, bool keep, std::vector<int> &first) -> int32_t {
    const int ka = seg * every, kb = std::min(steps, ka + every);
    first.assign(kb - ka, 0);
    int slot = (seg & 1) * A.per_set;
    for (int k = ka; k < kb; ++k) {
      first[k - ka] = slot;
      const bool st = h->history_steps >= 2.0;
      void *const *tape = keep ? &A.tape[slot] : nullptr;
      int32_t r = st ? steady_step(h, tape) : startup_step(h, tape);
      if (r) return r;
      slot += st ? 1 : 3;
    }
    return PCF_OK;
  };
  std::vector<int> first_rev, first_next;
  int32_t r;
  for (int seg = 0; seg < nseg; ++seg) {
    if ((e = snap_save(h, A.snaps[seg]))) return hip_fail(e, "snapshot");
    if ((r = run_seg(seg, seg == nseg - 1, first_rev))) return r;
  }
  pcf_handle::Snap &fin = A.snaps[nseg];
  if ((e = snap_save(h, fin))) return hip_fail(e, "snapshot");
  if ((e = ensure_u12(h))) return hip_fail(e, "reconstruction");
  PCF_HIP(hipEventRecord(A.ev_seed, h->stream));
  PCF_HIP(hipStreamWaitEvent(A.stream, A.ev_seed, 0));
  const size_t bytesA = 3 * h->cstride * sizeof(double2);
  if (seed == PCF_SEED_ENERGY) {
    const double scale = 2.0 * h->cfg.ly * h->cfg.lz;
    e = pcf::launch_energy_seed(g, A.QB, A.QD, scale, h->u0[h->so[0]], h->u1, h->u2, A.mu_u[0], A.mu_u[1],
                                A.mu_u[2], A.stream);
    if (e == hipSuccess)
      e = pcf::launch_recon_adj(g, A.RcT, A.ldrT, A.mu_u[0], A.mu_u[1], A.mu_u[2], nullptr, A.muX, A.stream);
    // Emag: b and j of the final A at the nodes, mu_b = 2 Ly Lz fejer b, mu_j = 0
    if (e == hipSuccess)
      e = pcf::launch_nodal_mhd_mfma<double>(g, h->md.mview, h->A[h->so[0]], h->cstride, A.ws, A.stream);
    double2 *bj = reinterpret_cast<double2 *>(A.ws.X) + 6 * pl;
    if (e == hipSuccess) e = pcf::launch_seed_b(g, h->dfejer, scale, bj, A.stream);
    if (e == hipSuccess) e = a_nodal_adj(h, bj, A.muA);
  } else {
    e = pcf::launch_recon_adj(g, A.RcT, A.ldrT, A.cot_u[0], A.cot_u[1], A.cot_u[2], A.cot_g, A.muX, A.stream);
    if (e == hipSuccess)
      e = A.cotA_set ? hipMemcpyAsync(A.muA, A.cotA, bytesA, hipMemcpyDeviceToDevice, A.stream) : a_zero(h, A.muA);
  }
  if (e != hipSuccess) return hip_fail(e, "seed");
  for (int i = 0; i < 3; ++i) {
    if ((e = zero_prim(h, A.rho[i])) || (e = a_zero(h, A.rhoA[i]))) return hip_fail(e, "memset");
    A.rr[i] = i;
  }
  // the recompute overwrites the state the seed reads
  PCF_HIP(hipEventRecord(A.ev_seed, A.stream));
  PCF_HIP(hipStreamWaitEvent(h->stream, A.ev_seed, 0));
  // tape set b is rewritten only after the reverse that read it (ev_rev[b])
  // and read only after the recompute that wrote it (ev_rc[b])
  bool rev_pending[2] = {false, false};
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int ka = seg * every, kb = std::min(steps, ka + every);
    if (seg < nseg - 1) {
      first_rev.swap(first_next);
      PCF_HIP(hipStreamWaitEvent(A.stream, A.ev_rc[seg & 1], 0));
    }
    if (seg > 0) {
      const int b = (seg - 1) & 1;
      if (rev_pending[b]) PCF_HIP(hipStreamWaitEvent(h->stream, A.ev_rev[b], 0));
      if ((e = snap_load(h, A.snaps[seg - 1]))) return hip_fail(e, "restore");
      if ((r = run_seg(seg - 1, true, first_next))) return r;
      PCF_HIP(hipEventRecord(A.ev_rc[b], h->stream));
    }
    for (int k = kb - 1; k >= ka; --k)
      if ((e = reverse_step_mhd(h, k, h0 + k >= 2, &A.tape[first_rev[k - ka]]))) return hip_fail(e, "reverse step");
    PCF_HIP(hipEventRecord(A.ev_rev[seg & 1], A.stream));
    rev_pending[seg & 1] = true;
  }
  // history leaves of the initial state: levels -1 and -2 (flow as pcf_vjp, then A)
  {
    const double e1 = -dt * 18.0 / 11.0, e2 = dt * 6.0 / 11.0;
    const pcf::Prim &r0 = A.rho[A.rr[0]], &r1 = A.rho[A.rr[1]];
    pcf::PrimSum s{};
    for (int i = 0; i < 2; ++i)
      if ((e = zero_prim(h, A.gh[i]))) return hip_fail(e, "memset");
    ps_add(s, r0, C_SBDF[1]);
    ps_add(s, r1, C_SBDF[2]);
    if ((e = pcf::launch_adj_mass(g, h->sm, 0.0, pcf::PrimC{}, s, A.gh[0], A.stream))) return hip_fail(e, "history");
    s = pcf::PrimSum{};
    ps_add(s, r0, C_SBDF[2]);
    if ((e = pcf::launch_adj_mass(g, h->sm, 0.0, pcf::PrimC{}, s, A.gh[1], A.stream))) return hip_fail(e, "history");
    s = pcf::PrimSum{};
    ps_add(s, r0, e1);
    ps_add(s, r1, e2);
    if ((e = pcf::launch_prim_sum(g, s, A.gh[2], A.stream))) return hip_fail(e, "history");
    s = pcf::PrimSum{};
    ps_add(s, r0, e2);
    if ((e = pcf::launch_prim_sum(g, s, A.gh[3], A.stream))) return hip_fail(e, "history");
    const double2 *a0 = A.rhoA[A.rr[0]], *a1 = A.rhoA[A.rr[1]];
    if ((e = a_mass(h, amass(0.0, nullptr, {{a0, C_SBDF[1]}, {a1, C_SBDF[2]}}), A.ghA[0])) ||
        (e = a_mass(h, amass(0.0, nullptr, {{a0, C_SBDF[2]}}), A.ghA[1])) ||
        (e = a_sum(h, {{a0, e1}, {a1, e2}}, A.ghA[2])) ||
        (e = a_sum(h, {{a0, e2}}, A.ghA[3])))
      return hip_fail(e, "history");
  }
  PCF_HIP(hipStreamSynchronize(A.stream));
  if ((e = snap_load(h, fin))) return hip_fail(e, "restore");
  PCF_HIP(hipStreamSynchronize(h->stream));
  A.have_grad = true;
  return PCF_OK;
}

// Cotangent of a full-spectrum jax gradient g into the half layout:
// mu(k) = (conj g(k) + g(-k)) / 2 for every stored k.
void half_from_cotangent(const pcf::Geo &g, int rows, const double *full, std::vector<double2> &half) {
  half.assign((size_t)rows * g.m, double2{0, 0});
  for (int n = 0; n < rows; ++n)
    for (int iy = 0; iy < g.ny; ++iy)
      for (int kz = 0; kz < g.kzs; ++kz) {
        const int my_ = (g.ny - iy) % g.ny, mz_ = (g.nz - kz) % g.nz;
        const double *a = full + (((size_t)n * g.ny + iy) * g.nz + kz) * 2;
        const double *b = full + (((size_t)n * g.ny + my_) * g.nz + mz_) * 2;
        half[((size_t)n * g.ny + iy) * g.kzs + kz] = double2{0.5 * (a[0] + b[0]), 0.5 * (-a[1] + b[1])};
      }
}

// Half-layout cotangent -> jax gradient on the full spectrum: conj, then the
// Hermitian expansion; spanwise Nyquist plane zero.
void full_from_half_grad(const pcf::Geo &g, int rows, const std::vector<double2> &half, double *full) {
  for (int n = 0; n < rows; ++n)
    for (int iy = 0; iy < g.ny; ++iy) {
      double *row = full + ((size_t)n * g.ny + iy) * g.nz * 2;
      const double2 *hr = half.data() + ((size_t)n * g.ny + iy) * g.kzs;
      const double2 *mr = half.data() + ((size_t)n * g.ny + (g.ny - iy) % g.ny) * g.kzs;
      for (int kz = 0; kz < g.nz; ++kz) {
        double re = 0, im = 0;
        if (kz < g.kzs) {
          re = hr[kz].x;
          im = -hr[kz].y;
        } else if (kz > g.kzs) {
          re = mr[g.nz - kz].x;
          im = mr[g.nz - kz].y;
        }
        row[2 * kz] = re;
        row[2 * kz + 1] = im;
      }
    }
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
  h->mhd = cfg->mhd != 0;
  h->insul = h->mhd && cfg->walls == PCF_WALLS_INSULATING;
  h->arows = h->insul ? g.n0 : g.nd;
  for (int c = 0; c < 3; ++c) g.b0[c] = h->mhd ? cfg->b0[c] : 0.0;
  g.omega2 = h->mhd ? 2.0 * cfg->omega : 0.0;
  if (g.my % 2) return bail(fail(PCF_E_INVALID, "padded streamwise grid must be even"));
  if (g.mz < 2 * g.kzs) return bail(fail(PCF_E_INVALID, "spanwise padding must not truncate"));
  if (!make_plan(g.my, h->py) || !make_plan(g.mz, h->pz))
    return bail(fail(PCF_E_INVALID, "padded grid sizes must factor into 2, 3, 4"));
  if (!pcf::fft_fits(g, h->py, h->pz, cfg->precision == PCF_PREC_FP64 ? 16 : 8, h->mhd))
    return bail(fail(PCF_E_INVALID, "FFT sizes exceed the kernels' register budget"));
  {
    const int eb = cfg->precision == PCF_PREC_FP64 ? 16 : 8;
    h->geo0 = g;
    h->geo0.my = g.ny;
    h->geo0.mz = g.nz;
    h->maxima_ok = make_plan(g.ny, h->py0) && make_plan(g.nz, h->pz0) &&
                   pcf::fft_fits(h->geo0, h->py0, h->pz0, eb, h->mhd) &&
                   pcf::absmax_fits(h->geo0, h->pz0, h->mhd ? 6 : 3, eb);
  }

  h->device = cfg->device;
  if (hipSetDevice(cfg->device) != hipSuccess) return bail(fail(PCF_E_HIP, "hipSetDevice failed"));
  {
    double *sp = nullptr;  // the no-alloc scratch of this device, looked up at set-up
    if (int32_t r0 = scratch_on_device(cfg->device, &sp)) return bail(r0);
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(PCF_E_HIP, "hipStreamCreate failed"));
  h->own_stream = true;
  // the side stream pays with insulating walls (their bordered A_y/A_z solve
  // runs beside the flow solve): 5 890 -> 6 220 steps/s at 65x64x64; with
  // conducting walls the cross-stream edges cost more than the overlap gains
  // (6 830 -> 6 480 steps/s), so those run on one stream
  if (h->insul && (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess ||
                 hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
                 hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess))
    return bail(fail(PCF_E_HIP, "side stream / event create failed"));

  int32_t r;
  const size_t M = g.m;
  {
    // the whole state in one block, laid out as state_pieces() lists it (each
    // piece 256-byte aligned), so that a snapshot is a single copy
    const size_t c = sizeof(double2), rb = sizeof(double);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t bu0 = al(g.nb * M * c), bg = al(g.nd * M * c), bv = al(g.nd * rb);
    h->state_bytes = 3 * (bu0 + bg + 2 * bv) + 4 * (bu0 + bg + 2 * bv) + 2 * bg;
    unsigned char *blk = nullptr;
    if ((r = h->alloc(&blk, h->state_bytes))) return bail(r);
    h->state_block = blk;
    size_t off = 0;
    auto take = [&](size_t bytes) {
      unsigned char *q = blk + off;
      off += bytes;
      return q;
    };
    for (int s = 0; s < 3; ++s) {
      h->u0[s] = reinterpret_cast<double2 *>(take(bu0));
      h->g[s] = reinterpret_cast<double2 *>(take(bg));
      h->v[s] = reinterpret_cast<double *>(take(bv));
      h->w[s] = reinterpret_cast<double *>(take(bv));
    }
    for (int s = 0; s < 4; ++s) {
      h->Nu[s] = reinterpret_cast<double2 *>(take(bu0));
      h->Ng[s] = reinterpret_cast<double2 *>(take(bg));
      h->Nv[s] = reinterpret_cast<double *>(take(bv));
      h->Nw[s] = reinterpret_cast<double *>(take(bv));
    }
    h->u1 = reinterpret_cast<double2 *>(take(bg));
    h->u2 = reinterpret_cast<double2 *>(take(bg));
  }
  if (h->mhd) {
    h->cstride = (size_t)h->arows * M;
    for (int s = 0; s < 3; ++s)
      if ((r = h->alloc(&h->A[s], 3 * h->cstride))) return bail(r);
    for (int s = 0; s < 4; ++s)
      if ((r = h->alloc(&h->NA[s], 3 * h->cstride))) return bail(r);
  }
  const size_t ve = (cfg->precision == PCF_PREC_FP64) ? sizeof(double2) : sizeof(float2);
  unsigned char *wsp = nullptr;
  const size_t fin = h->mhd ? 12 : 6, fout = h->mhd ? 6 : 3;
  const size_t nX = fin * g.n0 * M, nY = fin * g.n0 * g.my * g.kzs;
  const size_t nHz = fout * g.n0 * g.my * g.kzs, nHx = fout * g.n0 * M;
  if ((r = h->alloc(&wsp, (nX + nY + nHz + nHx) * ve))) return bail(r);
  h->ws.X = wsp;
  h->ws.Y = wsp + nX * ve;
  h->ws.Hz = wsp + (nX + nY) * ve;
  h->ws.Hx = wsp + (nX + nY + nHz) * ve;

  if (cfg->precision == PCF_PREC_FP64) {
    if ((r = build_nodal<double>(h, h->md, 16, 16))) return bail(r);
  } else {
    if ((r = build_nodal<float>(h, h->mf, 16, 16))) return bail(r);
  }
  if ((r = build_struct_mats(h))) return bail(r);
  {
    hipError_t e = pcf::mfma_init(h->geo, cfg->precision == PCF_PREC_FP64);
    if (e != hipSuccess)
      return bail(fail(PCF_E_UNSUPPORTED, std::string("kernel LDS limits: ") + hipGetErrorString(e)));
  }
  const int ti = cfg->time_integrator;
  const bool stages = (ti == PCF_TI_SBDF3 || ti == PCF_TI_IMEXRK3);  // IMEXRK3 stages run
  if ((r = build_factors(h, h->sm, cfg->dt * 6.0 / 11.0, h->fsteady))) return bail(r);
  if (ti == PCF_TI_IMEXRK222 && (r = build_factors(h, h->sm, cfg->dt * ARS_G, h->fars))) return bail(r);
  for (int rk = 0; stages && rk < 3; ++rk)
    if ((r = build_factors(h, h->sm, stage_gamma(rk, cfg->dt), h->fstage[rk]))) return bail(r);
  if (ti == PCF_TI_CNAB2 && (r = build_factors(h, h->sm, 0.5 * cfg->dt, h->fcn))) return bail(r);
  if (h->mhd) {
    h->geo_eta = h->geo;
    h->geo_eta.nu = cfg->eta;
    if ((r = build_factors_A(h, cfg->dt * 6.0 / 11.0, h->fA))) return bail(r);
    if (ti == PCF_TI_IMEXRK222 && (r = build_factors_A(h, cfg->dt * ARS_G, h->fA_ars))) return bail(r);
    for (int rk = 0; stages && rk < 3; ++rk)
      if ((r = build_factors_A(h, stage_gamma(rk, cfg->dt), h->fA_stage[rk]))) return bail(r);
    if (ti == PCF_TI_CNAB2 && (r = build_factors_A(h, 0.5 * cfg->dt, h->fA_cn))) return bail(r);
    if ((r = h->alloc(&h->ddiagm, 5))) return bail(r);
    if (h->insul) {
      if ((r = h->upload(&h->dVC, h->ops.VC.a))) return bail(r);
      if ((r = h->upload(&h->dVC1, h->ops.VC1.a))) return bail(r);
      if ((r = h->alloc(&h->dresid, (size_t)pcf::insul_residual_blocks(g)))) return bail(r);
    }
  }
  if ((r = h->upload(&h->dVB, h->ops.VB.a))) return bail(r);
  if ((r = h->upload(&h->dVD, h->ops.VD.a))) return bail(r);
  if ((r = h->upload(&h->dVB1, h->ops.VB1.a))) return bail(r);
  if ((r = h->upload(&h->dVD1, h->ops.VD1.a))) return bail(r);
  if ((r = h->upload(&h->dfejer, h->ops.fejer))) return bail(r);
  if ((r = h->upload(&h->dxn, h->ops.x))) return bail(r);
  if ((r = h->alloc(&h->ddiag, 6))) return bail(r);
  if ((r = h->alloc(&h->dflag, 1))) return bail(r);
  if (h->maxima_ok) {
    const int nf = h->mhd ? 6 : 3, eb = cfg->precision == PCF_PREC_FP64 ? 16 : 8;
    if ((r = h->alloc(&h->dblockmax, (size_t)pcf::absmax_blocks(h->geo0, nf, eb) * nf))) return bail(r);
  }
  *out = h;
  return PCF_OK;
}

int32_t pcf_destroy(pcf_handle *h) {
  if (!h) return PCF_OK;
  if (h->adj.stream) (void)hipStreamSynchronize(h->adj.stream);
  for (hipEvent_t ev : {h->adj.ev_seed, h->adj.ev_rc[0], h->adj.ev_rc[1], h->adj.ev_rev[0], h->adj.ev_rev[1]})
    if (ev) (void)hipEventDestroy(ev);
  if (h->adj.stream) (void)hipStreamDestroy(h->adj.stream);
  if (h->side) (void)hipStreamSynchronize(h->side);
  for (hipEvent_t ev : {h->ev_fork, h->ev_join})
    if (ev) (void)hipEventDestroy(ev);
  if (h->side) (void)hipStreamDestroy(h->side);
  for (void *p : h->adj_allocs) (void)hipFree(p);
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto &kv : h->seg_graphs) (void)hipGraphExecDestroy(kv.second);
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
  h->graph_evictions += (int64_t)(h->graphs.size() + h->seg_graphs.size());
  ++h->graph_generation;
  for (auto &kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  h->graphs.clear();
  for (auto &kv : h->seg_graphs) (void)hipGraphExecDestroy(kv.second);
  h->seg_graphs.clear();
  return PCF_OK;
}

int32_t pcf_graph_cache_info(pcf_handle *h, int64_t *out, int32_t n) {
  if (!h || !out) return fail(PCF_E_INVALID, "null argument");
  if (n < 5) return fail(PCF_E_INVALID, "the counters need at least 5 slots");
  std::vector<int64_t> lengths;
  for (const auto *c : {&h->graphs, &h->seg_graphs})
    for (const auto &kv : *c) lengths.push_back((int64_t)((kv.first >> 16) & 0xFFFFFF));
  std::sort(lengths.begin(), lengths.end());
  lengths.erase(std::unique(lengths.begin(), lengths.end()), lengths.end());
  out[0] = h->graph_generation;
  out[1] = (int64_t)(h->graphs.size() + h->seg_graphs.size());
  out[2] = h->graph_hits;
  out[3] = h->graph_misses;
  out[4] = h->graph_evictions;
  for (int32_t i = 5; i < n; ++i) out[i] = (size_t)(i - 5) < lengths.size() ? lengths[i - 5] : 0;
  return (int32_t)lengths.size();
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
  // full (rows, ny, nz) -> half [rows][ny][kzs]: keep the first nz/2 spanwise
  // modes, packed on the host so that one contiguous copy crosses the bus
  {
    const size_t lines = (size_t)fr.rows * g.ny, w = 2 * (size_t)g.kzs;
    std::vector<double> half(lines * w);
    for (size_t r = 0; r < lines; ++r) std::copy(host + r * 2 * g.nz, host + r * 2 * g.nz + w, half.begin() + r * w);
    PCF_HIP(hipMemcpy(fr.ptr, half.data(), half.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (field == PCF_F_U1 || field == PCF_F_U2) {
    // the mean-mode vector of the current level is Re u[:, 0, 0]
    std::vector<double> vec(g.nd);
    const size_t plane = (size_t)g.ny * g.nz * 2;
    for (int i = 0; i < g.nd; ++i) vec[i] = host[i * plane];
    double *dst = (field == PCF_F_U1) ? h->v[h->so[0]] : h->w[h->so[0]];
    PCF_HIP(hipMemcpy(dst, vec.data(), g.nd * sizeof(double), hipMemcpyHostToDevice));
    h->u12_given = true;  // the next evaluation reads the state's own u1, u2
    h->u12_valid = true;
    h->u12_checked = false;
  } else if (field == PCF_F_U0 || field == PCF_F_G) {
    if (!h->u12_given) h->u12_valid = false;  // derived u1, u2 follow the new (u0, g)
    h->u12_checked = false;
  }
  return PCF_OK;
}

int32_t pcf_get_field(pcf_handle *h, int32_t field, double *host) {
  if (!h || !host) return fail(PCF_E_INVALID, "null argument");
  FieldRef fr;
  if (!field_ref(h, field, fr)) return fail(PCF_E_INVALID, "unknown field");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipSetDevice(h->device));
  if (field == PCF_F_U1 || field == PCF_F_U2) PCF_HIP(ensure_u12(h));
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

// Device-memory state exchange in the boundary layout, ordered on the
// handle's stream (no host staging, no synchronisation).
int32_t pcf_set_field_device(pcf_handle *h, int32_t field, const double *dev) {
  if (!h || !dev) return fail(PCF_E_INVALID, "null argument");
  FieldRef fr;
  if (!field_ref(h, field, fr)) return fail(PCF_E_INVALID, "unknown field");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipSetDevice(h->device));
  if (!fr.complex_) {
    PCF_HIP(hipMemcpyAsync(fr.ptr, dev, (size_t)fr.rows * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return PCF_OK;
  }
  PCF_HIP(hipMemcpy2DAsync(fr.ptr, (size_t)g.kzs * sizeof(double2), dev, (size_t)g.nz * sizeof(double2),
                           (size_t)g.kzs * sizeof(double2), (size_t)fr.rows * g.ny, hipMemcpyDeviceToDevice,
                           h->stream));
  if (field == PCF_F_U1 || field == PCF_F_U2) {
    // the mean-mode vector of the current level: Re u[:, 0, 0], one double per plane
    double *dst = (field == PCF_F_U1) ? h->v[h->so[0]] : h->w[h->so[0]];
    PCF_HIP(hipMemcpy2DAsync(dst, sizeof(double), dev, (size_t)g.ny * g.nz * sizeof(double2), sizeof(double),
                             (size_t)g.nd, hipMemcpyDeviceToDevice, h->stream));
    h->u12_given = true;
    h->u12_valid = true;
    h->u12_checked = false;
  } else if (field == PCF_F_U0 || field == PCF_F_G) {
    if (!h->u12_given) h->u12_valid = false;
    h->u12_checked = false;
  }
  return PCF_OK;
}

int32_t pcf_get_field_device(pcf_handle *h, int32_t field, double *dev) {
  if (!h || !dev) return fail(PCF_E_INVALID, "null argument");
  FieldRef fr;
  if (!field_ref(h, field, fr)) return fail(PCF_E_INVALID, "unknown field");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipSetDevice(h->device));
  if (field == PCF_F_U1 || field == PCF_F_U2) PCF_HIP(ensure_u12(h));
  if (!fr.complex_) {
    PCF_HIP(hipMemcpyAsync(dev, fr.ptr, (size_t)fr.rows * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return PCF_OK;
  }
  hipError_t e = pcf::launch_expand_half(reinterpret_cast<const double2 *>(fr.ptr), reinterpret_cast<double2 *>(dev),
                                         fr.rows, g.ny, g.nz, g.kzs, h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("expand launch: ") + hipGetErrorString(e));
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
  if (int32_t r0 = resolve_given(h)) return r0;
  const int ti = h->cfg.time_integrator;
  for (int s = 0; s < nsteps; ++s) {
    int32_t r;
    if (ti == PCF_TI_IMEXRK222) {
      r = ars222_step(h);
    } else if (ti == PCF_TI_CNAB2) {
      r = cnab2_step(h);
    } else if (ti == PCF_TI_IMEXRK3) {
      // the IMEXRK3 stages of the start-up step; the history it writes is scratch
      const double hs = h->history_steps, ho = h->have_old, pd = h->previous_dt;
      r = startup_step(h);
      h->history_steps = hs;
      h->have_old = ho;
      h->previous_dt = pd;
    } else {
      r = (h->history_steps < 2.0) ? startup_step(h) : steady_step(h);
    }
    if (r) return r;
  }
  return PCF_OK;
}

int32_t pcf_set_dt(pcf_handle *h, double dt) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(PCF_E_INVALID, "dt must be positive and finite");
  const int ti = h->cfg.time_integrator;
  if (std::fabs(dt - h->cfg.dt) <= 8.0 * DBL_EPSILON * std::fabs(h->cfg.dt)) return PCF_OK;
  if (ti == PCF_TI_SBDF3)
    return fail(PCF_E_FIXED_DT, "SBDF3 runs a fixed dt; start a fresh solver/history for a different timestep");
  PCF_HIP(hipSetDevice(h->device));
  // the integrator's factor sets in place (storage sized by the geometry alone)
  hipError_t e = hipSuccess;
  auto refactor = [&](double gamma, pcf::FactorSet &fs, pcf::FactorSet &fsA) {
    if (e == hipSuccess) e = pcf::launch_factor(h->geo, h->sm, gamma, fs, false, h->stream);
    if (e == hipSuccess && h->mhd) e = pcf::launch_factor(h->geo_eta, h->sm, gamma, fsA, true, h->stream);
  };
  if (ti == PCF_TI_IMEXRK222) refactor(dt * ARS_G, h->fars, h->fA_ars);
  if (ti == PCF_TI_IMEXRK3)
    for (int rk = 0; rk < 3; ++rk) refactor(stage_gamma(rk, dt), h->fstage[rk], h->fA_stage[rk]);
  if (ti == PCF_TI_CNAB2) refactor(0.5 * dt, h->fcn, h->fA_cn);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("factorisation: ") + hipGetErrorString(e));
  h->cfg.dt = dt;
  return PCF_OK;
}

int32_t pcf_step_steady(pcf_handle *h, int32_t nsteps) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  if (h->cfg.time_integrator != PCF_TI_SBDF3) return fail(PCF_E_UNSUPPORTED, "steady steps are SBDF3 steps");
  if (h->history_steps < 2.0) return fail(PCF_E_INVALID, "history not bootstrapped (history_steps < 2)");
  PCF_HIP(hipSetDevice(h->device));
  if (int32_t r0 = resolve_given(h)) return r0;
  for (int s = 0; s < nsteps; ++s) {
    int32_t r = steady_step(h);
    if (r) return r;
  }
  return PCF_OK;
}

int32_t pcf_step_steady_graph(pcf_handle *h, int32_t nsteps, int32_t replays) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  if (h->cfg.time_integrator != PCF_TI_SBDF3) return fail(PCF_E_UNSUPPORTED, "steady steps are SBDF3 steps");
  if (h->history_steps < 2.0) return fail(PCF_E_INVALID, "history not bootstrapped (history_steps < 2)");
  if (nsteps <= 0 || nsteps > (1 << 20)) return fail(PCF_E_INVALID, "graph length must be in [1, 2^20] steps");
  if (replays < 0) return fail(PCF_E_INVALID, "replays must be >= 0");
  PCF_HIP(hipSetDevice(h->device));
  if (int32_t r0 = resolve_given(h)) return r0;
  // one graph per (ring phase, length): a replay whose length is not a whole
  // number of ring periods (12 steps) leaves the rings at another phase, and so
  // can steps outside the graph (eager, start-up)
  hipGraphExec_t exec = nullptr;
  if (replays == 0) return cached_steady(h, h->graphs, graph_key(h, nsteps, -1), nsteps, -1, &exec);
  for (int rp = 0; rp < replays; ++rp) {
    int32_t r = cached_steady(h, h->graphs, graph_key(h, nsteps, -1), nsteps, -1, &exec);
    if (r) return r;
    PCF_HIP(hipGraphLaunch(exec, h->stream));
    advance_host(h, nsteps);
  }
  return PCF_OK;
}

int32_t pcf_diagnostics(pcf_handle *h, double *out) {
  if (!h || !out) return fail(PCF_E_INVALID, "null argument");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipSetDevice(h->device));
  PCF_HIP(ensure_u12(h));
  PCF_HIP(hipMemsetAsync(h->ddiag, 0, 6 * sizeof(double), h->stream));
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

int32_t pcf_mhd_diagnostics(pcf_handle *h, double *out) {
  if (!h || !out) return fail(PCF_E_INVALID, "null argument");
  if (!h->mhd) return fail(PCF_E_INVALID, "not an MHD handle");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipSetDevice(h->device));
  PCF_HIP(hipMemsetAsync(h->ddiagm, 0, 5 * sizeof(double), h->stream));
  const double2 *A = h->A[h->so[0]];
  hipError_t e = h->insul ? pcf::launch_diag_mhd(g, h->dVD, g.nd, h->dVC, h->dVC1, g.n0, h->dfejer, A, h->cstride,
                                                 h->ddiagm, h->stream)
                          : pcf::launch_diag_mhd(g, h->dVD, g.nd, h->dVD, h->dVD1, g.nd, h->dfejer, A, h->cstride,
                                                 h->ddiagm, h->stream);
  if (e == hipSuccess && h->insul) e = pcf::launch_insul_residual(g, A, h->cstride, h->dresid, h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("diag launch: ") + hipGetErrorString(e));
  double s[5];
  std::vector<double> bm(h->insul ? pcf::insul_residual_blocks(g) : 0);
  PCF_HIP(hipMemcpyAsync(s, h->ddiagm, sizeof(s), hipMemcpyDeviceToHost, h->stream));
  if (h->insul) PCF_HIP(hipMemcpyAsync(bm.data(), h->dresid, bm.size() * sizeof(double), hipMemcpyDeviceToHost,
                                       h->stream));
  PCF_HIP(hipStreamSynchronize(h->stream));
  const double area = h->cfg.ly * h->cfg.lz;
  const double *b0 = g.b0;
  out[0] = area * s[0];
  out[1] = std::sqrt(area * s[1]);
  // int |b + b0|^2 = int |b|^2 + 2 b0 . int b + |b0|^2 V, V = 2 Ly Lz
  out[2] = out[0] + 2.0 * area * (b0[0] * s[2] + b0[1] * s[3] + b0[2] * s[4]) +
           (b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]) * 2.0 * area;
  double r = 0.0;
  for (double v : bm) r = std::max(r, v);
  out[3] = r;
  return PCF_OK;
}

int32_t pcf_field_maxima(pcf_handle *h, double *out) {
  if (!h || !out) return fail(PCF_E_INVALID, "null argument");
  if (!h->maxima_ok)
    return fail(PCF_E_UNSUPPORTED, "field maxima need ny, nz that factor into 2, 3, 4 within the FFT kernels");
  PCF_HIP(hipSetDevice(h->device));
  if (int32_t r0 = resolve_given(h)) return r0;
  const bool fp64 = h->cfg.precision == PCF_PREC_FP64;
  const int nf = h->mhd ? 6 : 3;
  hipError_t e = fp64 ? field_maxima_stages<double>(h, h->md, h->dblockmax)
                      : field_maxima_stages<float>(h, h->mf, h->dblockmax);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("field maxima launch: ") + hipGetErrorString(e));
  std::vector<double> bm((size_t)pcf::absmax_blocks(h->geo0, nf, fp64 ? 16 : 8) * nf);
  PCF_HIP(hipMemcpyAsync(bm.data(), h->dblockmax, bm.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  PCF_HIP(hipStreamSynchronize(h->stream));
  for (int f = 0; f < 6; ++f) out[f] = 0.0;
  for (size_t b = 0; b < bm.size() / nf; ++b)
    for (int f = 0; f < nf; ++f) out[f] = std::max(out[f], bm[b * nf + f]);
  return PCF_OK;
}

int32_t pcf_field_maxima_noalloc(pcf_handle *h, double *out) {
  if (!h || !out) return fail(PCF_E_INVALID, "null argument");
  if (!h->maxima_ok)
    return fail(PCF_E_UNSUPPORTED, "field maxima need ny, nz that factor into 2, 3, 4 within the FFT kernels");
  PCF_HIP(hipSetDevice(h->device));
  double *sp = nullptr;
  if (int32_t r0 = scratch_on_device(h->device, &sp)) return r0;
  const bool fp64 = h->cfg.precision == PCF_PREC_FP64;
  const int nf = h->mhd ? 6 : 3;
  const size_t nblk = (size_t)pcf::absmax_blocks(h->geo0, nf, fp64 ? 16 : 8);
  if (nblk * nf + nf > NOALLOC_SCRATCH_DOUBLES)
    return fail(PCF_E_INVALID, "grid too large for pcf_field_maxima_noalloc: it needs " +
                                   std::to_string(nblk * nf + nf) + " doubles of block maxima, the scratch holds " +
                                   std::to_string(NOALLOC_SCRATCH_DOUBLES));
  if (int32_t r0 = resolve_given(h)) return r0;
  // every block writes its own maxima and the reduction reads only this
  // call's blocks, so nothing left by an earlier call needs clearing
  double *fin = sp + nblk * nf;
  hipError_t e = fp64 ? field_maxima_stages<double>(h, h->md, sp) : field_maxima_stages<float>(h, h->mf, sp);
  if (e == hipSuccess) e = pcf::launch_reduce_blockmax(sp, (int)nblk, nf, fin, h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("field maxima launch: ") + hipGetErrorString(e));
  for (int f = 0; f < 6; ++f) out[f] = 0.0;
  PCF_HIP(hipMemcpyAsync(out, fin, nf * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  PCF_HIP(hipStreamSynchronize(h->stream));
  return PCF_OK;
}

int32_t pcf_velocity_energies(pcf_handle *h, double *out3) {
  if (!h || !out3) return fail(PCF_E_INVALID, "null argument");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipSetDevice(h->device));
  PCF_HIP(ensure_u12(h));
  PCF_HIP(hipMemsetAsync(h->ddiag, 0, 6 * sizeof(double), h->stream));
  hipError_t e = pcf::launch_diag(g, h->dVB, h->dVD, h->dVB1, h->dVD1, h->dfejer, h->dxn, h->u0[h->so[0]], h->u1,
                                  h->u2, h->ddiag, h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("diag launch: ") + hipGetErrorString(e));
  double s[6];
  PCF_HIP(hipMemcpyAsync(s, h->ddiag, sizeof(s), hipMemcpyDeviceToHost, h->stream));
  PCF_HIP(hipStreamSynchronize(h->stream));
  const double LL = h->cfg.ly * h->cfg.lz;
  out3[0] = LL * s[4];
  out3[1] = LL * s[5];
  out3[2] = LL * (s[0] - s[4] - s[5]);
  return PCF_OK;
}

// Nodal-spectral arrays of the extended diagnostics, in this order in XDiag::NS
// (the first ten are also transformed to the physical grid).
enum XdField {
  XD_U = 0,      // u0, u1, u2
  XD_B = 3,      // b = curl A (fluctuation) in b_coeff_spaces
  XD_J = 6,      // curl b, unprojected (_current_physical)
  XD_DIVA = 9,   // div A
  XD_DU = 10,    // x-derivatives of u
  XD_DB = 13,    // x-derivatives of b
  XD_A1 = 16,    // A_y, A_z values (mean-flux traces)
  XD_A2 = 17,
  XD_COUNT = 18,
};

static int32_t xdiag_setup(pcf_handle *h) {
  auto &X = h->xd;
  if (X.ready) return PCF_OK;
  const pcf::Geo &g = h->geo;
  std::string err = pcf::build_diag_operators(g.n0, X.ops);
  if (!err.empty()) return fail(PCF_E_INVALID, "diagnostic operators: " + err);
  const int P = X.ops.P, n0 = g.n0;
  X.P = P;
  auto up = [&](double **dst, const std::vector<double> &src) -> int32_t {
    if (int32_t r = h->alloc(dst, src.size())) return r;
    PCF_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice));
    return PCF_OK;
  };
  const pcf::DiagOps &o = X.ops;
  int32_t r = 0;
  if ((r = up(&X.EB0, o.EB0.a)) || (r = up(&X.EB1, o.EB1.a)) || (r = up(&X.ED0, o.ED0.a)) ||
      (r = up(&X.ED1, o.ED1.a)) || (r = up(&X.ED2, o.ED2.a)) || (r = up(&X.EC0, o.EC0.a)) ||
      (r = up(&X.EC1, o.EC1.a)) || (r = up(&X.EC2, o.EC2.a)) || (r = up(&X.Dn, o.Dn.a)) || (r = up(&X.I, o.I.a)) ||
      (r = up(&X.xp, o.xp)))
    return r;
  // interior quadrature weights fejer * (Ly / ny) (Lz / nz); the wall rows carry none
  std::vector<double> wq(P, 0.0);
  for (int j = 0; j < n0; ++j) wq[j] = h->ops.fejer[j] * (h->cfg.ly / g.ny) * (h->cfg.lz / g.nz);
  if ((r = up(&X.wq, wq))) return r;
  auto twiddles = [&](double2 **dst, int n) -> int32_t {
    std::vector<double> t(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
      const long double a = 2.0L * 3.141592653589793238462643383279502884L * k / n;
      t[2 * k] = (double)cosl(a);
      t[2 * k + 1] = (double)sinl(a);
    }
    double *d = nullptr;
    if (int32_t q = up(&d, t)) return q;
    *dst = reinterpret_cast<double2 *>(d);
    return PCF_OK;
  };
  if ((r = twiddles(&X.twy, g.ny)) || (r = twiddles(&X.twz, g.nz))) return r;
  const size_t m = g.m, npl = (size_t)g.ny * g.nz;
  if ((r = h->alloc(&X.NS, (size_t)XD_COUNT * P * m)) || (r = h->alloc(&X.G, (size_t)10 * P * m)) ||
      (r = h->alloc(&X.phys, (size_t)10 * P * npl)) || (r = h->alloc(&X.E, (size_t)6 * n0 * npl)) ||
      (r = h->alloc(&X.EH, (size_t)6 * n0 * m)) || (r = h->alloc(&X.EC, (size_t)6 * n0 * m)) ||
      (r = h->alloc(&X.DIV, (size_t)2 * n0 * m)) || (r = h->alloc(&X.emf, (size_t)P * pcf::EMF_Q)))
    return r;
  // the Parseval jobs read only XDiag's own buffers, so they are set once
  const double *b0 = g.b0;
  auto ns = [&](int f) -> const double2 * { return X.NS + (size_t)f * P * m; };
  std::vector<pcf::ParsevalJob> jobs;
  auto job = [&](int kind, const double2 *a, const double2 *b, double aa = 0.0, double ab = 0.0) {
    pcf::ParsevalJob j;
    j.kind = kind;
    j.a = a;
    j.b = b ? b : a;
    j.add0a = aa;
    j.add0b = ab;
    jobs.push_back(j);
  };
  job(pcf::PV_DOT, ns(XD_U), ns(XD_U + 1));                              // 0 Reynolds
  job(pcf::PV_DOT, ns(XD_B), ns(XD_B + 1), b0[0], b0[1]);                // 1 Maxwell (total field)
  for (int i = 0; i < 3; ++i) job(pcf::PV_DOT, ns(XD_B + i), nullptr, b0[i], b0[i]);  // 2-4 |b + B0|^2
  for (int i = 0; i < 3; ++i) job(pcf::PV_MEAN, ns(XD_B + i), nullptr, b0[i]);        // 5-7 mean b + B0
  for (int i = 0; i < 3; ++i) job(pcf::PV_DOT, ns(XD_B + i), nullptr);                // 8-10 |b|^2
  for (int i = 0; i < 3; ++i) job(pcf::PV_DOT, ns(XD_DU + i), nullptr);               // 11-13 |du/dx|^2
  for (int i = 0; i < 3; ++i) job(pcf::PV_K2, ns(XD_U + i), nullptr);                 // 14-16 k^2 |u|^2
  for (int i = 0; i < 3; ++i) job(pcf::PV_DOT, ns(XD_DB + i), nullptr);               // 17-19 |db/dx|^2
  for (int i = 0; i < 3; ++i) job(pcf::PV_K2, ns(XD_B + i), nullptr);                 // 20-22 k^2 |b|^2
  for (int i = 0; i < 3; ++i) job(pcf::PV_KZ1, ns(XD_U + i), nullptr);                // 23-25
  for (int i = 0; i < 3; ++i) job(pcf::PV_KZ1, ns(XD_B + i), nullptr);                // 26-28
  job(pcf::PV_DOT, ns(XD_DIVA), nullptr);                                // 29 |div A|^2
  const double2 *di = X.DIV, *dr = X.DIV + (size_t)n0 * m;
  job(pcf::PV_DOT, di, nullptr);                                         // 30 |div E_ideal|^2
  job(pcf::PV_DOT, dr, nullptr);                                         // 31 |div E_resistive|^2
  job(pcf::PV_SUM2, di, dr);                                             // 32 |div E|^2
  X.njobs = (int)jobs.size();
  if ((r = h->alloc(&X.part, (size_t)X.njobs * n0))) return r;
  double *dj = nullptr;
  if ((r = h->alloc(&dj, jobs.size() * sizeof(pcf::ParsevalJob) / sizeof(double) + 1))) return r;
  PCF_HIP(hipMemcpy(dj, jobs.data(), jobs.size() * sizeof(pcf::ParsevalJob), hipMemcpyHostToDevice));
  X.jobs = reinterpret_cast<pcf::ParsevalJob *>(dj);
  X.ready = true;
  return PCF_OK;
}

int32_t pcf_mhd_extended_diagnostics(pcf_handle *h, double *out) {
  if (!h || !out) return fail(PCF_E_INVALID, "null argument");
  if (!h->mhd) return fail(PCF_E_INVALID, "not an MHD handle");
  if (h->cfg.lx[0] != -1.0 || h->cfg.lx[1] != 1.0) return fail(PCF_E_INVALID, "wall-normal domain must be (-1, 1)");
  PCF_HIP(hipSetDevice(h->device));
  if (int32_t r = xdiag_setup(h)) return r;
  PCF_HIP(ensure_u12(h));
  auto &X = h->xd;
  const pcf::Geo &g = h->geo;
  const int P = X.P, n0 = g.n0, nd = g.nd, nb = g.nb;
  const size_t m = g.m, fs = (size_t)P * m;
  const bool ins = h->insul;
  const int ka = ins ? n0 : nd;  // rows of A_y, A_z
  const double2 *A = h->A[h->so[0]], *A0 = A, *A1 = A + h->cstride, *A2 = A + 2 * h->cstride;
  const double2 *u[3] = {h->u0[h->so[0]], h->u1, h->u2};
  auto T = [](const double *M, const double2 *F, int K, int kf = 0, double c = 1.0) {
    pcf::EvalTerm t;
    t.M = M;
    t.F = F;
    t.K = K;
    t.kf = kf;
    t.c = c;
    return t;
  };
  auto out_of = [&](int f, int rows, std::initializer_list<pcf::EvalTerm> ts) {
    pcf::EvalOut o;
    o.dst = X.NS + (size_t)f * fs;
    o.rows = rows;
    o.nt = 0;
    for (const auto &t : ts) o.t[o.nt++] = t;
    return o;
  };
  // b = curl A in b_coeff_spaces (examples/pcf_mhd_jax.py:269-289; TC^3 for
  // insulating walls, examples/pcf_mhd_mri_shearpy_jax.py:548-557) and its
  // x-derivative at the P points, composed with the evaluation matrices.
  const double *EA0 = ins ? X.EC0 : X.ED0, *EA1 = ins ? X.EC1 : X.ED1, *EA2 = ins ? X.EC2 : X.ED2;
  std::vector<pcf::EvalOut> outs = {
      out_of(XD_U + 0, P, {T(X.EB0, u[0], nb)}),
      out_of(XD_U + 1, P, {T(X.ED0, u[1], nd)}),
      out_of(XD_U + 2, P, {T(X.ED0, u[2], nd)}),
      out_of(XD_DU + 0, P, {T(X.EB1, u[0], nb)}),
      out_of(XD_DU + 1, P, {T(X.ED1, u[1], nd)}),
      out_of(XD_DU + 2, P, {T(X.ED1, u[2], nd)}),
      // bx = i ky A2 - i kz A1 (in TD for conducting walls: the projection of a TD field is itself)
      out_of(XD_B + 0, P, {T(EA0, A2, ka, 1), T(EA0, A1, ka, 2, -1.0)}),
      out_of(XD_DB + 0, P, {T(EA1, A2, ka, 1), T(EA1, A1, ka, 2, -1.0)}),
      // by = i kz A0 - dA2/dx, bz = dA1/dx - i ky A0
      out_of(XD_B + 1, P, {T(X.ED0, A0, nd, 2), T(EA1, A2, ka, 0, -1.0)}),
      out_of(XD_DB + 1, P, {T(X.ED1, A0, nd, 2), T(EA2, A2, ka, 0, -1.0)}),
      out_of(XD_B + 2, P, {T(EA1, A1, ka), T(X.ED0, A0, nd, 1, -1.0)}),
      out_of(XD_DB + 2, P, {T(EA2, A1, ka), T(X.ED1, A0, nd, 1, -1.0)}),
      // div A (examples/pcf_mhd_mri_shearpy_jax.py:243-249), A_y and A_z
      out_of(XD_DIVA, P, {T(X.ED1, A0, nd), T(EA0, A1, ka, 1), T(EA0, A2, ka, 2)}),
      out_of(XD_A1, P, {T(EA0, A1, ka)}),
      out_of(XD_A2, P, {T(EA0, A2, ka)}),
  };
  hipError_t e = pcf::launch_eval_terms(g, outs.data(), (int)outs.size(), h->stream);
  if (e == hipSuccess) e = pcf::launch_curl_b(g, P, X.NS + XD_B * fs, X.NS + XD_DB * fs, X.NS + XD_J * fs, h->stream);
  // u, b, J, div A on the physical grid at all P rows
  if (e == hipSuccess) e = pcf::launch_dft_y(g, true, 10 * P, X.twy, X.NS, X.G, h->stream);
  if (e == hipSuccess) e = pcf::launch_dft_z_c2r(g, 10 * P, X.twz, X.G, X.phys, h->stream);
  if (e == hipSuccess) e = pcf::launch_emf(g, P, h->cfg.eta, X.xp, X.wq, X.phys, X.E, X.emf, h->stream);
  // TC projection of the ideal and resistive fields (mask_nyquist(TC.forward))
  // and their divergence at the nodes
  if (e == hipSuccess) e = pcf::launch_dft_z_r2c(g, 6 * n0, X.twz, X.E, X.EH, h->stream);
  if (e == hipSuccess) e = pcf::launch_dft_y(g, false, 6 * n0, X.twy, X.EH, X.EC, h->stream);
  pcf::EvalOut dv[2];
  for (int part = 0; part < 2; ++part) {
    const double2 *Ex = X.EC + (size_t)(3 * part) * n0 * m;
    pcf::EvalOut &o = dv[part];
    o.dst = X.DIV + (size_t)part * n0 * m;
    o.rows = n0;
    o.nt = 3;
    o.t[0] = T(X.Dn, Ex, n0);
    o.t[1] = T(X.I, Ex + (size_t)n0 * m, n0, 1);
    o.t[2] = T(X.I, Ex + (size_t)2 * n0 * m, n0, 2);
  }
  if (e == hipSuccess) e = pcf::launch_eval_terms(g, dv, 2, h->stream);
  if (e == hipSuccess) e = pcf::launch_parseval(g, X.jobs, X.njobs, X.part, h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("extended diagnostics launch: ") + hipGetErrorString(e));
  std::vector<double> part((size_t)X.njobs * n0), emf((size_t)P * pcf::EMF_Q);
  double2 wall[4];  // mode 0 of A_y, A_z at x = -1, +1
  PCF_HIP(hipMemcpyAsync(part.data(), X.part, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  PCF_HIP(hipMemcpyAsync(emf.data(), X.emf, emf.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  for (int k = 0; k < 2; ++k)
    for (int w = 0; w < 2; ++w)
      PCF_HIP(hipMemcpyAsync(&wall[2 * k + w], X.NS + (size_t)(XD_A1 + k) * fs + (size_t)(n0 + w) * m,
                             sizeof(double2), hipMemcpyDeviceToHost, h->stream));
  PCF_HIP(hipStreamSynchronize(h->stream));
  // assembly (examples/pcf_mhd_mri_shearpy_jax.py:178-215, 251-263, 355-429, 431-513)
  const double LL = h->cfg.ly * h->cfg.lz, V = 2.0 * LL, lx = 2.0, *b0 = g.b0;
  const std::vector<double> &fj = h->ops.fejer;
  double fsum = 0.0;
  for (int x = 0; x < n0; ++x) fsum += fj[x];
  auto I = [&](int j) {  // Ly Lz sum_x fejer_x part[j][x]: the volume integral
    double s = 0.0;
    for (int x = 0; x < n0; ++x) s += fj[x] * part[(size_t)j * n0 + x];
    return LL * s;
  };
  auto kz1 = [&](int j0) {
    double s = 0.0;
    for (int j = j0; j < j0 + 3; ++j)
      for (int x = 0; x < n0; ++x) s += (fj[x] / fsum) * part[(size_t)j * n0 + x];
    return g.nz >= 2 ? std::sqrt(2.0 * s) : 0.0;
  };
  auto sumrows = [&](int q) {
    double s = 0.0;
    for (int p = 0; p < n0; ++p) s += emf[(size_t)p * pcf::EMF_Q + q];
    return s;
  };
  auto maxrows = [&](int q, int p0, int p1) {
    double s = 0.0;
    for (int p = p0; p < p1; ++p) s = std::max(s, emf[(size_t)p * pcf::EMF_Q + q]);
    return s;
  };
  const double npl = (double)g.ny * g.nz;
  auto wmean = [&](int q, int wrow) { return emf[(size_t)(n0 + wrow) * pcf::EMF_Q + q] / npl; };
  out[0] = I(0) / V;                                   // reynolds_stress
  out[1] = -I(1) / V;                                  // maxwell_stress
  out[2] = I(2) + I(3) + I(4);                         // Emag_total
  for (int i = 0; i < 3; ++i) out[3 + i] = I(5 + i) / V;  // mean_b
  out[6] = I(8) + I(9) + I(10);                        // mag_energy_fluct_total
  double dk = 0.0, dm = 0.0;
  for (int j = 11; j < 17; ++j) dk += I(j);
  for (int j = 17; j < 23; ++j) dm += I(j);
  out[7] = h->cfg.nu * dk;                             // dissipation_kinetic
  out[8] = h->cfg.eta * dm;                            // dissipation_magnetic
  out[9] = kz1(23);                                    // channel_kz1_velocity_rms
  out[10] = kz1(26);                                   // channel_kz1_magnetic_rms
  out[11] = maxrows(pcf::EMF_BMAX, 0, n0);             // bmax
  out[12] = maxrows(pcf::EMF_BTMAX, 0, n0);            // bmax_total
  out[13] = std::sqrt(sumrows(pcf::EMF_IDEAL2));       // electric_ideal_l2
  out[14] = std::sqrt(sumrows(pcf::EMF_RES2));         // electric_resistive_l2
  out[15] = std::sqrt(sumrows(pcf::EMF_TOT2));         // electric_total_l2
  out[16] = std::sqrt(I(32));                          // divergence_e_l2
  out[17] = std::sqrt(I(30));                          // divergence_e_ideal_l2
  out[18] = std::sqrt(I(31));                          // divergence_e_resistive_l2
  out[19] = std::sqrt(I(29));                          // divergence_a_l2
  out[20] = maxrows(pcf::EMF_WDIVA, n0, P);            // divergence_a_wall_linf
  out[21] = maxrows(pcf::EMF_WIDEAL, n0, P);           // electric_ideal_wall_tangential_linf
  out[22] = maxrows(pcf::EMF_WRES, n0, P);             // electric_resistive_wall_tangential_linf
  out[23] = maxrows(pcf::EMF_WTOT, n0, P);             // electric_wall_tangential_linf
  out[24] = (wmean(pcf::EMF_WSUM2, 1) - wmean(pcf::EMF_WSUM2, 0)) / lx;    // faraday_mean_by_tendency
  out[25] = -(wmean(pcf::EMF_WSUM1, 1) - wmean(pcf::EMF_WSUM1, 0)) / lx;   // faraday_mean_bz_tendency
  out[26] = b0[0];                                     // mean_b*_trace (magnetic_flux_from_a)
  out[27] = -(wall[3].x - wall[2].x) / lx + b0[1];
  out[28] = (wall[1].x - wall[0].x) / lx + b0[2];
  for (int q = 0; q < 9; ++q) out[29 + q] = sumrows(pcf::EMF_MEAN + q) / V;  // mean_e{x,y,z}_{ideal,resistive,total}
  return PCF_OK;
}

int32_t pcf_physical_fields(pcf_handle *h, double *out) {
  if (!h || !out) return fail(PCF_E_INVALID, "null argument");
  if (!h->maxima_ok)
    return fail(PCF_E_UNSUPPORTED, "physical fields need ny, nz that factor into 2, 3, 4 within the FFT kernels");
  const pcf::Geo &g0 = h->geo0;
  const int nf = h->mhd ? 6 : 3;
  const size_t count = (size_t)nf * g0.n0 * g0.my * g0.mz;
  // the output region: the workspace's forward-transform buffers (free between steps)
  const size_t ve = (h->cfg.precision == PCF_PREC_FP64) ? sizeof(double2) : sizeof(float2);
  const size_t fout = h->mhd ? 6 : 3;
  const size_t room = fout * (size_t)h->geo.n0 * ((size_t)h->geo.my * h->geo.kzs + h->geo.m) * ve;
  if (count * sizeof(double) > room) return fail(PCF_E_UNSUPPORTED, "physical fields exceed the workspace");
  PCF_HIP(hipSetDevice(h->device));
  double *phys = reinterpret_cast<double *>(h->ws.Hz);
  const bool fp64 = h->cfg.precision == PCF_PREC_FP64;
  hipError_t e = fp64 ? field_maxima_stages<double>(h, h->md, h->dblockmax, phys)
                      : field_maxima_stages<float>(h, h->mf, h->dblockmax, phys);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("physical fields launch: ") + hipGetErrorString(e));
  PCF_HIP(hipMemcpyAsync(out, phys, count * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  PCF_HIP(hipStreamSynchronize(h->stream));
  return PCF_OK;
}

int32_t pcf_stage_count(void) { return NSTAGE; }

const char *pcf_stage_name(int32_t i) { return (i >= 0 && i < NSTAGE) ? STAGE_NAMES[i] : ""; }

int32_t pcf_profile_steady(pcf_handle *h, int32_t nsteps, double *stage_ms) {
  if (!h || !stage_ms) return fail(PCF_E_INVALID, "null argument");
  if (h->cfg.time_integrator != PCF_TI_SBDF3) return fail(PCF_E_UNSUPPORTED, "steady steps are SBDF3 steps");
  if (h->history_steps < 2.0) return fail(PCF_E_INVALID, "history not bootstrapped (history_steps < 2)");
  if (nsteps <= 0) return fail(PCF_E_INVALID, "nsteps must be positive");
  PCF_HIP(hipSetDevice(h->device));
  if (int32_t r0 = resolve_given(h)) return r0;
  StageTimer t{};
  for (int i = 0; i < NSTAGE; ++i)
    for (int j = 0; j < 2; ++j) PCF_HIP(hipEventCreate(&t.ev[i][j]));
  double acc[NSTAGE] = {0, 0, 0, 0, 0, 0};
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

// Timing of the reverse pass's steady steps outside pcf_vjp: nsteps eager
// steady steps keep their tape (handle stream, no graph), then, after a host
// synchronisation, the energy seed and nsteps reverse steady steps run on the
// reverse stream alone, each bracketed by HIP events on that stream.  No
// recompute overlaps the reverse, nothing is captured.  The state is put back
// afterwards.  out[0] = ms per reverse steady step, out[1] = ms per forward
// steady step with its tape write.
int32_t pcf_profile_reverse(pcf_handle *h, int32_t nsteps, double *out) {
  if (!h || !out) return fail(PCF_E_INVALID, "null argument");
  if (h->mhd) return fail(PCF_E_UNSUPPORTED, "the reverse mode covers the hydrodynamic step only");
  if (h->cfg.time_integrator != PCF_TI_SBDF3) return fail(PCF_E_UNSUPPORTED, "the reverse mode covers SBDF3 only");
  if (h->history_steps < 2.0) return fail(PCF_E_INVALID, "history not bootstrapped (history_steps < 2)");
  if (nsteps < 1 || nsteps > 256) return fail(PCF_E_INVALID, "nsteps must be in [1, 256]");
  PCF_HIP(hipSetDevice(h->device));
  if (int32_t r0 = resolve_given(h)) return r0;
  int32_t r;
  if ((r = adj_setup(h))) return r;
  if ((r = adj_reserve(h, nsteps, nsteps))) return r;
  auto &A = h->adj;
  const pcf::Geo &g = h->geo;
  hipError_t e;
  if ((e = snap_save(h, A.snaps[0]))) return fail(PCF_E_HIP, std::string("snapshot: ") + hipGetErrorString(e));
  hipEvent_t ev[4] = {};
  for (auto &x : ev) PCF_HIP(hipEventCreate(&x));
  auto cleanup = [&]() {
    for (auto &x : ev) (void)hipEventDestroy(x);
  };
  PCF_HIP(hipEventRecord(ev[0], h->stream));
  for (int i = 0; i < nsteps; ++i)
    if ((r = steady_step(h, &A.tape[i]))) {
      cleanup();
      return r;
    }
  PCF_HIP(hipEventRecord(ev[1], h->stream));
  if ((e = ensure_u12(h)) == hipSuccess) e = hipStreamSynchronize(h->stream);
  const double scale = 2.0 * h->cfg.ly * h->cfg.lz;
  if (e == hipSuccess)
    e = pcf::launch_energy_seed(g, A.QB, A.QD, scale, h->u0[h->so[0]], h->u1, h->u2, A.mu_u[0], A.mu_u[1],
                                A.mu_u[2], A.stream);
  if (e == hipSuccess)
    e = pcf::launch_recon_adj(g, A.RcT, A.ldrT, A.mu_u[0], A.mu_u[1], A.mu_u[2], nullptr, A.muX, A.stream);
  for (int i = 0; i < 3 && e == hipSuccess; ++i) {
    e = zero_prim(h, A.rho[i]);
    A.rr[i] = i;
  }
  if (e == hipSuccess) e = hipEventRecord(ev[2], A.stream);
  // steps 1..nsteps-1 of the run (index k > 0: the step-0 branch reads the
  // initial state's leaves)
  for (int k = nsteps - 1; k >= 1 && e == hipSuccess; --k) e = reverse_step(h, k, true, &A.tape[k]);
  if (e == hipSuccess) e = hipEventRecord(ev[3], A.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(A.stream);
  float fwd = 0.0f, rev = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&fwd, ev[0], ev[1]);
  if (e == hipSuccess) e = hipEventElapsedTime(&rev, ev[2], ev[3]);
  cleanup();
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("reverse profile: ") + hipGetErrorString(e));
  if ((e = snap_load(h, A.snaps[0])) == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("restore: ") + hipGetErrorString(e));
  A.have_grad = false;
  out[0] = nsteps > 1 ? (double)rev / (nsteps - 1) : 0.0;
  out[1] = (double)fwd / nsteps;
  return PCF_OK;
}

int32_t pcf_profile_stage(pcf_handle *h, int32_t stage, int32_t reps, double *ms_per_launch) {
  if (!h || !ms_per_launch) return fail(PCF_E_INVALID, "null argument");
  if (stage < 0 || stage > 4) return fail(PCF_E_INVALID, "stage must be a nonlinear stage (0..4)");
  if (reps <= 0) return fail(PCF_E_INVALID, "reps must be positive");
  PCF_HIP(hipSetDevice(h->device));
  if (int32_t r0 = resolve_given(h)) return r0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PCF_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(PCF_E_HIP, "event create failed");
  }
  hipError_t e = (h->cfg.precision == PCF_PREC_FP64) ? stage_repeat<double>(h, h->md, stage, reps, e0, e1)
                                                     : stage_repeat<float>(h, h->mf, stage, reps, e0, e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(PCF_E_HIP, std::string("stage profile: ") + hipGetErrorString(e));
  *ms_per_launch = ms / reps;
  return PCF_OK;
}

int32_t pcf_set_cotangent(pcf_handle *h, int32_t field, const double *host) {
  if (!h || !host) return fail(PCF_E_INVALID, "null argument");
  int idx;
  switch (field) {
    case PCF_F_U0: idx = 0; break;
    case PCF_F_U1: idx = 1; break;
    case PCF_F_U2: idx = 2; break;
    case PCF_F_G: idx = 3; break;
    case PCF_F_AX:
    case PCF_F_AY:
    case PCF_F_AZ: idx = 4 + (field - PCF_F_AX); break;
    default: return fail(PCF_E_INVALID, "cotangents are taken for U0, U1, U2, G (and AX, AY, AZ under MHD) only");
  }
  if (idx >= 4 && !h->mhd) return fail(PCF_E_INVALID, "no vector potential: not an MHD handle");
  PCF_HIP(hipSetDevice(h->device));
  int32_t r = adj_setup(h);
  if (r) return r;
  if (idx >= 4) {
    if (h->insul || h->cfg.precision != PCF_PREC_FP64)
      return fail(PCF_E_UNSUPPORTED, "the MHD reverse mode covers conducting walls in FP64");
    if ((r = adj_setup_mhd(h))) return r;
    std::vector<double2> half;
    half_from_cotangent(h->geo, h->geo.nd, host, half);
    PCF_HIP(hipStreamSynchronize(h->stream));
    PCF_HIP(hipMemcpy(h->adj.cotA + (idx - 4) * h->cstride, half.data(), half.size() * sizeof(double2),
                      hipMemcpyHostToDevice));
    h->adj.cotA_set = true;
    return PCF_OK;
  }
  const pcf::Geo &g = h->geo;
  const int rows = (idx == 0) ? g.nb : g.nd;
  std::vector<double2> half;
  half_from_cotangent(g, rows, host, half);
  double2 *dst = (idx < 3) ? h->adj.cot_u[idx] : h->adj.cot_g;
  PCF_HIP(hipStreamSynchronize(h->stream));
  PCF_HIP(hipMemcpy(dst, half.data(), half.size() * sizeof(double2), hipMemcpyHostToDevice));
  h->adj.cot_set[idx] = true;
  return PCF_OK;
}

int32_t pcf_vjp(pcf_handle *h, int32_t steps, int32_t checkpoint_every, int32_t seed) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  if (h->mhd && h->insul)
    return fail(PCF_E_UNSUPPORTED, "the MHD reverse mode covers conducting walls (A in TD^3) only");
  if (h->mhd && h->cfg.precision != PCF_PREC_FP64)
    return fail(PCF_E_UNSUPPORTED, "the MHD reverse mode runs in FP64");
  if (h->cfg.time_integrator != PCF_TI_SBDF3) return fail(PCF_E_UNSUPPORTED, "the reverse mode covers SBDF3 only");
  if (steps < 1) return fail(PCF_E_INVALID, "steps must be >= 1");
  if (checkpoint_every < 1) return fail(PCF_E_INVALID, "checkpoint_every must be >= 1");
  if (seed != PCF_SEED_ENERGY && seed != PCF_SEED_COTANGENT) return fail(PCF_E_INVALID, "unknown seed");
  if (seed == PCF_SEED_COTANGENT &&
      !(h->adj.cot_set[0] || h->adj.cot_set[1] || h->adj.cot_set[2] || h->adj.cot_set[3] || h->adj.cotA_set))
    return fail(PCF_E_INVALID, "no cotangent set (pcf_set_cotangent)");
  PCF_HIP(hipSetDevice(h->device));
  int32_t r;
  if ((r = adj_setup(h))) return r;
  if (h->mhd && (r = adj_setup_mhd(h))) return r;
  if ((r = adj_reserve(h, steps, checkpoint_every))) return r;
  if (h->mhd) return vjp_mhd(h, steps, checkpoint_every, seed);
  auto &A = h->adj;
  const pcf::Geo &g = h->geo;
  const int every = checkpoint_every;
  const int nseg = (steps + every - 1) / every;
  const int h0 = (int)std::floor(std::min(2.0, std::max(0.0, h->history_steps)));
  const double dt = h->cfg.dt;
  A.have_grad = false;
  auto hip_fail = [&](hipError_t e, const char *what) {
    return fail(PCF_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e;
  // Runs segment seg's steps on the handle stream from the current state,
  // keeping the tape in set seg & 1 when keep is set; first[i] receives the
  // tape slot of step ka + i.
  auto sweep = [&](int seg, bool keep, std::vector<int> &first) -> int32_t {
    const int ka = seg * every, kb = std::min(steps, ka + every);
    const int base = (seg & 1) * A.per_set;
    first.assign(kb - ka, 0);
    int slot = base;
    for (int k = ka; k < kb;) {
      if (h->history_steps >= 2.0) {  // steady to the end of the segment
        for (int i = 0; i < kb - k; ++i) first[k - ka + i] = slot + i;
        return steady_run(h, kb - k, keep ? slot : -1);
      }
      first[k - ka] = slot;
      int32_t r = startup_step(h, keep ? &A.tape[slot] : nullptr);
      if (r) return r;
      slot += 3;
      ++k;
    }
    return PCF_OK;
  };
  // forward, keeping a snapshot at the start of every segment; the last
  // segment keeps its tape, so it is not recomputed
  std::vector<int> first_rev, first_next;
  for (int seg = 0; seg < nseg; ++seg) {
    if ((e = snap_save(h, A.snaps[seg]))) return hip_fail(e, "snapshot");
    if ((r = sweep(seg, seg == nseg - 1, first_rev))) return r;
  }
  pcf_handle::Snap &fin = A.snaps[nseg];
  if ((e = snap_save(h, fin))) return hip_fail(e, "snapshot");
  if ((e = ensure_u12(h))) return hip_fail(e, "reconstruction");  // the energy seed reads u1, u2
  // seed (on the reverse stream, after the forward): mu of the final state's
  // (u, g) -> mu_{x_K}
  PCF_HIP(hipEventRecord(A.ev_seed, h->stream));
  PCF_HIP(hipStreamWaitEvent(A.stream, A.ev_seed, 0));
  if (seed == PCF_SEED_ENERGY) {
    const double scale = 2.0 * h->cfg.ly * h->cfg.lz;
    if ((e = pcf::launch_energy_seed(g, A.QB, A.QD, scale, h->u0[h->so[0]], h->u1, h->u2, A.mu_u[0],
                                     A.mu_u[1], A.mu_u[2], A.stream)))
      return hip_fail(e, "energy seed");
    e = pcf::launch_recon_adj(g, A.RcT, A.ldrT, A.mu_u[0], A.mu_u[1], A.mu_u[2], nullptr, A.muX, A.stream);
  } else {
    e = pcf::launch_recon_adj(g, A.RcT, A.ldrT, A.cot_u[0], A.cot_u[1], A.cot_u[2], A.cot_g, A.muX, A.stream);
  }
  if (e) return hip_fail(e, "seed");
  for (int i = 0; i < 3; ++i) {
    if ((e = zero_prim(h, A.rho[i]))) return hip_fail(e, "memset");
    A.rr[i] = i;
  }
  // the recompute overwrites the state the seed reads
  PCF_HIP(hipEventRecord(A.ev_seed, A.stream));
  PCF_HIP(hipStreamWaitEvent(h->stream, A.ev_seed, 0));
  // reverse, segment by segment from the last; the recompute of segment
  // seg - 1 (handle stream) overlaps the reverse of segment seg (reverse
  // stream).  Tape set b is rewritten only after the reverse that read it
  // (ev_rev[b]) and read only after the recompute that wrote it (ev_rc[b]).
  bool rev_pending[2] = {false, false};
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int ka = seg * every, kb = std::min(steps, ka + every);
    if (seg < nseg - 1) {
      first_rev.swap(first_next);
      PCF_HIP(hipStreamWaitEvent(A.stream, A.ev_rc[seg & 1], 0));
    }
    if (seg > 0) {
      const int b = (seg - 1) & 1;
      if (rev_pending[b]) PCF_HIP(hipStreamWaitEvent(h->stream, A.ev_rev[b], 0));
      if ((e = snap_load(h, A.snaps[seg - 1]))) return hip_fail(e, "restore");
      if ((r = sweep(seg - 1, true, first_next))) return r;
      PCF_HIP(hipEventRecord(A.ev_rc[b], h->stream));
    }
    for (int k = kb - 1; k >= ka; --k)
      if ((e = reverse_step(h, k, h0 + k >= 2, &A.tape[first_rev[k - ka]]))) return hip_fail(e, "reverse step");
    PCF_HIP(hipEventRecord(A.ev_rev[seg & 1], A.stream));
    rev_pending[seg & 1] = true;
  }
  // history leaves of the initial state: levels -1 and -2
  {
    const double e1 = -dt * 18.0 / 11.0, e2 = dt * 6.0 / 11.0;
    const pcf::Prim &r0 = A.rho[A.rr[0]], &r1 = A.rho[A.rr[1]];
    pcf::PrimSum s{};
    for (int i = 0; i < 2; ++i)
      if ((e = zero_prim(h, A.gh[i]))) return hip_fail(e, "memset");
    ps_add(s, r0, C_SBDF[1]);
    ps_add(s, r1, C_SBDF[2]);
    if ((e = pcf::launch_adj_mass(g, h->sm, 0.0, pcf::PrimC{}, s, A.gh[0], A.stream))) return hip_fail(e, "history");
    s = pcf::PrimSum{};
    ps_add(s, r0, C_SBDF[2]);
    if ((e = pcf::launch_adj_mass(g, h->sm, 0.0, pcf::PrimC{}, s, A.gh[1], A.stream))) return hip_fail(e, "history");
    s = pcf::PrimSum{};
    ps_add(s, r0, e1);
    ps_add(s, r1, e2);
    if ((e = pcf::launch_prim_sum(g, s, A.gh[2], A.stream))) return hip_fail(e, "history");
    s = pcf::PrimSum{};
    ps_add(s, r0, e2);
    if ((e = pcf::launch_prim_sum(g, s, A.gh[3], A.stream))) return hip_fail(e, "history");
  }
  PCF_HIP(hipStreamSynchronize(A.stream));
  if ((e = snap_load(h, fin))) return hip_fail(e, "restore");
  PCF_HIP(hipStreamSynchronize(h->stream));
  A.have_grad = true;
  return PCF_OK;
}

int32_t pcf_get_gradient(pcf_handle *h, int32_t field, double *host) {
  if (field < 0 || field >= PCF_F_COUNT) return fail(PCF_E_INVALID, "no such field");
  if (!h || !host) return fail(PCF_E_INVALID, "null argument");
  // the vector-potential leaves have a gradient under MHD only
  if (field >= PCF_F_AX && !h->mhd) return fail(PCF_E_INVALID, "no gradient for this field (flow leaves only)");
  auto &A = h->adj;
  if (!A.have_grad) return fail(PCF_E_INVALID, "no gradient: call pcf_vjp first");
  const pcf::Geo &g = h->geo;
  PCF_HIP(hipSetDevice(h->device));
  PCF_HIP(hipStreamSynchronize(h->stream));
  const size_t M = g.m;
  auto fetch = [&](const void *src, size_t count, std::vector<double2> &dst) -> hipError_t {
    dst.resize(count);
    return hipMemcpy(dst.data(), src, count * sizeof(double2), hipMemcpyDeviceToHost);
  };
  std::vector<double2> half, extra;
  int rows = g.nd;
  if (field >= PCF_F_AX) {
    // A, then the histories (NL_A_OLD, NL_A_OLDER, A_OLD, A_OLDER) of the initial state
    const int grp = (field - PCF_F_AX) / 3, comp = (field - PCF_F_AX) % 3;
    const double2 *src = (grp == 0) ? A.muA : (grp == 1) ? A.ghA[2] : (grp == 2) ? A.ghA[3] : (grp == 3) ? A.ghA[0]
                                                                                                       : A.ghA[1];
    PCF_HIP(fetch(src + comp * h->cstride, (size_t)rows * M, half));
    full_from_half_grad(g, rows, half, host);
    return PCF_OK;
  }
  if (field == PCF_F_U0 || field == PCF_F_U1 || field == PCF_F_U2 || field == PCF_F_G) {
    if (field == PCF_F_U0) {
      rows = g.nb;
      PCF_HIP(fetch(A.mu_u[0], rows * M, half));
      PCF_HIP(fetch(A.muX.u0, rows * M, extra));
      for (size_t i = 0; i < half.size(); ++i) {
        half[i].x += extra[i].x;
        half[i].y += extra[i].y;
      }
    } else if (field == PCF_F_G) {
      PCF_HIP(fetch(A.muX.g, rows * M, half));
    } else {
      const int c = (field == PCF_F_U1) ? 1 : 2;
      PCF_HIP(fetch(A.mu_u[c], rows * M, half));
      std::vector<double> vec(rows);
      PCF_HIP(hipMemcpy(vec.data(), c == 1 ? A.muX.v : A.muX.w, rows * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < rows; ++i) half[(size_t)i * M].x += vec[i];
    }
    full_from_half_grad(g, rows, half, host);
    return PCF_OK;
  }
  int grp, comp;
  if (field >= PCF_F_SOL_OLDER_U) grp = 1, comp = field - PCF_F_SOL_OLDER_U;
  else if (field >= PCF_F_SOL_OLD_U) grp = 0, comp = field - PCF_F_SOL_OLD_U;
  else if (field >= PCF_F_NL_OLDER_U) grp = 3, comp = field - PCF_F_NL_OLDER_U;
  else grp = 2, comp = field - PCF_F_NL_OLD_U;
  const pcf::Prim &P = A.gh[grp];
  if (comp >= 2) {
    PCF_HIP(hipMemcpy(host, comp == 2 ? P.v : P.w, g.nd * sizeof(double), hipMemcpyDeviceToHost));
    return PCF_OK;
  }
  rows = (comp == 0) ? g.nb : g.nd;
  PCF_HIP(fetch(comp == 0 ? (const void *)P.u0 : (const void *)P.g, rows * M, half));
  full_from_half_grad(g, rows, half, host);
  return PCF_OK;
}

int32_t pcf_kernels_per_step(pcf_handle *h) {
  if (!h) return fail(PCF_E_INVALID, "null handle");
  // MHD adds the b/j transform, the induction rows and the A solve
  return h->mhd ? NSTAGE + 3 : NSTAGE;
}

}  // extern "C"
