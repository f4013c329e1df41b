// C ABI of libnsp: handle lifetime, state transfer, the time-step loop,
// restart files, initial conditions and diagnostics.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <random>

#include "nsp_internal.h"

using namespace nsp;

static thread_local std::string g_err;

static nsp_status fail(nsp_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(NSP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

struct nsp_handle {
  nsp_params P;
  Geom g;
  int device;
  hipStream_t stream = nullptr;
  RadialTables host;
  DevTables t{};
  DevInfluence b{};
  NlBufs nb{};
  // fields: groups of three components, N complex each
  double2* u = nullptr;     // current state (r/theta)
  double2* ui = nullptr;    // iterate
  double2* un = nullptr;    // next iterate
  double2* Nl = nullptr;    // nonlinear term
  double2* E = nullptr;     // explicit part of the step (+/-), sweep-side layout
  double2* SR = nullptr;    // sweep scratch (3 fields), sweep-side layout
  double2* SP = nullptr;    // sweep scratch (1 field), sweep-side layout
  double2* DH = nullptr;    // refined form: div(H u^n) of the step, sweep-side layout
  double* F = nullptr;      // banded LU factors (u+ all modes, u- row NT/2, uz and p rows <= NT/2)
  float* F32 = nullptr;     // mixed form: FP32 copies of the u+, u-, uz sets (read for |m| >= kM32)
  double* PH = nullptr;     // homogeneous pressure data (rows it <= NT/2), sweep-side layout
  double* Dc = nullptr;     // centred D table
  double* Luc = nullptr;    // centred Lu table
  unsigned long long* norm = nullptr;  // device [2]
  unsigned long long* norm_h = nullptr;  // pinned host [2]
  double* part = nullptr;   // diagnostics partials [MODES][4]
  // whole-step graphs: state buffers U[0..2] (u, ui, un point into them),
  // U[loc] holds the state; one graph per starting buffer
  double2* U[3] = {};
  int loc = 0;
  bool graphs = true;
  StepRec* rec = nullptr;    // device
  StepRec* rec_h = nullptr;  // pinned host
  // captured step from state slot a: predictor and the first corrector
  // (gexec), the remaining correctors (gexec2, launched only when the first
  // corrector has not converged)
  hipGraphExec_t gexec[3] = {};
  hipGraphExec_t gexec2[3] = {};
  std::vector<void*> allocs;
  // diagnostics files: friction, qcross, ucross, meanpr
  FILE* io[4] = {};
  Profiler prof;
  std::string tune_report = "{\"tuned\": false}";
};

namespace {

template <typename T>
nsp_status dalloc(nsp_handle* h, T** ptr, size_t count) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) return fail(NSP_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  h->allocs.push_back(p);
  *ptr = reinterpret_cast<T*>(p);
  return NSP_OK;
}

void drop_graphs(nsp_handle* h) {
  for (int a = 0; a < 3; a++) {
    if (h->gexec[a]) (void)hipGraphExecDestroy(h->gexec[a]);
    if (h->gexec2[a]) (void)hipGraphExecDestroy(h->gexec2[a]);
    h->gexec[a] = h->gexec2[a] = nullptr;
  }
}

void close_io(nsp_handle* h) {
  for (FILE*& f : h->io) {
    if (f) std::fclose(f);
    f = nullptr;
  }
}

void free_all(nsp_handle* h) {
  close_io(h);
  drop_graphs(h);
  if (h->rec_h) (void)hipHostFree(h->rec_h);
  h->rec_h = nullptr;
  for (void* p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  if (h->norm_h) (void)hipHostFree(h->norm_h);
  h->norm_h = nullptr;
  if (h->stream) (void)hipStreamDestroy(h->stream);
  h->stream = nullptr;
  if (h->nb.s2) (void)hipStreamDestroy(h->nb.s2);
  h->nb.s2 = nullptr;
  for (hipEvent_t& e : h->nb.ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
}

nsp_status check_params(const nsp_params* p) {
  if (!p) return fail(NSP_ERR_INVALID, "null params");
  if (p->nr < 2 * IW + 2) return fail(NSP_ERR_INVALID, "nr must be >= 8");
  if (p->nt < 2 || p->nt % 2) return fail(NSP_ERR_INVALID, "nt must be even and >= 2");
  if (p->nz < 3 || p->nz % 2 == 0) return fail(NSP_ERR_INVALID, "nz must be odd and >= 3");
  if (!(p->lz > 0) || !(p->re > 0) || !(p->dt > 0)) return fail(NSP_ERR_INVALID, "lz, re, dt must be > 0");
  if (p->maxit < 1) return fail(NSP_ERR_INVALID, "maxit must be >= 1");
  if (p->precision != NSP_FP64 && p->precision != NSP_MIXED && p->precision != NSP_MIXED_REFINED)
    return fail(NSP_ERR_INVALID, "bad precision");
  const int nzf = 3 * (p->nz - 1);
  if (nzf > 36 * 256) return fail(NSP_ERR_INVALID, "3(nz-1) must be <= 9216");
  return NSP_OK;
}

std::vector<double2> twiddles(int L, int count) {
  std::vector<double2> w(count);
  for (int e = 0; e < count; e++) {
    const double a = -PI2 * double(e) / double(L);
    w[e] = make_double2(std::cos(a), std::sin(a));
  }
  return w;
}

template <typename T>
nsp_status upload(nsp_handle* h, T** dst, const std::vector<T>& src) {
  nsp_status st = dalloc(h, dst, src.size());
  if (st) return st;
  HIPCHK(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return NSP_OK;
}

nsp_status copy_fields(nsp_handle* h, double2* dst, const double* a, const double* b, const double* c, int where) {
  const size_t bytes = h->g.n * sizeof(double2);
  const double* src[3] = {a, b, c};
  const hipMemcpyKind kind = where == NSP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  for (int q = 0; q < 3; q++) {
    if (!src[q]) return fail(NSP_ERR_INVALID, "null field pointer");
    HIPCHK(hipMemcpyAsync(dst + q * h->g.n, src[q], bytes, kind, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return NSP_OK;
}

nsp_status fetch_fields(nsp_handle* h, const double2* src, double* a, double* b, double* c, int where) {
  const size_t bytes = h->g.n * sizeof(double2);
  double* dst[3] = {a, b, c};
  const hipMemcpyKind kind = where == NSP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  for (int q = 0; q < 3; q++) {
    if (!dst[q]) return fail(NSP_ERR_INVALID, "null field pointer");
    HIPCHK(hipMemcpyAsync(dst[q], src + q * h->g.n, bytes, kind, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return NSP_OK;
}

// The mixed form's sweep reads the u+, u-, uz LU factors rounded to FP32 for
// |m| >= kM32; the low-|m| sets (near-axis stiffness), the pressure factors
// and all arithmetic stay FP64.
constexpr int kM32 = 4;

// un: the step's state u^n (the refined form's solves are for u - u^n)
Lin2Args lin_args(nsp_handle* h, const Geom& g, const double2* N, const double2* u, double2* out, bool predictor,
                  const double2* un) {
  Lin2Args a;
  a.g = g;
  a.r = h->t.r;
  a.rdr = h->t.rdr;
  a.uc = h->t.uc;
  a.qc = h->t.qc;
  a.Dc = h->Dc;
  a.Luc = h->Luc;
  a.Lu = h->t.Lu;
  a.Lp = h->t.Lp;
  a.F = h->F;
  a.F32 = h->F32;
  a.m32 = kM32;
  a.s32 = h->nb.mixed ? 1 : 0;
  a.form64 = h->nb.mixed ? 0 : 1;
  a.inc = h->P.precision == NSP_MIXED_REFINED ? 1 : 0;
  a.DH = h->DH;
  a.un = un;
  a.PH = h->PH;
  a.invc = h->b.invc;
  a.N = N;
  a.n32 = h->nb.mixed ? 1 : 0;
  a.E = h->E;
  a.u = u;
  a.out = out;
  a.SR = h->SR;
  a.SP = h->SP;
  a.norm = h->norm;
  a.predictor = predictor ? 1 : 0;
  a.ord = h->nb.var.s_ord;
  return a;
}

std::vector<float2> to_float2(const std::vector<double2>& v) {
  std::vector<float2> o(v.size());
  for (size_t q = 0; q < v.size(); q++) o[q] = make_float2(float(v[q].x), float(v[q].y));
  return o;
}

double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof d);
  return d;
}

}  // namespace

namespace {
// The "chosen" object of the tuning report.
std::string variant_json(const NlVariant& v) {
  auto tvj = [](const ThetaVar& t) {
    char o[112];
    std::snprintf(o, sizeof o, "{\"alg\": %d, \"tl\": %d, \"nthr\": %d, \"pad\": %d, \"ord\": %d}", t.alg, t.tl,
                  t.nthr, t.pad, t.ord);
    return std::string(o);
  };
  char z[176];
  std::snprintf(z, sizeof z,
                ", \"z_alg\": %d, \"z_nthr\": %d, \"z_pad\": %d, \"z_pf\": %d, \"s_ord\": %d, \"nl_chunks\": %d}",
                v.z_alg, v.z_nthr, v.z_pad, v.z_pf, v.s_ord, v.nl_chunks);
  return "{\"theta_inv\": " + tvj(v.inv) + ", \"theta_fwd\": " + tvj(v.fwd) + z;
}

// Integer value of "key": n in text, searched from `from` (-1 if absent).
long json_int(const std::string& t, size_t from, const char* key, size_t* at = nullptr) {
  const std::string k = std::string("\"") + key + "\"";
  const size_t p = t.find(k, from);
  if (p == std::string::npos) return -1;
  size_t c = t.find(':', p + k.size());
  if (c == std::string::npos) return -1;
  c++;
  while (c < t.size() && (t[c] == ' ' || t[c] == '\t')) c++;
  if (c < t.size() && t[c] == '{') {  // object: caller reads its fields
    if (at) *at = c;
    return 0;
  }
  char* end = nullptr;
  const long v = std::strtol(t.c_str() + c, &end, 10);
  if (end == t.c_str() + c) return -1;
  return v;
}

}  // namespace

extern "C" {

const char* nsp_last_error(void) { return g_err.c_str(); }

nsp_status nsp_default_params(nsp_params* p, int nr, int nt, int nz, double lz, double re, double dt) {
  if (!p) return fail(NSP_ERR_INVALID, "null params");
  p->nr = nr;
  p->nt = nt;
  p->nz = nz;
  p->lz = lz;
  p->re = re;
  p->dt = dt;
  p->d_im = 0.51;
  p->maxit = 10;
  p->tol = 5e-8;
  p->qdes = 0.5;
  p->precision = NSP_FP64;
  return NSP_OK;
}

// share: a handle with identical parameters whose operator data (LU factors,
// pressure homogeneous solution, influence-matrix column, flux profile) this
// one uses instead of computing its own; it must outlive this handle.
static nsp_status create_handle(const nsp_params* p, int device, const nsp_handle* share, nsp_handle** out) {
  if (!out) return fail(NSP_ERR_INVALID, "null handle pointer");
  *out = nullptr;
  nsp_status st = check_params(p);
  if (st) return st;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(NSP_ERR_HIP, "no HIP device");
  if (device < 0 || device >= ndev) return fail(NSP_ERR_INVALID, "device index out of range");
  HIPCHK(hipSetDevice(device));
  // the device's address of the allocation-free diagnostics' scratch array
  // (once per device, with the device current)
  HIPCHK(diag_scratch_init(device));

  nsp_handle* h = new nsp_handle();
  h->P = *p;
  h->device = device;
  Geom& g = h->g;
  g.nr = p->nr;
  g.nt = p->nt;
  g.nz = p->nz;
  g.modes = p->nt * p->nz;
  g.ntp = 3 * p->nt / 2;
  g.nzf = 3 * (p->nz - 1);
  g.nh = g.nzf / 2;
  g.n = size_t(g.nr) * g.modes;
  g.kz0 = PI2 / p->lz;
  g.re = p->re;
  g.dt = p->dt;
  g.d_im = p->d_im;
  g.qdes = p->qdes;
  g.skip = nullptr;
  g.lv0 = 0;
  g.nlv = g.nr;

  auto bail = [&](nsp_status s) {
    free_all(h);
    delete h;
    return s;
  };
#define TRY(x)                 \
  do {                         \
    nsp_status s_ = (x);       \
    if (s_) return bail(s_);   \
  } while (0)
#define HTRY(x)                                                                                \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) return bail(fail(NSP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_))); \
  } while (0)

  HTRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HTRY(hipStreamCreateWithFlags(&h->nb.s2, hipStreamNonBlocking));  // default priority, as h->stream
  for (hipEvent_t& e : h->nb.ev) HTRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));

  // radial tables
  build_radial_tables(g.nr, h->host);
  TRY(upload(h, &h->t.r, h->host.r));
  TRY(upload(h, &h->t.D, h->host.D));
  TRY(upload(h, &h->t.Lu, h->host.Lu));
  TRY(upload(h, &h->t.Lp, h->host.Lp));
  TRY(upload(h, &h->t.rdr, h->host.rdr));
  TRY(upload(h, &h->t.dr0, h->host.dr0));
  TRY(upload(h, &h->t.dr1, h->host.dr1));
  if (share) {
    h->t.uc = share->t.uc;
    h->t.qc = share->t.qc;
  } else {
    TRY(dalloc(h, &h->t.uc, g.nr));
    TRY(dalloc(h, &h->t.qc, 1));
  }

  // FFT plans
  std::vector<int> perm_t, perm_h;
  if (!make_fft_plan(g.ntp, h->nb.pt, perm_t) || !make_fft_plan(g.nh, h->nb.ph, perm_h))
    return bail(fail(NSP_ERR_INVALID, "transform lengths must factor into 2, 3 and 5"));
  TRY(upload(h, &h->nb.perm_t, perm_t));
  TRY(upload(h, &h->nb.perm_h, perm_h));
  TRY(upload(h, &h->nb.tw_t, twiddles(g.ntp, g.ntp)));
  TRY(upload(h, &h->nb.tw_h, twiddles(g.nh, g.nh)));
  TRY(upload(h, &h->nb.tw_z, twiddles(g.nzf, g.nh + 1)));
  h->nb.wstride = size_t(g.nr) * tp_plane(g);  // tile-major planes (tp_plane)
  h->nb.use_ct = false;
  h->nb.mixed = p->precision != NSP_FP64;
  h->nb.var = NlVariant{ThetaVar{0, 8, 256, 0, 1}, ThetaVar{0, 8, 256, 0, 1}, 1, 256, 0, 0, 1, 1};
  if (!ct_has_zreg(g.nh)) {  // long axial rows: the 512-thread LDS form
    h->nb.var.z_alg = 0;
    h->nb.var.z_nthr = 512;
  }
  const bool have_ct = ct_has_theta(g.ntp) && ct_has_zhalf(g.nh);
  if (h->nb.mixed && !have_ct)
    return bail(fail(NSP_ERR_INVALID, "the mixed form needs compiled transform plans for 3NT/2 and 3(NZ-1)/2"));
  if (have_ct && (h->nb.mixed || !getenv("NSP_GENERIC_FFT"))) {
    const auto twt = pass_twiddles(ct_radices(g.ntp, true));
    // the half-length plan's pass twiddles, then those of the pair form's
    // full-length plan and the compact table of its last pass
    auto twh = pass_twiddles(ct_radices(g.nh, false));
    const auto rpair = ct_pair_radices(g.nh);
    if (!rpair.empty()) {
      const auto twp = pass_twiddles(rpair);
      twh.insert(twh.end(), twp.begin(), twp.end());
      const auto twc = compact_twiddles(rpair);
      twh.insert(twh.end(), twc.begin(), twc.end());
    }
    if (h->nb.mixed) {
      TRY(upload(h, &h->nb.ct_tw_t32, to_float2(twt)));
      TRY(upload(h, &h->nb.ct_tw_h32, to_float2(twh)));
      TRY(upload(h, &h->nb.tw_z32, to_float2(twiddles(g.nzf, g.nh + 1))));
    } else {
      TRY(upload(h, &h->nb.ct_tw_t, twt));
      TRY(upload(h, &h->nb.ct_tw_h, twh));
    }
    h->nb.use_ct = true;
  }

  // fields
  const size_t n = g.n;
  TRY(dalloc(h, &h->u, 3 * n));
  TRY(dalloc(h, &h->ui, 3 * n));
  TRY(dalloc(h, &h->un, 3 * n));
  h->U[0] = h->u;
  h->U[1] = h->ui;
  h->U[2] = h->un;
  h->loc = 0;
  TRY(dalloc(h, &h->rec, 1));
  {
    void* ph = nullptr;
    if (hipHostMalloc(&ph, sizeof(StepRec)) != hipSuccess) return fail(NSP_ERR_NOMEM, "pinned step record");
    h->rec_h = static_cast<StepRec*>(ph);
  }
  {
    const char* ng = std::getenv("NSP_NO_GRAPH");
    h->graphs = !(ng && ng[0] && ng[0] != '0');
  }
  // N(u) and d_r u are FP32 pairs in the mixed form
  // N(u): three tile-major components (dr_cs), FP32 pairs in the mixed form
  TRY(dalloc(h, &h->Nl, h->nb.mixed ? (3 * dr_cs(g) + 1) / 2 : 3 * dr_cs(g)));
  // sweep-side layout (E, SR, SP, F, PH): per-workgroup columns, padded to
  // whole chunks of 64 axial modes; SR also serves as 3N scratch elsewhere
  TRY(dalloc(h, &h->E, 3 * sw_field(g)));
  TRY(dalloc(h, &h->SR, 3 * sw_field(g)));
  TRY(dalloc(h, &h->SP, sw_field(g)));
  if (p->precision == NSP_MIXED_REFINED) TRY(dalloc(h, &h->DH, sw_field(g)));
  if (share) {
    h->F = share->F;
    h->F32 = h->nb.mixed ? share->F32 : nullptr;  // an FP64 member never reads FP32 factors
    h->PH = share->PH;
  } else {
    TRY(dalloc(h, &h->F, factor_count(g)));
    if (h->nb.mixed) TRY(dalloc(h, &h->F32, factor32_count(g)));
    TRY(dalloc(h, &h->PH, sw_ph_count(g)));
  }
  TRY(upload(h, &h->Dc, centred_table(g.nr, h->host.D)));
  TRY(upload(h, &h->Luc, centred_table(g.nr, h->host.Lu)));
  if (share)
    h->b.invc = share->b.invc;
  else
    TRY(dalloc(h, &h->b.invc, 4 * size_t(g.modes)));
  // d_r u: three tile-major components (dr_cs), FP32 pairs in the mixed form
  TRY(dalloc(h, &h->nb.dR, h->nb.mixed ? (3 * dr_cs(g) + 1) / 2 : 3 * dr_cs(g)));
  TRY(dalloc(h, &h->norm, 2));
  TRY(dalloc(h, &h->part, 4 * size_t(g.modes)));
  HTRY(hipHostMalloc((void**)&h->norm_h, 2 * sizeof(unsigned long long), hipHostMallocDefault));
  HTRY(hipMemsetAsync(h->u, 0, 3 * n * sizeof(double2), h->stream));

  // influence matrix, homogeneous solutions and flux profile (setBCon,
  // setCublasFlux) with the reference's solver order; its full-field
  // homogeneous solutions are only needed here
  if (share) {
    HTRY(hipStreamSynchronize(h->stream));
  } else {
    double2 *uh = nullptr, *uhP = nullptr, *P = nullptr;
    double* U = nullptr;
    hipError_t e1 = hipMalloc(&uh, 3 * n * sizeof(double2));
    hipError_t e2 = hipMalloc(&uhP, 3 * n * sizeof(double2));
    hipError_t e3 = hipMalloc(&P, n * sizeof(double2));
    hipError_t e4 = hipMalloc(&U, 4 * n * sizeof(double));
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
      (void)hipFree(uh); (void)hipFree(uhP); (void)hipFree(P); (void)hipFree(U);
      return bail(fail(NSP_ERR_NOMEM, "hipMalloc failed (setup buffers)"));
    }
    h->b.uh = uh;
    h->b.uhP = uhP;
    launch_setup_influence(g, h->t, h->b, P, U, h->PH, h->stream);
    launch_factor(g, h->t.r, h->t.Lu, h->t.Lp, h->F, h->F32, h->stream);
    hipError_t e5 = hipGetLastError();
    hipError_t e6 = hipStreamSynchronize(h->stream);
    (void)hipFree(uh); (void)hipFree(uhP); (void)hipFree(P); (void)hipFree(U);
    h->b.uh = h->b.uhP = nullptr;
    if (e5 != hipSuccess || e6 != hipSuccess)
      return bail(fail(NSP_ERR_HIP, std::string("setup: ") + hipGetErrorString(e5 != hipSuccess ? e5 : e6)));
  }
  // transform intermediates after the setup buffers are freed (lower peak)
  // W carries nh spare elements: the z pass's buffer loads of a row stop at
  // its nz stored modes by the resource's range check, and the spare tail
  // keeps even the last row's unchecked extent (nh elements) inside the block.
  // W starts zeroed and no kernel writes its columns l >= nz (the padding of
  // the last 64-column tile), which the pair z pass reads as zero modes.
  if (h->nb.mixed) {
    TRY(dalloc(h, &h->nb.W32, 9 * h->nb.wstride + g.nh));
    TRY(dalloc(h, &h->nb.Y32, 3 * h->nb.wstride));
    HTRY(hipMemsetAsync(h->nb.W32, 0, (9 * h->nb.wstride + g.nh) * sizeof(*h->nb.W32), h->stream));
  } else {
    TRY(dalloc(h, &h->nb.W, 9 * h->nb.wstride + g.nh));
    TRY(dalloc(h, &h->nb.Y, 3 * h->nb.wstride));
    HTRY(hipMemsetAsync(h->nb.W, 0, (9 * h->nb.wstride + g.nh) * sizeof(*h->nb.W), h->stream));
  }
#undef TRY
#undef HTRY
  *out = h;
  return NSP_OK;
}

nsp_status nsp_create(const nsp_params* p, int device, nsp_handle** out) {
  return create_handle(p, device, nullptr, out);
}

nsp_status nsp_destroy(nsp_handle* h) {
  if (!h) return NSP_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_all(h);
  delete h;
  return NSP_OK;
}

nsp_status nsp_set_state(nsp_handle* h, const double* ur, const double* ut, const double* uz, int where) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->device));
  return copy_fields(h, h->u, ur, ut, uz, where);
}

nsp_status nsp_get_state(nsp_handle* h, double* ur, double* ut, double* uz, int where) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->device));
  return fetch_fields(h, h->u, ur, ut, uz, where);
}

void* nsp_stream(nsp_handle* h) { return h ? (void*)h->stream : nullptr; }

nsp_status nsp_synchronize(nsp_handle* h) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipDeviceSynchronize());
  return NSP_OK;
}

nsp_status nsp_nonlinear(nsp_handle* h, const double* ur, const double* ut, const double* uz, double* dr,
                         double* dt, double* dz, int where) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->device));
  nsp_status st = copy_fields(h, h->ui, ur, ut, uz, where);
  if (st) return st;
  launch_nonlinear(h->g, h->t, h->nb, h->ui, h->Nl, h->stream, nullptr);
  HIPCHK(hipGetLastError());
  // N(u) is stored tile-major (FP32 pairs in the mixed form): into the
  // reference layout in the sweep scratch (3N capacity)
  launch_untile(h->g, h->Nl, h->nb.mixed, h->SR, h->stream);
  HIPCHK(hipGetLastError());
  return fetch_fields(h, h->SR, dr, dt, dz, where);
}

nsp_status nsp_profile_enable(nsp_handle* h, int enable) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  h->prof.on = enable != 0;
  h->prof.reset();
  return NSP_OK;
}

int nsp_profile_count(void) { return K_COUNT; }

const char* nsp_profile_name(int i) { return kernel_name(i); }

nsp_status nsp_profile_read(nsp_handle* h, double* total_ms, long long* launches) {
  if (!h || !total_ms || !launches) return fail(NSP_ERR_INVALID, "null argument");
  for (int q = 0; q < K_COUNT; q++) {
    total_ms[q] = h->prof.total_ms[q];
    launches[q] = h->prof.launches[q];
  }
  return NSP_OK;
}

// Times each compiled kernel variant of the nonlinear stages on the handle's
// own sizes and current state, and keeps the fastest per stage.
nsp_status nsp_autotune(nsp_handle* h, int reps) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->device));
  if (!h->nb.use_ct) {
    h->tune_report = "{\"tuned\": false}";
    return NSP_OK;
  }
  if (reps < 1) reps = 3;
  const Geom& g = h->g;
  hipStream_t s = h->stream;
  NlBufs& nb = h->nb;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto stage = [&](int which, const NlVariant& v) -> bool {
    if (which == 0) {
      launch_dr3(g, h->t.D, h->u, nb.dR, nb.mixed, s);
      return nb.mixed ? ct_theta_inv<float>(v, g, (const cpx<float>*)nb.ct_tw_t32, h->t.r, h->u, nb.dR,
                                            (cpx<float>*)nb.W32, nb.wstride, s)
                      : ct_theta_inv<double>(v, g, (const cpx<double>*)nb.ct_tw_t, h->t.r, h->u, nb.dR,
                                             (cpx<double>*)nb.W, nb.wstride, s);
    }
    if (which == 1)
      return nb.mixed ? ct_zpass<float>(v, g, (const cpx<float>*)nb.ct_tw_h32, (const cpx<float>*)nb.tw_z32, h->t.r,
                                        (const cpx<float>*)nb.W32, nb.wstride, (cpx<float>*)nb.Y32, s)
                      : ct_zpass<double>(v, g, (const cpx<double>*)nb.ct_tw_h, (const cpx<double>*)nb.tw_z, h->t.r,
                                         (const cpx<double>*)nb.W, nb.wstride, (cpx<double>*)nb.Y, s);
    if (which == 2)
      return nb.mixed ? ct_theta_fwd<float>(v, g, (const cpx<float>*)nb.ct_tw_t32, (const cpx<float>*)nb.Y32,
                                            nb.wstride, h->t.D, h->u, h->Nl, s)
                      : ct_theta_fwd<double>(v, g, (const cpx<double>*)nb.ct_tw_t, (const cpx<double>*)nb.Y,
                                             nb.wstride, h->t.D, h->u, h->Nl, s);
    if (which == 4) {  // the whole nonlinear stage with v (chunked pipeline or not)
      const NlVariant keep = nb.var;
      nb.var = v;
      launch_nonlinear(g, h->t, nb, h->u, h->Nl, s, nullptr);
      nb.var = keep;
      return true;
    }
    // a predictor sweep of the current state into the free iterate buffer
    // (E and the iterate are recomputed by the next step)
    Lin2Args a = lin_args(h, g, h->Nl, h->u, h->ui, true, h->u);
    a.ord = v.s_ord;
    launch_sweep2(a, s);
    return true;
  };
  std::string rep = "{\"tuned\": true, \"candidates_ms\": {";
  bool first = true;
  auto time_it = [&](int which, const NlVariant& v, const char* label) -> double {
    if (!stage(which, v)) return -1.0;  // not compiled for this size
    if (hipStreamSynchronize(s) != hipSuccess) return -1.0;
    (void)hipEventRecord(e0, s);
    for (int q = 0; q < reps; q++) stage(which, v);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double t = double(ms) / reps;
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s\"%s\": %.4f", first ? "" : ", ", label, t);
    rep += buf;
    first = false;
    return t;
  };
  NlVariant best = nb.var;
  double bt = 1e30;
  // one stage at a time over all its compiled forms
  const int treg[4][2] = {{4, 192}, {8, 256}, {8, 384}, {16, 256}};
  for (int stage : {0, 2}) {
    ThetaVar& tb = stage == 0 ? best.inv : best.fwd;
    const char* nm = stage == 0 ? "theta_inv" : "theta_fwd";
    bt = 1e30;
    std::vector<ThetaVar> cands;
    for (int ord = 0; ord < 2; ord++) {
      for (int pad = 0; pad < 4; pad++)
        for (int tl : {4, 8}) cands.push_back(ThetaVar{0, tl, 256, pad, ord});
      cands.push_back(ThetaVar{0, 16, 256, 0, ord});
      cands.push_back(ThetaVar{0, 16, 512, 0, ord});
      cands.push_back(ThetaVar{0, 16, 1024, 0, ord});
      if (stage == 0) cands.push_back(ThetaVar{0, 32, 1024, 0, ord});
      cands.push_back(ThetaVar{0, 16, 256, 3, ord});
      for (const auto& tr : treg) cands.push_back(ThetaVar{1, tr[0], tr[1], 0, ord});
    }
    for (const ThetaVar& tv : cands) {
      NlVariant v = best;
      (stage == 0 ? v.inv : v.fwd) = tv;
      char lab[80];
      std::snprintf(lab, sizeof lab, "%s %s tl=%d nthr=%d pad=%d ord=%d", nm, tv.alg ? "register" : "scatter", tv.tl,
                    tv.nthr, tv.pad, tv.ord);
      const double t = time_it(stage, v, lab);
      if (t >= 0 && t < bt) { bt = t; tb = tv; }
    }
  }
  bt = 1e30;
  for (int pad = 0; pad < 3; pad++)
    for (int nthr : {256, 512}) {
      NlVariant v = best;
      v.z_alg = 0;
      v.z_nthr = nthr;
      v.z_pad = pad;
      char lab[64];
      std::snprintf(lab, sizeof lab, "zpass scatter nthr=%d pad=%d", nthr, pad);
      const double t = time_it(1, v, lab);
      if (t >= 0 && t < bt) { bt = t; best.z_alg = 0; best.z_nthr = nthr; best.z_pad = pad; }
    }
  for (int pf = 0; pf < 2; pf++) {
    NlVariant v = best;
    v.z_alg = 1;
    v.z_pf = pf;
    char lab[64];
    std::snprintf(lab, sizeof lab, "zpass register pf=%d", pf);
    const double t = time_it(1, v, lab);
    if (t >= 0 && t < bt) { bt = t; best.z_alg = 1; best.z_pf = pf; }
  }
  for (int pad = 0; pad < 3; pad++) {
    NlVariant v = best;
    v.z_alg = 2;
    v.z_pad = pad;
    char lab[64];
    std::snprintf(lab, sizeof lab, "zpass triples pad=%d", pad);
    const double t = time_it(1, v, lab);
    if (t >= 0 && t < bt) { bt = t; best.z_alg = 2; best.z_pad = pad; }
  }
  for (int alg : {3, 4})
    for (int pad = 0; pad < 3; pad++)
      for (int pf = 0; pf < 2; pf++) {
        NlVariant v = best;
        v.z_alg = alg;
        v.z_pad = pad;
        v.z_pf = pf;
        char lab[64];
        std::snprintf(lab, sizeof lab, "zpass pairs%s pad=%d pf=%d", alg == 4 ? " 2 rows" : "", pad, pf);
        const double t = time_it(1, v, lab);
        if (t >= 0 && t < bt) { bt = t; best.z_alg = alg; best.z_pad = pad; best.z_pf = pf; }
      }
  // the nonlinear stage in level chunks on two streams (z pass of one chunk
  // beside the theta inverse of the next) or stage after stage
  bt = 1e30;
  for (int kc : {1, 2, 3, 4, 6, 8}) {
    if (kc > g.nr) continue;
    NlVariant v = best;
    v.nl_chunks = kc;
    char lab[64];
    std::snprintf(lab, sizeof lab, "nonlinear chunks=%d", kc);
    const double t = time_it(4, v, lab);
    if (t >= 0 && t < bt) { bt = t; best.nl_chunks = kc; }
  }
  // linear sweep: order of its (m, -m, axial chunk) workgroups
  bt = 1e30;
  for (int so = 0; so < 2; so++) {
    NlVariant v = best;
    v.s_ord = so;
    char lab[64];
    std::snprintf(lab, sizeof lab, "sweep ord=%d", so);
    const double t = time_it(3, v, lab);
    if (t >= 0 && t < bt) { bt = t; best.s_ord = so; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  nb.var = best;
  drop_graphs(h);  // captured steps hold the old variants
  rep += "}, \"chosen\": " + variant_json(best) + "}";
  h->tune_report = rep;
  return NSP_OK;
}

const char* nsp_tuning(nsp_handle* h) { return h ? h->tune_report.c_str() : ""; }

nsp_status nsp_set_tuning(nsp_handle* h, const char* json) {
  if (!h || !json) return fail(NSP_ERR_INVALID, "null argument");
  if (!h->nb.use_ct) return NSP_OK;  // generic-plan sizes have no variants
  const std::string t(json);
  size_t base = t.find("\"chosen\"");
  if (base == std::string::npos) base = 0;
  NlVariant v = h->nb.var;
  for (int st = 0; st < 2; st++) {
    size_t at = std::string::npos;
    if (json_int(t, base, st == 0 ? "theta_inv" : "theta_fwd", &at) < 0 || at == std::string::npos)
      return fail(NSP_ERR_INVALID, "tuning text lacks theta_inv / theta_fwd objects");
    ThetaVar& tv = st == 0 ? v.inv : v.fwd;
    const long a = json_int(t, at, "alg"), tl = json_int(t, at, "tl"), n = json_int(t, at, "nthr"),
               pd = json_int(t, at, "pad"), o = json_int(t, at, "ord");
    if (a < 0 || tl < 0 || n < 0 || pd < 0 || o < 0) return fail(NSP_ERR_INVALID, "incomplete theta variant");
    tv = ThetaVar{int(a), int(tl), int(n), int(pd), int(o)};
  }
  const long za = json_int(t, base, "z_alg"), zn = json_int(t, base, "z_nthr"), zp = json_int(t, base, "z_pad"),
             zf = json_int(t, base, "z_pf");
  if (za < 0 || zn < 0 || zp < 0 || zf < 0) return fail(NSP_ERR_INVALID, "incomplete z-pass variant");
  v.z_alg = int(za);
  v.z_nthr = int(zn);
  v.z_pad = int(zp);
  v.z_pf = int(zf);
  const long so = json_int(t, base, "s_ord");  // absent in older tuning text: keep the current order
  if (so >= 0) v.s_ord = so ? 1 : 0;
  const long nc = json_int(t, base, "nl_chunks");  // absent: stage after stage
  v.nl_chunks = nc >= 1 ? int(std::min<long>(nc, kNlMaxChunks)) : 1;
  // accept only variants that are compiled for this size: one nonlinear
  // evaluation of the current state into the scratch of the nonlinear term
  const NlVariant old = h->nb.var;
  h->nb.var = v;
  HIPCHK(hipSetDevice(h->device));
  bool ok = true;
  {
    NlBufs& nb = h->nb;
    const Geom& g = h->g;
    hipStream_t s = h->stream;
    ok = nb.mixed ? ct_theta_inv<float>(v, g, (const cpx<float>*)nb.ct_tw_t32, h->t.r, h->u, nb.dR,
                                        (cpx<float>*)nb.W32, nb.wstride, s)
                  : ct_theta_inv<double>(v, g, (const cpx<double>*)nb.ct_tw_t, h->t.r, h->u, nb.dR,
                                         (cpx<double>*)nb.W, nb.wstride, s);
    ok = ok && (nb.mixed ? ct_zpass<float>(v, g, (const cpx<float>*)nb.ct_tw_h32, (const cpx<float>*)nb.tw_z32,
                                           h->t.r, (const cpx<float>*)nb.W32, nb.wstride, (cpx<float>*)nb.Y32, s)
                         : ct_zpass<double>(v, g, (const cpx<double>*)nb.ct_tw_h, (const cpx<double>*)nb.tw_z,
                                            h->t.r, (const cpx<double>*)nb.W, nb.wstride, (cpx<double>*)nb.Y, s));
    ok = ok && (nb.mixed ? ct_theta_fwd<float>(v, g, (const cpx<float>*)nb.ct_tw_t32, (const cpx<float>*)nb.Y32,
                                               nb.wstride, h->t.D, h->u, h->Nl, s)
                         : ct_theta_fwd<double>(v, g, (const cpx<double>*)nb.ct_tw_t, (const cpx<double>*)nb.Y,
                                                nb.wstride, h->t.D, h->u, h->Nl, s));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  if (!ok) {
    h->nb.var = old;
    return fail(NSP_ERR_INVALID, "tuning names a kernel variant that is not compiled for this size");
  }
  h->tune_report = "{\"tuned\": \"set\", \"chosen\": " + variant_json(v) + "}";
  drop_graphs(h);
  return NSP_OK;
}

namespace {

// Captures one whole time step starting from the state in U[a]: the predictor
// writes U[a+1], correctors alternate U[a+1] -> U[a+2] -> U[a+1] ...  Every
// corrector launch carries the converged flag and returns at once after the
// loop has ended.  The step is two graphs: the predictor with the first
// corrector, then correctors 2..maxit, which the host launches only when the
// step record says the first corrector did not end the loop (the common
// single-corrector step then carries no idle corrector launches).  Each graph
// ends by copying the step record to the host.
nsp_status capture_step(nsp_handle* h, int a) {
  hipStream_t s = h->stream;
  const Geom& g = h->g;
  Geom gk = g;
  gk.skip = &h->rec->st[0];
  double2* ua = h->U[a];
  double2* ub = h->U[(a + 1) % 3];
  double2* uc = h->U[(a + 2) % 3];
  auto corrector = [&](int it) {
    double2* src = (it % 2) ? ub : uc;
    double2* dst = (it % 2) ? uc : ub;
    launch_nonlinear(gk, h->t, h->nb, src, h->Nl, s, nullptr);
    launch_norm_reset(h->norm, gk.skip, s);
    launch_sweep2(lin_args(h, gk, h->Nl, src, dst, false, ua), s);
    launch_converge(h->norm, &h->rec->st[0], &h->rec->err, h->P.tol, it, h->P.maxit, s);
  };
  auto finish = [&](hipGraphExec_t* out, const char* what) -> nsp_status {
    (void)hipMemcpyAsync(h->rec_h, h->rec, sizeof(StepRec), hipMemcpyDeviceToHost, s);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess || !graph) return fail(NSP_ERR_HIP, std::string(what) + " capture: " + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) {
      *out = nullptr;
      return fail(NSP_ERR_HIP, std::string(what) + " instantiate: " + hipGetErrorString(e2));
    }
    return NSP_OK;
  };
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  launch_step_begin(&h->rec->st[0], &h->rec->err, s);
  launch_nonlinear(g, h->t, h->nb, ua, h->Nl, s, nullptr);
  launch_sweep2(lin_args(h, g, h->Nl, ua, ub, true, ua), s);
  if (h->P.maxit >= 1) corrector(1);
  nsp_status st = finish(&h->gexec[a], "step graph");
  if (st != NSP_OK || h->P.maxit < 2) return st;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int it = 2; it <= h->P.maxit; it++) corrector(it);
  st = finish(&h->gexec2[a], "corrector graph");
  if (st != NSP_OK) {
    (void)hipGraphExecDestroy(h->gexec[a]);
    h->gexec[a] = nullptr;
  }
  return st;
}

void set_state_slot(nsp_handle* h, int loc) {
  h->loc = loc;
  h->u = h->U[loc];
  h->ui = h->U[(loc + 1) % 3];
  h->un = h->U[(loc + 2) % 3];
}

}  // namespace

// integrate() loop body (mod_integrate.cu:112-131).  Default: one hipGraph
// launch per step (capture_step), and a second one when the first corrector
// does not end the corrector loop.  With per-kernel profiling on, or with
// NSP_NO_GRAPH=1, the kernels are launched one by one with a host-side
// convergence test after every corrector (HIP events recorded inside a
// captured graph do not yield elapsed times, so timing needs eager launches).
nsp_status nsp_step(nsp_handle* h, int nsteps, nsp_step_info* info) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  if (nsteps < 0) return fail(NSP_ERR_INVALID, "nsteps < 0");
  HIPCHK(hipSetDevice(h->device));
  hipStream_t s = h->stream;
  const Geom& g = h->g;
  Profiler* pf = h->prof.on ? &h->prof : nullptr;
  for (int step = 0; step < nsteps; step++) {
    int ite = 0, flags = 0;
    double err = 1e98;
    if (h->graphs && !h->prof.on) {
      const int a = h->loc;
      if (!h->gexec[a]) {
        const nsp_status st = capture_step(h, a);
        if (st != NSP_OK) return st;
      }
      HIPCHK(hipGraphLaunch(h->gexec[a], s));
      HIPCHK(hipStreamSynchronize(s));
      if (h->rec_h->st[0] == 0 && h->gexec2[a]) {  // the loop goes on past the first corrector
        HIPCHK(hipGraphLaunch(h->gexec2[a], s));
        HIPCHK(hipStreamSynchronize(s));
      }
      ite = h->rec_h->st[1];
      flags = h->rec_h->st[2];
      err = h->rec_h->err;
      // newest iterate: the predictor's output (no corrector) or the last corrector's
      const int newest = (ite % 2 == 1) ? (a + 2) % 3 : (a + 1) % 3;
      set_state_slot(h, newest);
    } else {
      // nonlinear(u,rhs); predictor(u,rhs,uw,rhsw,ui)
      launch_nonlinear(g, h->t, h->nb, h->u, h->Nl, s, pf);
      if (pf) pf->begin(K_LIN_PRED, s);
      launch_sweep2(lin_args(h, g, h->Nl, h->u, h->ui, true, h->u), s);
      if (pf) pf->end(s);
      HIPCHK(hipGetLastError());
      while (ite < h->P.maxit && err > h->P.tol) {
        ite += 1;
        launch_nonlinear(g, h->t, h->nb, h->ui, h->Nl, s, pf);
        launch_norm_reset(h->norm, nullptr, s);
        if (pf) pf->begin(K_LIN_CORR, s);
        launch_sweep2(lin_args(h, g, h->Nl, h->ui, h->un, false, h->u), s);
        if (pf) pf->end(s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h->norm_h, h->norm, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        err = bits_to_double(h->norm_h[0]) / bits_to_double(h->norm_h[1]);
        std::swap(h->ui, h->un);
      }
      // the last iterate becomes the state (ui); keep the U[] slots in step
      int newest = 0;
      for (int q = 0; q < 3; q++)
        if (h->U[q] == h->ui) newest = q;
      set_state_slot(h, newest);
      flags = (std::isnan(err) ? NSP_FLAG_NAN : 0) | ((!std::isnan(err) && err > h->P.tol) ? NSP_FLAG_MAXIT : 0);
      if (pf) pf->collect();
    }
    if (info) {
      info[step].n_it = ite;
      info[step].err = err;
      info[step].flags = flags;
    }
  }
  return NSP_OK;
}

nsp_status nsp_tables(nsp_handle* h, double* r, double* D, double* Lu, double* Lp, double* rdr, double* dr0,
                      double* dr1) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  const RadialTables& T = h->host;
  auto cp = [](double* dst, const std::vector<double>& v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  cp(r, T.r);
  cp(D, T.D);
  cp(Lu, T.Lu);
  cp(Lp, T.Lp);
  cp(rdr, T.rdr);
  cp(dr0, T.dr0);
  cp(dr1, T.dr1);
  return NSP_OK;
}

nsp_status nsp_radial_tables(int nr, double* r, double* D, double* Lu, double* Lp, double* rdr, double* dr0,
                             double* dr1) {
  if (nr < 2 * IW + 2) return fail(NSP_ERR_INVALID, "nr must be >= 8");
  RadialTables T;
  build_radial_tables(nr, T);
  auto cp = [](double* dst, const std::vector<double>& v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  cp(r, T.r);
  cp(D, T.D);
  cp(Lu, T.Lu);
  cp(Lp, T.Lp);
  cp(rdr, T.rdr);
  cp(dr0, T.dr0);
  cp(dr1, T.dr1);
  return NSP_OK;
}

nsp_status nsp_remesh_weights(int nr_old, int nr_new, int flags, double* RM) {
  if (nr_old < 2 * IW + 2 || nr_new < 2 * IW + 2) return fail(NSP_ERR_INVALID, "nr must be >= 8");
  if (!RM) return fail(NSP_ERR_INVALID, "null matrix");
  if (flags & ~NSP_REMESH_MIRROR) return fail(NSP_ERR_INVALID, "unknown remesh flags");
  std::vector<double> ro(nr_old), rn(nr_new);
  radial_grid(nr_old, ro.data());
  radial_grid(nr_new, rn.data());
  const std::vector<double> M = remesh_matrix(ro, rn, (flags & NSP_REMESH_MIRROR) != 0);
  std::memcpy(RM, M.data(), M.size() * sizeof(double));
  return NSP_OK;
}

nsp_status nsp_remesh(nsp_handle* dst, nsp_handle* src, int flags) {
  if (!dst || !src) return fail(NSP_ERR_INVALID, "null handle");
  if (dst == src) return fail(NSP_ERR_INVALID, "remesh needs two handles");
  if (dst->device != src->device) return fail(NSP_ERR_INVALID, "remesh handles are on different devices");
  if (flags & ~NSP_REMESH_MIRROR) return fail(NSP_ERR_INVALID, "unknown remesh flags");
  const Geom &gn = dst->g, &go = src->g;
  const std::vector<double> RM = remesh_matrix(src->host.r, dst->host.r, (flags & NSP_REMESH_MIRROR) != 0);
  // nonzero column band of each new row over both parity halves
  std::vector<int2> band(gn.nr);
  for (int i = 0; i < gn.nr; i++) {
    int lo = go.nr, hi = -1;
    for (int h2 = 0; h2 < 2; h2++)
      for (int j = 0; j < go.nr; j++)
        if (RM[((size_t)h2 * gn.nr + i) * go.nr + j] != 0.0) {
          lo = std::min(lo, j);
          hi = std::max(hi, j);
        }
    band[i] = make_int2(lo, hi);
  }
  // the matrix and bands live in dst's sweep scratch for the launch
  const size_t mbytes = RM.size() * sizeof(double), bbytes = band.size() * sizeof(int2);
  if (mbytes + bbytes > 3 * gn.n * sizeof(double2)) return fail(NSP_ERR_INVALID, "remesh matrix exceeds scratch");
  HIPCHK(hipSetDevice(dst->device));
  HIPCHK(hipStreamSynchronize(src->stream));
  char* scr = reinterpret_cast<char*>(dst->SR);
  HIPCHK(hipMemcpyAsync(scr, RM.data(), mbytes, hipMemcpyHostToDevice, dst->stream));
  HIPCHK(hipMemcpyAsync(scr + mbytes, band.data(), bbytes, hipMemcpyHostToDevice, dst->stream));
  launch_remesh(gn, go, reinterpret_cast<const double*>(scr), reinterpret_cast<const int2*>(scr + mbytes), src->u,
                dst->u, dst->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(dst->stream));
  return NSP_OK;
}

#define CHECK(x)             \
  do {                       \
    nsp_status s_ = (x);     \
    if (s_) return s_;       \
  } while (0)

// Workspace of the physical-space output: the nonlinear-term intermediates,
// which hold nothing between steps.  Wt (NR*NTP*NZ complex) first, then the
// physical rows (NR*NTP rows of NZF+2 doubles).
static nsp_status phys_scratch(nsp_handle* h, double2** Wt, double** F) {
  const Geom& g = h->g;
  char* base = h->nb.mixed ? reinterpret_cast<char*>(h->nb.W32) : reinterpret_cast<char*>(h->nb.W);
  const size_t cap = 9 * h->nb.wstride * (h->nb.mixed ? sizeof(float2) : sizeof(double2));
  const size_t wt = (size_t)g.nr * g.ntp * g.nz * sizeof(double2);
  const size_t fb = (size_t)g.nr * g.ntp * (g.nzf + 2) * sizeof(double);
  if (wt + fb > cap) return fail(NSP_ERR_INVALID, "physical-space output exceeds the workspace");
  *Wt = reinterpret_cast<double2*>(base);
  *F = reinterpret_cast<double*>(base + wt);
  return NSP_OK;
}

nsp_status nsp_physical_field(nsp_handle* h, int comp, double* out, int where) {
  if (!h || !out) return fail(NSP_ERR_INVALID, "null argument");
  if (comp < 0 || comp > 2) return fail(NSP_ERR_INVALID, "comp must be 0, 1 or 2");
  if (where != NSP_HOST && where != NSP_DEVICE) return fail(NSP_ERR_INVALID, "bad location");
  HIPCHK(hipSetDevice(h->device));
  const Geom& g = h->g;
  double2* Wt;
  double* F;
  CHECK(phys_scratch(h, &Wt, &F));
  const size_t pitch = g.nzf + 2;
  double* dst = where == NSP_DEVICE ? out : F;
  launch_physical(g, h->nb, h->u + (size_t)comp * g.n, Wt, dst, pitch, 0, h->stream);
  HIPCHK(hipGetLastError());
  if (where == NSP_HOST)
    HIPCHK(hipMemcpyAsync(out, F, (size_t)g.nr * g.ntp * pitch * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return NSP_OK;
}

nsp_status nsp_cross_sections(nsp_handle* h, double* q, double* u2) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  HIPCHK(hipSetDevice(h->device));
  const Geom& g = h->g;
  double2* Wt;
  double* F;
  CHECK(phys_scratch(h, &Wt, &F));
  const size_t pitch = g.nzf + 2;
  double* q1 = reinterpret_cast<double*>(h->SR);
  double* qd = q1 + (size_t)g.nr * g.nzf;
  for (int which = 0; which < 2; which++) {
    double* dst = which == 0 ? q : u2;
    if (!dst) continue;
    if (which == 0) {  // q0kernel_k: ur^2 + ut^2
      launch_physical(g, h->nb, h->u, Wt, F, pitch, 1, h->stream);
      launch_physical(g, h->nb, h->u + g.n, Wt, F, pitch, 2, h->stream);
    } else {  // u0kernel_k: uz^2
      launch_physical(g, h->nb, h->u + 2 * g.n, Wt, F, pitch, 1, h->stream);
    }
    launch_cross(g, F, pitch, h->t.rdr, q1, qd, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dst, qd, g.nzf * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return NSP_OK;
}

nsp_status nsp_mean_profile(nsp_handle* h, double* prof) {
  if (!h || !prof) return fail(NSP_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->device));
  const Geom& g = h->g;
  // real part of uz at mode (0,0) of every level (wrt_mnp_k)
  HIPCHK(hipMemcpy2DAsync(prof, sizeof(double), h->u + 2 * g.n, g.modes * sizeof(double2), sizeof(double), g.nr,
                          hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return NSP_OK;
}

nsp_status nsp_io_open(nsp_handle* h, const char* dir) {
  if (!h || !dir) return fail(NSP_ERR_INVALID, "null argument");
  close_io(h);
  static const char* names[4] = {"io_friction.txt", "io_qcross.txt", "io_ucross.txt", "io_meanpr.txt"};
  for (int f = 0; f < 4; f++) {
    const std::string path = std::string(dir) + "/" + names[f];
    h->io[f] = std::fopen(path.c_str(), "w");
    if (!h->io[f]) {
      close_io(h);
      return fail(NSP_ERR_IO, "cannot create " + path);
    }
  }
  // header rows: the axial sample positions (with a leading 0), the radii
  const Geom& g = h->g;
  const double lz = PI2 / g.kz0;
  for (int f = 1; f <= 2; f++) {
    std::fprintf(h->io[f], "%.8e  ", 0.0);
    for (int n = 0; n < g.nzf; n++) std::fprintf(h->io[f], "%.8e  ", lz * double(n) / double(g.nzf));
    std::fprintf(h->io[f], "\n");
  }
  std::fprintf(h->io[3], "%.8e  ", 0.0);
  for (int i = 0; i < g.nr; i++) std::fprintf(h->io[3], "%.8e  ", h->host.r[i]);
  std::fprintf(h->io[3], "\n");
  for (FILE* f : h->io) std::fflush(f);
  return NSP_OK;
}

nsp_status nsp_io_write(nsp_handle* h, double t, int mask) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  if (!h->io[0]) return fail(NSP_ERR_INVALID, "diagnostics files are not open");
  const Geom& g = h->g;
  if (mask & NSP_IO_FRICTION) {
    double d[7];
    CHECK(nsp_diagnostics(h, d));
    std::fprintf(h->io[0], "%.8e  %.8e  %.8e\n", t, d[0], d[1]);
    std::fflush(h->io[0]);
  }
  if (mask & NSP_IO_CROSS) {
    std::vector<double> q(g.nzf), u2(g.nzf);
    CHECK(nsp_cross_sections(h, q.data(), u2.data()));
    for (int f = 1; f <= 2; f++) {
      const std::vector<double>& v = f == 1 ? q : u2;
      std::fprintf(h->io[f], "%.8e  ", t);
      for (double x : v) std::fprintf(h->io[f], "%.8e  ", x);
      std::fprintf(h->io[f], "\n");
      std::fflush(h->io[f]);
    }
  }
  if (mask & NSP_IO_MEANPR) {
    std::vector<double> prof(g.nr);
    CHECK(nsp_mean_profile(h, prof.data()));
    std::fprintf(h->io[3], "%.8e  ", t);
    for (double x : prof) std::fprintf(h->io[3], "%.8e  ", x);
    std::fprintf(h->io[3], "\n");
    std::fflush(h->io[3]);
  }
  return NSP_OK;
}

nsp_status nsp_io_close(nsp_handle* h) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  close_io(h);
  return NSP_OK;
}

#undef CHECK

nsp_status nsp_diagnostics(nsp_handle* h, double out[7]) {
  if (!h || !out) return fail(NSP_ERR_INVALID, "null argument");
  HIPCHK(hipSetDevice(h->device));
  const Geom& g = h->g;
  launch_diagnostics(g, h->t, h->u, h->part, h->stream);
  HIPCHK(hipGetLastError());
  std::vector<double> part(4 * size_t(g.modes));
  std::vector<double2> col(3 * size_t(g.nr));
  HIPCHK(hipMemcpyAsync(part.data(), h->part, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  // mode (0,0) uz column for the centre-line, friction and flux values
  HIPCHK(hipMemcpy2DAsync(col.data(), sizeof(double2), h->u + 2 * g.n, g.modes * sizeof(double2), sizeof(double2),
                          g.nr, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const RadialTables& T = h->host;
  double uc = 0.0, du = 0.0, q = 0.0;
  for (int i = 0; i < IW + 1; i++) uc += col[i].x * T.dr0[i];  // mod_io_nonlinear.cu:45
  for (int i = 0; i < STEN; i++) du += col[g.nr - STEN + i].x * T.dr1[i];  // :47
  du = std::sqrt(std::fabs(du) / g.re);
  for (int i = 0; i < g.nr; i++) q += T.rdr[i] * col[i].x;
  double ke = 0.0, d2 = 0.0, wall = 0.0, wdiv = 0.0;
  for (int k = 0; k < g.modes; k++) {
    ke += part[4 * size_t(k)];
    d2 += part[4 * size_t(k) + 1];
    wall = std::max(wall, part[4 * size_t(k) + 2]);
    wdiv = std::max(wdiv, part[4 * size_t(k) + 3]);
  }
  out[0] = uc;
  out[1] = du;
  out[2] = ke;
  out[3] = std::sqrt(d2);
  out[4] = wall;
  out[5] = wdiv;
  out[6] = q;
  return NSP_OK;
}

// Same values as nsp_diagnostics, reduced entirely on the device into out
// (device memory): no allocation, no host synchronisation, no memset (every
// block writes its own partial set in the static scratch array).
nsp_status nsp_diagnostics_noalloc(nsp_handle* h, double* out) {
  if (!h || !out) return fail(NSP_ERR_INVALID, "null argument");
  const Geom& g = h->g;
  if (g.modes > kDiagMaxModes)
    return fail(NSP_ERR_INVALID, "nsp_diagnostics_noalloc: NT*NZ = " + std::to_string(g.modes) +
                                     " exceeds its scratch limit of " + std::to_string(kDiagMaxModes) +
                                     " modes (use nsp_diagnostics)");
  HIPCHK(hipSetDevice(h->device));
  double* part = diag_scratch(h->device);
  if (!part) return fail(NSP_ERR_HIP, "diagnostics scratch not initialised on this device");
  launch_diagnostics_noalloc(g, h->t, h->u, part, out, h->stream);
  HIPCHK(hipGetLastError());
  return NSP_OK;
}

// ---------------------------------------------------------------- restart files
nsp_status nsp_write_state_files(nsp_handle* h, const char* a, const char* b, const char* c) {
  if (!h || !a || !b || !c) return fail(NSP_ERR_INVALID, "null argument");
  std::vector<double> buf(2 * h->g.n);
  const char* names[3] = {a, b, c};
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int q = 0; q < 3; q++) {
    HIPCHK(hipMemcpy(buf.data(), h->u + q * h->g.n, h->g.n * sizeof(double2), hipMemcpyDeviceToHost));
    FILE* f = std::fopen(names[q], "wb");
    if (!f) return fail(NSP_ERR_IO, std::string("writing error: ") + names[q]);
    const size_t w = std::fwrite(buf.data(), sizeof(double2), h->g.n, f);
    std::fclose(f);
    if (w != h->g.n) return fail(NSP_ERR_IO, std::string("writing error: ") + names[q]);
  }
  return NSP_OK;
}

nsp_status nsp_read_state_files(nsp_handle* h, const char* a, const char* b, const char* c) {
  if (!h || !a || !b || !c) return fail(NSP_ERR_INVALID, "null argument");
  std::vector<double> buf(6 * h->g.n);
  const char* names[3] = {a, b, c};
  for (int q = 0; q < 3; q++) {
    FILE* f = std::fopen(names[q], "rb");
    if (!f) return fail(NSP_ERR_IO, std::string("reading error: ") + names[q]);
    const size_t r = std::fread(buf.data() + 2 * q * h->g.n, sizeof(double2), h->g.n, f);
    std::fclose(f);
    if (r != h->g.n) return fail(NSP_ERR_IO, std::string("reading error (short file): ") + names[q]);
  }
  return nsp_set_state(h, buf.data(), buf.data() + 2 * h->g.n, buf.data() + 4 * h->g.n, NSP_HOST);
}

// Restart directory: the three raw state files of the reference plus a
// metadata file, restart.txt, of "key value" lines.
static const char* const kRestartMeta = "restart.txt";

nsp_status nsp_write_restart(nsp_handle* h, const char* dir, double t, long long step) {
  if (!h || !dir) return fail(NSP_ERR_INVALID, "null argument");
  const std::string d(dir);
  nsp_status st = nsp_write_state_files(h, (d + "/ur.bin").c_str(), (d + "/ut.bin").c_str(), (d + "/uz.bin").c_str());
  if (st) return st;
  const std::string path = d + "/" + kRestartMeta;
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return fail(NSP_ERR_IO, "writing error: " + path);
  const nsp_params& P = h->P;
  std::fprintf(f, "format nspipe-restart-1\n");
  std::fprintf(f, "time %.17g\nstep %lld\n", t, step);
  std::fprintf(f, "nr %d\nnt %d\nnz %d\n", P.nr, P.nt, P.nz);
  std::fprintf(f, "lz %.17g\nre %.17g\ndt %.17g\nd_im %.17g\nmaxit %d\ntol %.17g\nqdes %.17g\nprecision %d\n", P.lz,
               P.re, P.dt, P.d_im, P.maxit, P.tol, P.qdes, P.precision);
  const bool ok = std::fclose(f) == 0;
  if (!ok) return fail(NSP_ERR_IO, "writing error: " + path);
  return NSP_OK;
}

nsp_status nsp_read_restart(nsp_handle* h, const char* dir, double* t, long long* step) {
  if (!h || !dir) return fail(NSP_ERR_INVALID, "null argument");
  const std::string d(dir);
  double tt = 0.0;
  long long ss = 0;
  const std::string path = d + "/" + kRestartMeta;
  // The metadata is required: the raw files alone carry no grid, and another
  // grid with the same NR*NT*NZ would read without error.  Headerless
  // reference restarts go through nsp_read_state_files.
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return fail(NSP_ERR_IO, "missing restart metadata " + path + " (raw reference files: nsp_read_state_files)");
  {
    char key[64];
    char val[128];
    int nr = -1, nt = -1, nz = -1;
    bool format_ok = false;
    while (std::fscanf(f, "%63s %127s", key, val) == 2) {
      const std::string k(key);
      if (k == "format") format_ok = std::string(val) == "nspipe-restart-1";
      else if (k == "time") tt = std::strtod(val, nullptr);
      else if (k == "step") ss = std::strtoll(val, nullptr, 10);
      else if (k == "nr") nr = std::atoi(val);
      else if (k == "nt") nt = std::atoi(val);
      else if (k == "nz") nz = std::atoi(val);
    }
    std::fclose(f);
    if (!format_ok) return fail(NSP_ERR_IO, "not a restart metadata file: " + path);
    if (nr != h->P.nr || nt != h->P.nt || nz != h->P.nz)
      return fail(NSP_ERR_INVALID, "restart grid " + std::to_string(nr) + "x" + std::to_string(nt) + "x" +
                                       std::to_string(nz) + " differs from the handle's (convert it with nsp_remesh)");
  }
  nsp_status st = nsp_read_state_files(h, (d + "/ur.bin").c_str(), (d + "/ut.bin").c_str(), (d + "/uz.bin").c_str());
  if (st) return st;
  if (t) *t = tt;
  if (step) *step = ss;
  return NSP_OK;
}

// ---------------------------------------------------------------- initial conditions
// Reference IC (initField, restart == 0, mod_io_nonlinear.cu:424-449): the
// physical perturbation of init_per is projected on the retained modes by an
// exact direct transform on the host (separable: sin/cos(theta) times a
// z profile), equal to the reference's D2Z + normalize + padBackward.
nsp_status nsp_reference_ic(nsp_handle* h, double a_p) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  const Geom& g = h->g;
  const int ntp = g.ntp, nzf = g.nzf;
  const std::vector<double>& r = h->host.r;
  std::vector<double2> f(3 * g.n, make_double2(0.0, 0.0));
  // z profile e^{-10 sin^2(pi z/L)} sampled at z_l = l/nzf, forward DFT
  std::vector<double> prof(nzf);
  for (int l = 0; l < nzf; l++) {
    const double z = double(l) / double(nzf);
    double ux = std::sin(0.5 * PI2 * z);
    ux = ux * ux;
    prof[l] = std::exp(-10.0 * ux);
  }
  std::vector<double2> zc(g.nz);
  for (int l = 0; l < g.nz; l++) {
    double sr = 0.0, si = 0.0;
    for (int n = 0; n < nzf; n++) {
      const double a = -PI2 * double((long long)l * n % nzf) / double(nzf);
      sr += prof[n] * std::cos(a);
      si += prof[n] * std::sin(a);
    }
    zc[l] = make_double2(sr / nzf, si / nzf);
  }
  // sin(theta) -> m=+1: -i/2, m=-1: +i/2 ; cos(theta) -> m=+-1: 1/2
  for (int i = 0; i < g.nr; i++) {
    const double rr = r[i];
    const double ur = (1.0 - rr * rr) * (1.0 - rr * rr);
    const double ut = ur - 4.0 * rr * rr * (1.0 - rr * rr);
    for (int l = 0; l < g.nz - 1; l++) {
      const double2 z = zc[l];
      // m = +1 (it = 1) and m = -1 (it = nt-1)
      const size_t hp = (size_t)i * g.modes + 1 * g.nz + l;
      const size_t hm = (size_t)i * g.modes + (size_t)(g.nt - 1) * g.nz + l;
      f[hp] = make_double2(a_p * ur * 0.5 * z.y, -a_p * ur * 0.5 * z.x);        // (-i/2) z
      f[hm] = make_double2(-a_p * ur * 0.5 * z.y, a_p * ur * 0.5 * z.x);        // (+i/2) z
      f[g.n + hp] = make_double2(a_p * ut * 0.5 * z.x, a_p * ut * 0.5 * z.y);
      f[g.n + hm] = make_double2(a_p * ut * 0.5 * z.x, a_p * ut * 0.5 * z.y);
    }
    f[2 * g.n + (size_t)i * g.modes].x = 1.0 - rr * rr;  // base_k
  }
  if (g.nt == 2) {  // m=+-1 would alias onto the zeroed -NT/2 row
    for (size_t q = 0; q < 2 * g.n; q++) f[q] = make_double2(0.0, 0.0);
  }
  (void)ntp;
  std::vector<double> host(6 * g.n);
  return nsp_set_state(h, (const double*)f.data(), (const double*)(f.data() + g.n),
                       (const double*)(f.data() + 2 * g.n), NSP_HOST);
}

// Seeded, analytically divergence-free, no-slip perturbation on top of the
// flux-matched Poiseuille profile (see nsp.h).  For a mode (m, kz) with
// q = |m| and w(r) = (1 - r^2)^2:
//   kz != 0, q >= 1: ur = a r^(q-1) w, ut = i sgn(m) a r^(q-1) w,
//                    uz = -4 i a r^q (1 - r^2) / kz;  (1/r)(r ur)' + i m ut/r = -4 a r^q (1-r^2)
//   kz != 0, q == 0: ur = a r w, ut = b r w, uz = i a (2w + r w') / kz
//   kz == 0, q >= 1: ur = a r^(q-1) w, ut = i (r ur)'/m, uz = c r^q (1 - r^2)
//   kz == 0, q == 0: ur = 0, ut = b r w (real), uz = 0
// Modes with l = 0, m < 0 are the conjugates of m > 0 (real field).
nsp_status nsp_seeded_ic(nsp_handle* h, unsigned long long seed, int mmax, int lmax, int kind, double rms) {
  if (!h) return fail(NSP_ERR_INVALID, "null handle");
  const Geom& g = h->g;
  const bool add = (kind & NSP_IC_ADD) != 0;
  kind &= ~NSP_IC_ADD;
  if (mmax < 0 || lmax < 0 || kind < 0 || kind > 2) return fail(NSP_ERR_INVALID, "bad IC arguments");
  mmax = std::min(mmax, g.nt / 2 - 1);
  lmax = std::min(lmax, g.nz - 2);
  const std::vector<double>& r = h->host.r;
  const std::vector<double>& rdr = h->host.rdr;
  std::mt19937_64 rng(seed);
  auto gauss = [&]() -> double2 {  // Box-Muller
    const double u1 = (double(rng() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = (double(rng() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double rad = std::sqrt(-2.0 * std::log(u1)), ang = PI2 * u2;
    return make_double2(rad * std::cos(ang), rad * std::sin(ang));
  };
  std::vector<double2> f(3 * g.n, make_double2(0.0, 0.0));
  auto cm = [](double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
  // kind 2 (localized puff): one random amplitude triple per |m|, shared by all
  // l with the spectrum exp(-(kz sigma)^2 / 2) exp(-i kz LZ/2) of a Gaussian
  // axial window of width 5R (sigma = 2.5R) centred at z = LZ/2.  The stored
  // (-m, l > 0) coefficients are the conjugates of (m, -l), so they carry the
  // conjugate amplitude (real for m = 0): every azimuthal mode is then the
  // window times a fixed radial/azimuthal shape, localized in z.
  std::vector<double2> pa(2 * mmax + 1), pb(2 * mmax + 1), pc(2 * mmax + 1);
  if (kind == 2)
    for (int m = 0; m <= mmax; m++) {
      pa[m + mmax] = gauss();
      pb[m + mmax] = gauss();
      pc[m + mmax] = gauss();
      if (m == 0) {
        pa[mmax].y = pb[mmax].y = pc[mmax].y = 0.0;
      } else {
        pa[mmax - m] = make_double2(pa[m + mmax].x, -pa[m + mmax].y);
        pb[mmax - m] = make_double2(pb[m + mmax].x, -pb[m + mmax].y);
        pc[mmax - m] = make_double2(pc[m + mmax].x, -pc[m + mmax].y);
      }
    }
  const double sigma = 2.5, lz = PI2 / g.kz0;
  for (int l = 0; l <= lmax; l++) {
    for (int m = -mmax; m <= mmax; m++) {
      if (l == 0 && m < 0) continue;
      const double kz = g.kz0 * double(l);
      const int q = std::abs(m);
      double2 env = make_double2(1.0, 0.0);
      if (kind == 1) {
        const double kk = std::sqrt(kz * kz + double(m * m));
        env.x = kk > 0 ? std::pow(kk, -5.0 / 3.0) : 0.0;
      } else if (kind == 2) {
        const double amp = std::exp(-0.5 * kz * kz * sigma * sigma);
        env = make_double2(amp * std::cos(kz * 0.5 * lz), -amp * std::sin(kz * 0.5 * lz));
      }
      double2 a, b, c;
      if (kind == 2) {
        a = pa[m + mmax];
        b = pb[m + mmax];
        c = pc[m + mmax];
      } else {
        a = gauss();
        b = gauss();
        c = gauss();
      }
      a = cm(a, env);
      b = cm(b, env);
      c = cm(c, env);
      if (l == 0) {  // kz = 0 modes of a real field: keep a real radial structure
        if (m == 0) { a = make_double2(0, 0); b.y = 0.0; c = make_double2(0, 0); }
      }
      const int it = m >= 0 ? m : m + g.nt;
      for (int i = 0; i < g.nr; i++) {
        const double rr = r[i], w = (1 - rr * rr) * (1 - rr * rr), wp = -4.0 * rr * (1 - rr * rr);
        double2 ur, ut, uz;
        if (l > 0 && q >= 1) {
          const double s = std::pow(rr, q - 1) * w;
          ur = make_double2(a.x * s, a.y * s);
          const double sg = m > 0 ? 1.0 : -1.0;
          ut = make_double2(-sg * a.y * s, sg * a.x * s);
          const double t = -4.0 * std::pow(rr, q) * (1 - rr * rr) / kz;
          uz = make_double2(-a.y * t, a.x * t);  // i * a * t
        } else if (l > 0) {
          ur = make_double2(a.x * rr * w, a.y * rr * w);
          ut = make_double2(b.x * rr * w, b.y * rr * w);
          const double t = (2 * w + rr * wp) / kz;
          uz = make_double2(-a.y * t, a.x * t);
        } else if (q >= 1) {
          const double s = std::pow(rr, q - 1) * w;
          ur = make_double2(a.x * s, a.y * s);
          // (r ur)' = a [q r^(q-1) w + r^q w']
          const double dd = (q * std::pow(rr, q - 1) * w + std::pow(rr, q) * wp) / double(m);
          ut = make_double2(-a.y * dd, a.x * dd);  // i a dd
          const double t = std::pow(rr, q) * (1 - rr * rr);
          uz = make_double2(c.x * t, c.y * t);
        } else {
          ur = make_double2(0, 0);
          ut = make_double2(b.x * rr * w, 0.0);
          uz = make_double2(0, 0);
        }
        const size_t hh = (size_t)i * g.modes + (size_t)it * g.nz + l;
        f[hh] = ur;
        f[g.n + hh] = ut;
        f[2 * g.n + hh] = uz;
        if (l == 0 && m > 0) {  // conjugate partner
          const size_t hc = (size_t)i * g.modes + (size_t)(g.nt - m) * g.nz;
          f[hc] = make_double2(ur.x, -ur.y);
          f[g.n + hc] = make_double2(ut.x, -ut.y);
          f[2 * g.n + hc] = make_double2(uz.x, -uz.y);
        }
      }
    }
  }
  // rms of the perturbation: sqrt(sum_modes w_l sum_r rdr |u|^2)
  double e = 0.0;
  for (int i = 0; i < g.nr; i++)
    for (int k = 0; k < g.modes; k++) {
      const double wl = (k % g.nz == 0) ? 1.0 : 2.0;
      for (int c2 = 0; c2 < 3; c2++) {
        const double2 v = f[c2 * g.n + (size_t)i * g.modes + k];
        e += wl * rdr[i] * (v.x * v.x + v.y * v.y);
      }
    }
  const double scale = e > 0 ? rms / std::sqrt(e) : 0.0;
  for (auto& v : f) v = make_double2(v.x * scale, v.y * scale);
  if (add) {  // perturbation only, added to the current state on the device
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipMemcpyAsync(h->SR, f.data(), 3 * g.n * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    launch_add(h->u, h->SR, 3 * g.n, h->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return NSP_OK;
  }
  // flux-matched Poiseuille
  double q0 = 0.0;
  for (int i = 0; i < g.nr; i++) q0 += rdr[i] * (1 - r[i] * r[i]);
  const double alpha = h->P.qdes / q0;
  for (int i = 0; i < g.nr; i++) f[2 * g.n + (size_t)i * g.modes].x += alpha * (1 - r[i] * r[i]);
  return nsp_set_state(h, (const double*)f.data(), (const double*)(f.data() + g.n),
                       (const double*)(f.data() + 2 * g.n), NSP_HOST);
}


// ------------------------------------------------------------------ ensembles
// Independent pipes stepped together: every stage is one launch per class of
// members with the same compiled transform plans and precision (the radial
// derivative and the four sweep passes: one launch for all members), with
// blockIdx.y selecting the member through device tables of per-member
// arguments.  Members converge independently: each has its own converged flag
// in the corrector launches.  The ensemble step is captured like a pipe's
// (two graphs, see ens_capture); the tables (which state buffer plays which
// role) are refreshed from pinned host memory by the first graph's first
// node, so the same graphs serve every step.
struct nsp_ensemble {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<nsp_handle*> m;      // members in caller order
  std::vector<int> owner;          // member whose operator data it uses (itself if none)
  struct Cls {
    std::vector<int> idx;          // members (caller indices), table order
    int off = 0;                   // first table slot
    int max_nr = 0, max_nz = 0, max_modes = 0, max_nt = 0;
  };
  std::vector<Cls> cls;
  std::vector<int> order;          // table slot -> member
  int max_nr = 0, max_nt = 0, max_nz = 0, max_modes = 0, maxit = 0;
  bool fp64 = false, mixed = false;  // some member is of that form
  double tol = 0.0;
  // device: NL tables per input role [3][M], sweep tables [3][M], records [M], norm pointers [M]
  NlMember* nl = nullptr;
  Lin2Args* sw = nullptr;
  StepRec* rec = nullptr;
  unsigned long long** norms = nullptr;
  // pinned host staging of the tables, and of the records read back
  char* stage_h = nullptr;
  size_t stage_bytes = 0;
  StepRec* rec_h = nullptr;
  hipGraphExec_t gexec = nullptr;   // tables, predictor, first corrector
  hipGraphExec_t gexec2 = nullptr;  // correctors 2..maxit (when a member needs them)
};

namespace {

void ens_drop_graphs(nsp_ensemble* e) {
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->gexec2) (void)hipGraphExecDestroy(e->gexec2);
  e->gexec = e->gexec2 = nullptr;
}

void ens_free(nsp_ensemble* e) {
  ens_drop_graphs(e);
  if (e->nl) (void)hipFree(e->nl);
  if (e->rec) (void)hipFree(e->rec);
  if (e->norms) (void)hipFree(e->norms);
  if (e->stage_h) (void)hipHostFree(e->stage_h);
  if (e->rec_h) (void)hipHostFree(e->rec_h);
  e->nl = nullptr;
  e->sw = nullptr;
  e->rec = nullptr;
  e->norms = nullptr;
  e->stage_h = nullptr;
  e->rec_h = nullptr;
  // members that use another's operator data go first
  for (size_t j = 0; j < e->m.size(); j++)
    if (e->m[j] && e->owner[j] != (int)j) {
      (void)nsp_destroy(e->m[j]);
      e->m[j] = nullptr;
    }
  for (size_t j = 0; j < e->m.size(); j++)
    if (e->m[j]) (void)nsp_destroy(e->m[j]);
  e->m.clear();
  if (e->stream) (void)hipStreamDestroy(e->stream);
  e->stream = nullptr;
}

bool same_pipe(const nsp_params& a, const nsp_params& b) {
  return a.nr == b.nr && a.nt == b.nt && a.nz == b.nz && a.lz == b.lz && a.re == b.re && a.dt == b.dt &&
         a.d_im == b.d_im && a.precision == b.precision;
}

NlMember nl_member(nsp_handle* h, const Geom& g, const double2* u) {
  NlMember mb;
  mb.g = g;
  mb.r = h->t.r;
  mb.D = h->t.D;
  mb.u = u;
  mb.dR = h->nb.dR;
  mb.W = h->nb.mixed ? (void*)h->nb.W32 : (void*)h->nb.W;
  mb.Y = h->nb.mixed ? (void*)h->nb.Y32 : (void*)h->nb.Y;
  mb.N = h->Nl;
  mb.ws = h->nb.wstride;
  mb.f32 = h->nb.mixed ? 1 : 0;
  return mb;
}

// Table contents for the members' current state buffers, into pinned staging.
void ens_stage(nsp_ensemble* e) {
  const int M = (int)e->m.size();
  NlMember* nl = reinterpret_cast<NlMember*>(e->stage_h);
  Lin2Args* sw = reinterpret_cast<Lin2Args*>(e->stage_h + 3 * M * sizeof(NlMember));
  for (int j = 0; j < M; j++) {
    nsp_handle* h = e->m[e->order[j]];
    const int a = h->loc;
    double2* ua = h->U[a];
    double2* ub = h->U[(a + 1) % 3];
    double2* uc = h->U[(a + 2) % 3];
    Geom g0 = h->g, gk = h->g;
    g0.skip = nullptr;
    gk.skip = &e->rec[j].st[0];
    nl[0 * M + j] = nl_member(h, g0, ua);
    nl[1 * M + j] = nl_member(h, gk, ub);
    nl[2 * M + j] = nl_member(h, gk, uc);
    sw[0 * M + j] = lin_args(h, g0, h->Nl, ua, ub, true, ua);
    sw[1 * M + j] = lin_args(h, gk, h->Nl, ub, uc, false, ua);
    sw[2 * M + j] = lin_args(h, gk, h->Nl, uc, ub, false, ua);
  }
}

// One nonlinear evaluation of every member, input role `role`.
bool ens_nonlinear(nsp_ensemble* e, int role, hipStream_t s) {
  const int M = (int)e->m.size();
  NlMember* nl = e->nl + role * M;
  Batch all{nl, M, e->max_nr, e->max_nz, e->max_modes, e->max_nt};
  nsp_handle* h0 = e->m[e->order[0]];
  launch_dr3(h0->g, nullptr, nullptr, nullptr, false, s, &all);
  bool ok = true;
  for (const auto& c : e->cls) {
    nsp_handle* h = e->m[c.idx[0]];
    const NlBufs& nb = h->nb;
    const Geom& g = h->g;
    Batch bt{nl + c.off, (int)c.idx.size(), c.max_nr, c.max_nz, c.max_modes, c.max_nt};
    if (nb.mixed) {
      ok &= ct_theta_inv<float>(nb.var, g, (const cpx<float>*)nb.ct_tw_t32, nullptr, nullptr, nullptr, nullptr, 0, s,
                                &bt);
      ok &= ct_zpass<float>(nb.var, g, (const cpx<float>*)nb.ct_tw_h32, (const cpx<float>*)nb.tw_z32, nullptr,
                            nullptr, 0, nullptr, s, &bt);
      ok &= ct_theta_fwd<float>(nb.var, g, (const cpx<float>*)nb.ct_tw_t32, nullptr, 0, nullptr, nullptr, nullptr, s,
                                &bt);
    } else {
      ok &= ct_theta_inv<double>(nb.var, g, (const cpx<double>*)nb.ct_tw_t, nullptr, nullptr, nullptr, nullptr, 0, s,
                                 &bt);
      ok &= ct_zpass<double>(nb.var, g, (const cpx<double>*)nb.ct_tw_h, (const cpx<double>*)nb.tw_z, nullptr,
                             nullptr, 0, nullptr, s, &bt);
      ok &= ct_theta_fwd<double>(nb.var, g, (const cpx<double>*)nb.ct_tw_t, nullptr, 0, nullptr, nullptr, nullptr, s,
                                 &bt);
    }
  }
  return ok;
}

// Two graphs, as for one pipe: the table refresh, predictor and first
// corrector, then correctors 2..maxit, launched only when some member's loop
// goes on; both end by copying the step records to the host.
nsp_status ens_capture(nsp_ensemble* e) {
  const int M = (int)e->m.size();
  hipStream_t s = e->stream;
  bool ok = true;
  auto corrector = [&](int it) {
    const int role = (it % 2) ? 1 : 2;
    ok &= ens_nonlinear(e, role, s);
    launch_ens_norm_reset(e->norms, e->rec, M, s);
    launch_sweep_batch(e->sw + role * M, M, 0, e->max_nt, e->max_nz, e->fp64, e->mixed, s);
    launch_ens_converge(e->norms, e->rec, M, e->tol, it, e->maxit, s);
  };
  auto finish = [&](hipGraphExec_t* out) -> nsp_status {
    (void)hipMemcpyAsync(e->rec_h, e->rec, M * sizeof(StepRec), hipMemcpyDeviceToHost, s);
    hipGraph_t graph = nullptr;
    const hipError_t err = hipStreamEndCapture(s, &graph);
    if (err != hipSuccess || !graph) return fail(NSP_ERR_HIP, std::string("ensemble capture: ") + hipGetErrorString(err));
    if (!ok) {
      (void)hipGraphDestroy(graph);
      return fail(NSP_ERR_INVALID, "an ensemble class names a kernel variant that is not compiled");
    }
    const hipError_t e2 = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) {
      *out = nullptr;
      return fail(NSP_ERR_HIP, std::string("ensemble graph instantiate: ") + hipGetErrorString(e2));
    }
    return NSP_OK;
  };
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  (void)hipMemcpyAsync(e->nl, e->stage_h, e->stage_bytes, hipMemcpyHostToDevice, s);
  launch_ens_begin(e->rec, M, s);
  ok &= ens_nonlinear(e, 0, s);
  launch_sweep_batch(e->sw, M, 1, e->max_nt, e->max_nz, e->fp64, e->mixed, s);
  if (e->maxit >= 1) corrector(1);
  nsp_status st = finish(&e->gexec);
  if (st != NSP_OK || e->maxit < 2) return st;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int it = 2; it <= e->maxit; it++) corrector(it);
  st = finish(&e->gexec2);
  if (st != NSP_OK) ens_drop_graphs(e);
  return st;
}

}  // namespace

nsp_status nsp_ensemble_create(const nsp_params* params, int count, int device, nsp_ensemble** out) {
  if (!out || !params || count < 1) return fail(NSP_ERR_INVALID, "bad ensemble arguments");
  *out = nullptr;
  for (int j = 0; j < count; j++)
    if (params[j].precision == NSP_MIXED_REFINED)
      return fail(NSP_ERR_INVALID, "ensemble members take NSP_FP64 or NSP_MIXED (NSP_MIXED_REFINED: single pipes)");
  nsp_ensemble* e = new nsp_ensemble();
  e->device = device;
  auto bail = [&](nsp_status st) {
    ens_free(e);
    delete e;
    return st;
  };
  e->m.assign(count, nullptr);
  e->owner.assign(count, 0);
  for (int j = 0; j < count; j++) {
    int own = j;
    for (int q = 0; q < j; q++)
      if (e->owner[q] == q && same_pipe(params[q], params[j])) {
        own = q;
        break;
      }
    e->owner[j] = own;
    const nsp_status st = create_handle(&params[j], device, own == j ? nullptr : e->m[own], &e->m[j]);
    if (st) return bail(st);
    if (!e->m[j]->nb.use_ct) return bail(fail(NSP_ERR_INVALID, "ensemble members need compiled transform plans"));
    if (params[j].maxit != params[0].maxit || params[j].tol != params[0].tol)
      return bail(fail(NSP_ERR_INVALID, "ensemble members must share maxit and tol"));
  }
  e->maxit = params[0].maxit;
  e->tol = params[0].tol;
  // classes: same compiled plans and precision
  for (int j = 0; j < count; j++) {
    const Geom& g = e->m[j]->g;
    int c = -1;
    for (size_t q = 0; q < e->cls.size(); q++) {
      const nsp_handle* f = e->m[e->cls[q].idx[0]];
      if (f->g.ntp == g.ntp && f->g.nh == g.nh && f->nb.mixed == e->m[j]->nb.mixed) c = (int)q;
    }
    if (c < 0) {
      e->cls.emplace_back();
      c = (int)e->cls.size() - 1;
    }
    nsp_ensemble::Cls& C = e->cls[c];
    C.idx.push_back(j);
    C.max_nr = std::max(C.max_nr, g.nr);
    C.max_nz = std::max(C.max_nz, g.nz);
    C.max_modes = std::max(C.max_modes, g.modes);
    C.max_nt = std::max(C.max_nt, g.nt);
    e->max_nr = std::max(e->max_nr, g.nr);
    e->max_nt = std::max(e->max_nt, g.nt);
    e->max_nz = std::max(e->max_nz, g.nz);
    e->max_modes = std::max(e->max_modes, g.modes);
    (e->m[j]->nb.mixed ? e->mixed : e->fp64) = true;
  }
  for (auto& C : e->cls) {
    C.off = (int)e->order.size();
    for (int j : C.idx) e->order.push_back(j);
    // one variant per class (a later nsp_ensemble_autotune may change it)
    for (int j : C.idx) e->m[j]->nb.var = e->m[C.idx[0]]->nb.var;
  }
  if (hipSetDevice(device) != hipSuccess) return bail(fail(NSP_ERR_HIP, "hipSetDevice"));
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(NSP_ERR_HIP, "stream"));
  const int M = count;
  e->stage_bytes = 3 * M * sizeof(NlMember) + 3 * M * sizeof(Lin2Args);
  void* p = nullptr;
  if (hipMalloc(&p, e->stage_bytes) != hipSuccess) return bail(fail(NSP_ERR_NOMEM, "ensemble tables"));
  e->nl = static_cast<NlMember*>(p);
  e->sw = reinterpret_cast<Lin2Args*>(static_cast<char*>(p) + 3 * M * sizeof(NlMember));
  if (hipMalloc(&p, M * sizeof(StepRec)) != hipSuccess) return bail(fail(NSP_ERR_NOMEM, "ensemble records"));
  e->rec = static_cast<StepRec*>(p);
  if (hipMalloc(&p, M * sizeof(unsigned long long*)) != hipSuccess) return bail(fail(NSP_ERR_NOMEM, "norms"));
  e->norms = static_cast<unsigned long long**>(p);
  if (hipHostMalloc(&p, e->stage_bytes) != hipSuccess) return bail(fail(NSP_ERR_NOMEM, "pinned staging"));
  e->stage_h = static_cast<char*>(p);
  if (hipHostMalloc(&p, M * sizeof(StepRec)) != hipSuccess) return bail(fail(NSP_ERR_NOMEM, "pinned records"));
  e->rec_h = static_cast<StepRec*>(p);
  std::vector<unsigned long long*> nptr(M);
  for (int j = 0; j < M; j++) nptr[j] = e->m[e->order[j]]->norm;
  if (hipMemcpy(e->norms, nptr.data(), M * sizeof(unsigned long long*), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(NSP_ERR_HIP, "norm table"));
  *out = e;
  return NSP_OK;
}

nsp_status nsp_ensemble_destroy(nsp_ensemble* e) {
  if (!e) return NSP_OK;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_free(e);
  delete e;
  return NSP_OK;
}

int nsp_ensemble_size(nsp_ensemble* e) { return e ? (int)e->m.size() : 0; }

nsp_handle* nsp_ensemble_member(nsp_ensemble* e, int i) {
  if (!e || i < 0 || i >= (int)e->m.size()) return nullptr;
  return e->m[i];
}

nsp_status nsp_ensemble_autotune(nsp_ensemble* e, int reps) {
  if (!e) return fail(NSP_ERR_INVALID, "null ensemble");
  for (const auto& C : e->cls) {
    nsp_handle* f = e->m[C.idx[0]];
    const nsp_status st = nsp_autotune(f, reps);
    if (st) return st;
    for (int j : C.idx) {
      e->m[j]->nb.var = f->nb.var;
      drop_graphs(e->m[j]);
    }
  }
  ens_drop_graphs(e);
  return NSP_OK;
}

nsp_status nsp_ensemble_step(nsp_ensemble* e, int nsteps, nsp_step_info* info) {
  if (!e) return fail(NSP_ERR_INVALID, "null ensemble");
  if (nsteps < 0) return fail(NSP_ERR_INVALID, "nsteps < 0");
  HIPCHK(hipSetDevice(e->device));
  const int M = (int)e->m.size();
  for (int j = 0; j < M; j++) HIPCHK(hipStreamSynchronize(e->m[j]->stream));  // caller work on members
  if (!e->gexec) {
    const nsp_status st = ens_capture(e);
    if (st) return st;
  }
  for (int step = 0; step < nsteps; step++) {
    ens_stage(e);
    HIPCHK(hipGraphLaunch(e->gexec, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    bool more = false;
    for (int j = 0; j < M; j++) more |= e->rec_h[j].st[0] == 0;
    if (more && e->gexec2) {  // some member's loop goes on past the first corrector
      HIPCHK(hipGraphLaunch(e->gexec2, e->stream));
      HIPCHK(hipStreamSynchronize(e->stream));
    }
    for (int j = 0; j < M; j++) {
      const int mi = e->order[j];
      nsp_handle* h = e->m[mi];
      const StepRec& r = e->rec_h[j];
      const int a = h->loc;
      set_state_slot(h, (r.st[1] % 2 == 1) ? (a + 2) % 3 : (a + 1) % 3);
      if (info) {
        info[(size_t)step * M + mi].n_it = r.st[1];
        info[(size_t)step * M + mi].err = r.err;
        info[(size_t)step * M + mi].flags = r.st[2];
      }
    }
  }
  return NSP_OK;
}
}  // extern "C"
