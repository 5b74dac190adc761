// C ABI of libwignerhip.so (include/wignerhip.h): plans, argument validation,
// launch sequencing of the fused sweeps, hipGraph capture of the step loop.
//
// Replaces the reference's FFI layer cufft_wrapper.py (plans over libcufft.so)
// and the per-step host orchestration of GPU_Wigner2D_FFT.Run /
// GPU_Wigner2D_GPitaevskii.Run (GPU_Wigner2D_FFT.py:1754-1964, GP.py:1809-2055).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/wignerhip.h"
#include "wh_kernels.hpp"

namespace wh {
bool size_supported(int n);
int xsweep_rows(int n);
int psweep_cols(int prec, int n, int cl);
hipError_t launch_xsweep(int prec, const XSweepArgs& a, hipStream_t s);
hipError_t launch_psweep(int prec, const PSweepArgs& a, hipStream_t s);
hipError_t launch_colsum(int prec, const void* W, int X, int P, const double* p_host,
                         double* colsums, hipStream_t s);
hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s);
hipError_t launch_scale(int prec, void* W, size_t n, const double* f, hipStream_t s);
hipError_t launch_fft_rows(int prec, int N, int dir, void* d, int rows, int norm, const void* tw,
                           hipStream_t s);
hipError_t launch_fft_cols(int prec, int N, int dir, void* d, int X, int norm, const void* tw,
                           hipStream_t s);
hipError_t launch_set_int(int* p, int v, hipStream_t s);
hipError_t launch_set_scale(double* p, double re, double im, hipStream_t s);
hipError_t launch_gp_shift(int X, int P, const double* rho, const double* lambda_host,
                           const double* theta_host, double* G, const void* twX64, hipStream_t s);
hipError_t launch_colpart_reduce(const double* colpart, int nblk, int X, double factor, double* rho,
                                 double* obs, int obs_base, const int* step, double dx,
                                 hipStream_t s);
}  // namespace wh

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  int code = WH_EXEC_FAILED;
  if (e == hipErrorOutOfMemory) code = WH_ALLOC_FAILED;
  if (e == hipErrorNoDevice) code = WH_NO_DEVICE;
  if (e == hipErrorInvalidValue || e == hipErrorInvalidConfiguration) code = WH_LAUNCH_FAILED;
  return fail(code, "%s: %s", what, hipGetErrorString(e));
}

#define WH_HIP(call, what)                      \
  do {                                          \
    hipError_t e_ = (call);                     \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// device allocation owned by a plan
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    bytes = n;
    hipError_t e = hipMalloc(&p, n ? n : 16);
    if (e != hipSuccess) {
      p = nullptr;
      return hip_fail(e, "hipMalloc");
    }
    return WH_SUCCESS;
  }
  int upload(const void* host, size_t n) {
    int rc = alloc(n);
    if (rc) return rc;
    hipError_t e = hipMemcpy(p, host, n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy H2D");
    return WH_SUCCESS;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// exp(-2 pi i n / N), n = 0..N-1, evaluated in long double and rounded
int make_twiddles(DevBuf& buf, int N, int prec) {
  const long double two_pi = 6.283185307179586476925286766559L;
  if (prec == WH_C128) {
    std::vector<double> h(2 * size_t(N));
    for (int n = 0; n < N; ++n) {
      const long double a = two_pi * (long double)n / (long double)N;
      h[2 * n] = (double)cosl(a);
      h[2 * n + 1] = (double)(-sinl(a));
    }
    return buf.upload(h.data(), h.size() * sizeof(double));
  }
  std::vector<float> h(2 * size_t(N));
  for (int n = 0; n < N; ++n) {
    const long double a = two_pi * (long double)n / (long double)N;
    h[2 * n] = (float)cosl(a);
    h[2 * n + 1] = (float)(-sinl(a));
  }
  return buf.upload(h.data(), h.size() * sizeof(float));
}

int upload_opt(DevBuf& buf, const double* host, size_t n) {
  if (!host) return WH_SUCCESS;
  return buf.upload(host, n * sizeof(double));
}

}  // namespace

// ============================================================================ library
extern "C" const char* wh_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* wh_status_string(int s) {
  switch (s) {
    case WH_SUCCESS: return "success";
    case WH_INVALID_PLAN: return "invalid plan handle";
    case WH_ALLOC_FAILED: return "device allocation failed";
    case WH_INVALID_TYPE: return "unsupported transform type / precision";
    case WH_INVALID_VALUE: return "invalid argument value";
    case WH_INTERNAL_ERROR: return "internal error";
    case WH_EXEC_FAILED: return "execution failed on the device";
    case WH_SETUP_FAILED: return "library set-up failed";
    case WH_INVALID_SIZE: return "unsupported transform size";
    case WH_UNALIGNED_DATA: return "data not aligned for the transform";
    case WH_INVALID_DESCRIPTOR: return "invalid descriptor";
    case WH_LAUNCH_FAILED: return "kernel launch failed";
    case WH_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

extern "C" int wh_version(void) { return (0 << 16) | 3; }

// ============================================================================ FFT plans
struct wh_fft_plan {
  int n = 0, type = 0, batch = 0;
  bool columns = false;  // interleaved columns of an [n][batch] array
  DevBuf tw;
};

extern "C" int wh_fft_plan_many(wh_fft_plan** plan, int n, int istride, int idist, int type,
                                int batch) {
  if (!plan) return fail(WH_INVALID_VALUE, "wh_fft_plan_many: plan is NULL");
  *plan = nullptr;
  if (type != WH_C128 && type != WH_C64)
    return fail(WH_INVALID_TYPE, "wh_fft_plan_many: type %d is neither WH_C128 nor WH_C64", type);
  if (!wh::size_supported(n))
    return fail(WH_INVALID_SIZE, "wh_fft_plan_many: length %d is not a power of two in [64, 2048]", n);
  if (batch < 1) return fail(WH_INVALID_VALUE, "wh_fft_plan_many: batch %d < 1", batch);
  bool columns;
  if (istride == 1 && idist == n) {
    columns = false;
  } else if (istride == batch && idist == 1) {
    columns = true;
    const int c = wh::psweep_cols(type, n, 0);
    if (batch % c != 0)
      return fail(WH_INVALID_SIZE, "wh_fft_plan_many: column batch %d is not a multiple of %d", batch, c);
  } else {
    return fail(WH_INVALID_VALUE,
                "wh_fft_plan_many: layout istride=%d idist=%d is neither rows (1, n) nor columns (batch, 1)",
                istride, idist);
  }
  auto* p = new wh_fft_plan;
  p->n = n;
  p->type = type;
  p->batch = batch;
  p->columns = columns;
  int rc = make_twiddles(p->tw, n, type);
  if (rc) {
    delete p;
    return rc;
  }
  *plan = p;
  return WH_SUCCESS;
}

static int fft_exec(wh_fft_plan* plan, void* idata, void* odata, int dir, int norm, void* stream) {
  if (!plan) return fail(WH_INVALID_PLAN, "wh_fft_exec: plan is NULL");
  if (!idata || idata != odata)
    return fail(WH_INVALID_VALUE, "wh_fft_exec: only in-place transforms (idata == odata != NULL)");
  if (dir != WH_FORWARD && dir != WH_INVERSE)
    return fail(WH_INVALID_VALUE, "wh_fft_exec: direction %d", dir);
  const size_t align = plan->type == WH_C128 ? 16 : 8;
  if (reinterpret_cast<uintptr_t>(idata) % align)
    return fail(WH_UNALIGNED_DATA, "wh_fft_exec: data not %zu-byte aligned", align);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = plan->columns
                     ? wh::launch_fft_cols(plan->type, plan->n, dir, idata, plan->batch, norm, plan->tw.p, s)
                     : wh::launch_fft_rows(plan->type, plan->n, dir, idata, plan->batch, norm, plan->tw.p, s);
  if (e != hipSuccess) return hip_fail(e, "wh_fft_exec launch");
  return WH_SUCCESS;
}

extern "C" int wh_fft_exec(wh_fft_plan* plan, void* idata, void* odata, int direction, void* stream) {
  return fft_exec(plan, idata, odata, direction, 0, stream);
}
extern "C" int wh_fft_exec_normalized(wh_fft_plan* plan, void* idata, void* odata, int direction,
                                      void* stream) {
  return fft_exec(plan, idata, odata, direction, 1, stream);
}
extern "C" int wh_fft_destroy(wh_fft_plan* plan) {
  if (!plan) return fail(WH_INVALID_PLAN, "wh_fft_destroy: plan is NULL");
  delete plan;
  return WH_SUCCESS;
}

// ============================================================================ Wigner2D plan
struct GraphKey {
  void* W;
  double* obs;
  int first_step;
  hipStream_t stream;
  bool operator<(const GraphKey& o) const {
    return std::tie(W, obs, first_step, stream) < std::tie(o.W, o.obs, o.first_step, o.stream);
  }
};
struct GraphEntry {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int steps = 0;
};

struct wh_wigner2d {
  wh_wigner2d_desc d{};
  DevBuf twX, twP, twX64, pot, theta_decay, lambda_decay, x_mask, sigma_p, theta_host, lambda_host;
  DevBuf x_host, p_host, dv_host, v0_host;
  DevBuf colsums, colpart, scale, step, rho, G;
  hipStream_t capture_stream = nullptr;
  std::map<GraphKey, GraphEntry> graphs;
  int graph_chunk = 16;
  ~wh_wigner2d() {
    for (auto& kv : graphs) {
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
      if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }
};

static int validate_desc(const wh_wigner2d_desc* d) {
  if (!d) return fail(WH_INVALID_VALUE, "descriptor is NULL");
  if (d->struct_size != (int)sizeof(wh_wigner2d_desc))
    return fail(WH_INVALID_DESCRIPTOR, "descriptor struct_size %d != %d", d->struct_size,
                (int)sizeof(wh_wigner2d_desc));
  if (d->precision != WH_C128 && d->precision != WH_C64)
    return fail(WH_INVALID_TYPE, "precision %d is neither WH_C128 nor WH_C64", d->precision);
  if (!wh::size_supported(d->x_dim) || !wh::size_supported(d->p_dim))
    return fail(WH_INVALID_SIZE, "grid %d x %d: both sides must be powers of two in [64, 2048]",
                d->p_dim, d->x_dim);
  if (d->p_dim % wh::xsweep_rows(d->x_dim) != 0)
    return fail(WH_INVALID_SIZE, "P=%d must be a multiple of %d for X=%d", d->p_dim,
                wh::xsweep_rows(d->x_dim), d->x_dim);
  if (d->x_dim % wh::psweep_cols(d->precision, d->p_dim, d->damping) != 0)
    return fail(WH_INVALID_SIZE, "X=%d must be a multiple of %d for P=%d", d->x_dim,
                wh::psweep_cols(d->precision, d->p_dim, d->damping), d->p_dim);
  if (d->scheme != WH_SCHEME_LIE && d->scheme != WH_SCHEME_STRANG)
    return fail(WH_INVALID_DESCRIPTOR, "scheme %d", d->scheme);
  if (d->pot_deg < 0 || d->pot_deg > WH_MAX_DEG || d->pot_rows < 1 || !d->pot_coef)
    return fail(WH_INVALID_DESCRIPTOR, "potential: deg %d (max %d), rows %d, coef %p", d->pot_deg,
                WH_MAX_DEG, d->pot_rows, (const void*)d->pot_coef);
  if (d->kinetic_poly && (d->kinetic_deg < 0 || d->kinetic_deg > WH_MAX_DEG))
    return fail(WH_INVALID_DESCRIPTOR, "kinetic degree %d", d->kinetic_deg);
  if (!(d->dt == d->dt) || d->hbar == 0.0 || d->mass == 0.0)
    return fail(WH_INVALID_DESCRIPTOR, "dt/hbar/mass: %g %g %g", d->dt, d->hbar, d->mass);
  if (d->damping != 0 && d->damping != 1)
    return fail(WH_INVALID_DESCRIPTOR, "damping %d (0 none, 1 Caldeira-Leggett)", d->damping);
  if (d->damping == 1 && (!d->sigma_p || !d->theta_host))
    return fail(WH_INVALID_DESCRIPTOR, "Caldeira-Leggett needs sigma_p and theta_host");
  if (d->gp_coeff != 0.0 && (!d->lambda_host || !d->theta_host))
    return fail(WH_INVALID_DESCRIPTOR, "Gross-Pitaevskii needs lambda_host and theta_host");
  if (!d->x_host || !d->p_host)
    return fail(WH_INVALID_DESCRIPTOR, "observables need x_host and p_host");
  return WH_SUCCESS;
}

extern "C" int wh_wigner2d_create(const wh_wigner2d_desc* desc, wh_wigner2d** out) {
  if (!out) return fail(WH_INVALID_VALUE, "wh_wigner2d_create: out is NULL");
  *out = nullptr;
  int rc = validate_desc(desc);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(WH_NO_DEVICE, "wh_wigner2d_create: no HIP device");
  auto* p = new wh_wigner2d;
  p->d = *desc;
  const auto& d = *desc;
  const size_t X = size_t(d.x_dim), P = size_t(d.p_dim);
  std::vector<double> zeros(std::max(X, P), 0.0);
#define TRY(expr)   \
  do {              \
    rc = (expr);    \
    if (rc) {       \
      delete p;     \
      return rc;    \
    }               \
  } while (0)
  TRY(make_twiddles(p->twX, d.x_dim, d.precision));
  TRY(make_twiddles(p->twP, d.p_dim, d.precision));
  TRY(p->pot.upload(d.pot_coef, size_t(d.pot_rows) * (WH_MAX_DEG + 1) * sizeof(double)));
  TRY(upload_opt(p->theta_decay, d.theta_decay, P));
  TRY(upload_opt(p->lambda_decay, d.lambda_decay, X));
  TRY(upload_opt(p->x_mask, d.x_mask, X));
  TRY(upload_opt(p->sigma_p, d.sigma_p, P));
  TRY(upload_opt(p->theta_host, d.theta_host, P));
  TRY(upload_opt(p->lambda_host, d.lambda_host, X));
  TRY(p->x_host.upload(d.x_host, X * sizeof(double)));
  TRY(p->p_host.upload(d.p_host, P * sizeof(double)));
  TRY(p->dv_host.upload(d.dv_host ? d.dv_host : zeros.data(), X * sizeof(double)));
  TRY(p->v0_host.upload(d.v0_host ? d.v0_host : zeros.data(), X * sizeof(double)));
  TRY(p->colsums.alloc(X * wh::kNCol * sizeof(double)));
  TRY(p->scale.alloc(4 * sizeof(double)));
  TRY(p->step.alloc(4 * sizeof(int)));
  if (d.gp_coeff != 0.0) {
    TRY(make_twiddles(p->twX64, d.x_dim, WH_C128));
    TRY(p->rho.alloc(X * sizeof(double)));
    TRY(p->G.alloc(X * P * sizeof(double)));
    TRY(p->colpart.alloc((P / size_t(wh::xsweep_rows(d.x_dim))) * X * sizeof(double)));
  }
#undef TRY
  *out = p;
  return WH_SUCCESS;
}

extern "C" int wh_wigner2d_destroy(wh_wigner2d* plan) {
  if (!plan) return fail(WH_INVALID_PLAN, "wh_wigner2d_destroy: plan is NULL");
  delete plan;
  return WH_SUCCESS;
}

namespace {

wh::XSweepArgs xargs(wh_wigner2d* p, void* W, const double* scale) {
  const auto& d = p->d;
  wh::XSweepArgs a{};
  a.W = W;
  a.X = d.x_dim;
  a.P = d.p_dim;
  a.tw = p->twX.p;
  a.scale = scale;
  a.kinetic_poly = d.kinetic_poly;
  for (int n = 0; n <= wh::kMaxDeg; ++n) a.kcoef[n] = d.kin_coef[n];
  a.kdeg = d.kinetic_deg;
  a.dt = d.dt;
  a.inv_mass = 1.0 / d.mass;
  a.hbar = d.hbar;
  a.inv_hbar = 1.0 / d.hbar;
  a.dp = d.dp;
  a.dlambda = d.dlambda;
  a.lambda_decay = p->lambda_decay.as<double>();
  a.colpart = nullptr;
  return a;
}

wh::PSweepArgs pargs(wh_wigner2d* p, void* W, const double* scale, double frac) {
  const auto& d = p->d;
  wh::PSweepArgs a{};
  a.W = W;
  a.X = d.x_dim;
  a.P = d.p_dim;
  a.tw = p->twP.p;
  a.scale = scale;
  a.step = p->step.as<int>();
  a.step_add = 0;
  a.pot.coef = p->pot.as<double>();
  a.pot.rows = d.pot_rows;
  a.pot.deg = d.pot_deg;
  a.vphase = frac * d.dt / d.hbar;
  a.hbar_half = 0.5 * d.hbar;
  a.dx = d.dx;
  a.dtheta = d.dtheta;
  a.theta_decay = p->theta_decay.as<double>();
  a.xmask = p->x_mask.as<double>();
  a.gp_G = nullptr;
  a.gp_phase = -frac * d.dt * d.gp_coeff / d.hbar;
  a.cl = 0;
  a.sigma_p = p->sigma_p.as<double>();
  a.theta_host = p->theta_host.as<double>();
  a.cl_a = d.gamma_dt;
  a.colsums = nullptr;
  a.p_host = p->p_host.as<double>();
  a.merge = 0;
  return a;
}

wh::ReduceArgs rargs(wh_wigner2d* p, double* obs, int obs_base, int renorm) {
  const auto& d = p->d;
  wh::ReduceArgs r{};
  r.colsums = p->colsums.as<double>();
  r.X = d.x_dim;
  r.x_host = p->x_host.as<double>();
  r.dv_host = p->dv_host.as<double>();
  r.v0_host = p->v0_host.as<double>();
  r.dxdp = d.dxdp;
  r.inv_2m = d.inv_2m;
  r.renorm = renorm;
  r.obs = obs;
  r.obs_base = obs_base;
  r.record = obs ? 1 : 0;
  r.step = p->step.as<int>();
  r.step_inc = nullptr;
  r.scale_out = p->scale.as<double>();
  r.rho_out = nullptr;
  r.rho_factor = d.gp_rho_factor;
  r.dx = d.dx_host;
  return r;
}

#define WH_LAUNCH(expr, what)                     \
  do {                                            \
    hipError_t e_ = (expr);                       \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// ---- Lie:  [obs, K, V(+GP), CL, renorm] per iteration (GPU_Wigner2D_FFT.py:1843-1913)
// One iteration after the first observable row: K sweep (with the pending
// renormalisation factor folded into its load), V sweep (+ column sums),
// reduce (record row, next factor, step counter += 1).
int lie_iteration(wh_wigner2d* p, void* W, double* obs, int first_step, hipStream_t s) {
  const auto& d = p->d;
  const int prec = d.precision;
  const bool gp = d.gp_coeff != 0.0;
  if (gp) {
    // shifted densities of the pre-step state (FFT.py:1845-1848)
    WH_LAUNCH(wh::launch_gp_shift(d.x_dim, d.p_dim, p->rho.as<double>(),
                                  p->lambda_host.as<double>(), p->theta_host.as<double>(),
                                  p->G.as<double>(), p->twX64.p, s),
              "gp_shift");
  }
  WH_LAUNCH(wh::launch_xsweep(prec, xargs(p, W, p->scale.as<double>()), s), "xsweep");
  wh::PSweepArgs pa = pargs(p, W, nullptr, 1.0);
  pa.colsums = p->colsums.as<double>();
  pa.cl = d.damping == 1;
  if (gp) pa.gp_G = p->G.as<double>();
  WH_LAUNCH(wh::launch_psweep(prec, pa, s), "psweep");
  wh::ReduceArgs r = rargs(p, obs, first_step - 1, d.renormalize);
  r.step_inc = p->step.as<int>();
  if (gp) r.rho_out = p->rho.as<double>();
  WH_LAUNCH(wh::launch_reduce(r, s), "reduce");
  return WH_SUCCESS;
}

// ---- Strang: V/2, K, V/2, CL, renorm, obs (GP.py:1898-1999).  Without GP the
// second half step of iteration k and the first of k+1 share one p sweep, the
// observables being taken between them.
int strang_iteration(wh_wigner2d* p, void* W, double* obs, int first_step, bool merge,
                     hipStream_t s) {
  const auto& d = p->d;
  const int prec = d.precision;
  const bool gp = d.gp_coeff != 0.0;
  if (!gp) {
    WH_LAUNCH(wh::launch_xsweep(prec, xargs(p, W, p->scale.as<double>()), s), "xsweep");
    wh::PSweepArgs pa = pargs(p, W, nullptr, 0.5);
    pa.colsums = p->colsums.as<double>();
    pa.cl = d.damping == 1;
    pa.merge = merge ? 1 : 0;
    WH_LAUNCH(wh::launch_psweep(prec, pa, s), "psweep");
  } else {
    // V/2 + GP/2 with the density of the current state (rho, G prepared by the last reduce)
    WH_LAUNCH(wh::launch_gp_shift(d.x_dim, d.p_dim, p->rho.as<double>(),
                                  p->lambda_host.as<double>(), p->theta_host.as<double>(),
                                  p->G.as<double>(), p->twX64.p, s),
              "gp_shift");
    wh::PSweepArgs pa = pargs(p, W, p->scale.as<double>(), 0.5);
    pa.gp_G = p->G.as<double>();
    WH_LAUNCH(wh::launch_psweep(prec, pa, s), "psweep V/2");
    // K with partial column sums -> density after the kinetic step (GP.py:1933-1939)
    wh::XSweepArgs xa = xargs(p, W, nullptr);
    xa.colpart = p->colpart.as<double>();
    WH_LAUNCH(wh::launch_xsweep(prec, xa, s), "xsweep");
    const int nblk = d.p_dim / wh::xsweep_rows(d.x_dim);
    // density after K and the ProbabilitySquare record of this iteration
    WH_LAUNCH(wh::launch_colpart_reduce(p->colpart.as<double>(), nblk, d.x_dim, d.gp_rho_factor,
                                        p->rho.as<double>(), obs, first_step, p->step.as<int>(),
                                        d.dx_host, s),
              "colpart_reduce");
    WH_LAUNCH(wh::launch_gp_shift(d.x_dim, d.p_dim, p->rho.as<double>(),
                                  p->lambda_host.as<double>(), p->theta_host.as<double>(),
                                  p->G.as<double>(), p->twX64.p, s),
              "gp_shift");
    wh::PSweepArgs pb = pargs(p, W, nullptr, 0.5);
    pb.gp_G = p->G.as<double>();
    pb.colsums = p->colsums.as<double>();
    pb.cl = d.damping == 1;
    WH_LAUNCH(wh::launch_psweep(prec, pb, s), "psweep V/2");
  }
  wh::ReduceArgs r = rargs(p, obs, first_step, d.renormalize);
  r.step_inc = p->step.as<int>();
  if (gp) r.rho_out = p->rho.as<double>();
  WH_LAUNCH(wh::launch_reduce(r, s), "reduce");
  return WH_SUCCESS;
}

int run_iterations(wh_wigner2d* p, void* W, double* obs, int first_step, int count, bool merge_last,
                   hipStream_t s) {
  for (int k = 0; k < count; ++k) {
    int rc;
    if (p->d.scheme == WH_SCHEME_LIE)
      rc = lie_iteration(p, W, obs, first_step, s);
    else
      rc = strang_iteration(p, W, obs, first_step, (k < count - 1) || merge_last, s);
    if (rc) return rc;
  }
  return WH_SUCCESS;
}

// Replays `count` iterations (all with the merge flag of the middle of a run)
// from a captured graph of p->graph_chunk iterations.
int run_graph(wh_wigner2d* p, void* W, double* obs, int first_step, int count, hipStream_t s) {
  const int chunk = p->graph_chunk;
  if (count >= chunk) {
    GraphKey key{W, obs, first_step, s};
    auto it = p->graphs.find(key);
    if (it == p->graphs.end()) {
      if (!p->capture_stream)
        WH_HIP(hipStreamCreateWithFlags(&p->capture_stream, hipStreamNonBlocking), "hipStreamCreate");
      GraphEntry g;
      WH_HIP(hipStreamBeginCapture(p->capture_stream, hipStreamCaptureModeThreadLocal), "BeginCapture");
      int rc = run_iterations(p, W, obs, first_step, chunk, true, p->capture_stream);
      hipError_t e = hipStreamEndCapture(p->capture_stream, &g.graph);
      if (rc) return rc;
      if (e != hipSuccess) return hip_fail(e, "EndCapture");
      WH_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "GraphInstantiate");
      g.steps = chunk;
      it = p->graphs.emplace(key, g).first;
    }
    while (count >= chunk) {
      WH_HIP(hipGraphLaunch(it->second.exec, s), "GraphLaunch");
      count -= chunk;
    }
  }
  return run_iterations(p, W, obs, first_step, count, true, s);
}

}  // namespace

extern "C" int wh_wigner2d_run(wh_wigner2d* p, void* W, int first_step, int n_steps, double* obs,
                               int normalize_initial, int use_graph, void* stream) {
  if (!p) return fail(WH_INVALID_PLAN, "wh_wigner2d_run: plan is NULL");
  if (!W) return fail(WH_INVALID_VALUE, "wh_wigner2d_run: W is NULL");
  if (n_steps < 0 || first_step < 0) return fail(WH_INVALID_VALUE, "wh_wigner2d_run: n_steps %d first_step %d", n_steps, first_step);
  const size_t align = p->d.precision == WH_C128 ? 16 : 8;
  if (reinterpret_cast<uintptr_t>(W) % align) return fail(WH_UNALIGNED_DATA, "W not %zu-byte aligned", align);
  const auto& d = p->d;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int prec = d.precision;
  const size_t cells = size_t(d.x_dim) * size_t(d.p_dim);
  WH_LAUNCH(wh::launch_set_int(p->step.as<int>(), first_step, s), "set_step");
  // state before the first iteration: norm (and Lie row 0)
  WH_LAUNCH(wh::launch_colsum(prec, W, d.x_dim, d.p_dim, p->p_host.as<double>(), p->colsums.as<double>(), s), "colsum");
  {
    const bool lie = d.scheme == WH_SCHEME_LIE;
    wh::ReduceArgs r = rargs(p, lie ? obs : nullptr, first_step, normalize_initial);
    if (d.gp_coeff != 0.0) r.rho_out = p->rho.as<double>();
    WH_LAUNCH(wh::launch_reduce(r, s), "reduce");
  }
  if (n_steps == 0) {
    if (normalize_initial) WH_LAUNCH(wh::launch_scale(prec, W, cells, p->scale.as<double>(), s), "scale");
    return WH_SUCCESS;
  }
  if (d.scheme == WH_SCHEME_STRANG && d.gp_coeff == 0.0) {
    // first half step of the first iteration, then the merged middle
    wh::PSweepArgs pa = pargs(p, W, p->scale.as<double>(), 0.5);
    WH_LAUNCH(wh::launch_psweep(prec, pa, s), "psweep first half");
    WH_LAUNCH(wh::launch_set_scale(p->scale.as<double>(), 1.0, 0.0, s), "set_scale");
  }
  int rc;
  if (d.scheme == WH_SCHEME_STRANG && d.gp_coeff == 0.0) {
    if (use_graph)
      rc = run_graph(p, W, obs, first_step, n_steps - 1, s);
    else
      rc = run_iterations(p, W, obs, first_step, n_steps - 1, true, s);
    if (rc) return rc;
    rc = run_iterations(p, W, obs, first_step, 1, false, s);  // last iteration: no merge
  } else {
    rc = use_graph ? run_graph(p, W, obs, first_step, n_steps, s)
                   : run_iterations(p, W, obs, first_step, n_steps, false, s);
  }
  if (rc) return rc;
  if (d.renormalize) WH_LAUNCH(wh::launch_scale(prec, W, cells, p->scale.as<double>(), s), "scale");
  return WH_SUCCESS;
}

extern "C" int wh_wigner2d_kinetic(wh_wigner2d* p, void* W, void* stream) {
  if (!p) return fail(WH_INVALID_PLAN, "wh_wigner2d_kinetic: plan is NULL");
  if (!W) return fail(WH_INVALID_VALUE, "wh_wigner2d_kinetic: W is NULL");
  WH_LAUNCH(wh::launch_xsweep(p->d.precision, xargs(p, W, nullptr), static_cast<hipStream_t>(stream)), "xsweep");
  return WH_SUCCESS;
}

extern "C" int wh_wigner2d_potential(wh_wigner2d* p, void* W, int step, int fused, void* stream) {
  if (!p) return fail(WH_INVALID_PLAN, "wh_wigner2d_potential: plan is NULL");
  if (!W || step < 0) return fail(WH_INVALID_VALUE, "wh_wigner2d_potential: W %p step %d", W, step);
  hipStream_t s = static_cast<hipStream_t>(stream);
  WH_LAUNCH(wh::launch_set_int(p->step.as<int>(), step, s), "set_step");
  const double frac = p->d.scheme == WH_SCHEME_LIE ? 1.0 : 0.5;
  wh::PSweepArgs pa = pargs(p, W, nullptr, frac);
  if (fused) {
    pa.colsums = p->colsums.as<double>();
    pa.cl = p->d.damping == 1;
  }
  WH_LAUNCH(wh::launch_psweep(p->d.precision, pa, s), "psweep");
  return WH_SUCCESS;
}

extern "C" int wh_wigner2d_observe(wh_wigner2d* p, const void* W, double* obs_row, void* stream) {
  if (!p) return fail(WH_INVALID_PLAN, "wh_wigner2d_observe: plan is NULL");
  if (!W || !obs_row) return fail(WH_INVALID_VALUE, "wh_wigner2d_observe: W %p obs %p", W, (void*)obs_row);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const auto& d = p->d;
  WH_LAUNCH(wh::launch_set_int(p->step.as<int>(), 0, s), "set_step");
  WH_LAUNCH(wh::launch_colsum(d.precision, W, d.x_dim, d.p_dim, p->p_host.as<double>(), p->colsums.as<double>(), s), "colsum");
  wh::ReduceArgs r = rargs(p, obs_row, 0, 0);
  r.scale_out = nullptr;
  WH_LAUNCH(wh::launch_reduce(r, s), "reduce");
  return WH_SUCCESS;
}
