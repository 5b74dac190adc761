// C ABI of libwignerhip.so (include/wignerhip.h): plans, argument validation,
// launch sequencing of the fused sweeps, hipGraph capture of the step loop.
//
// Replaces the reference's FFI layer cufft_wrapper.py (plans over libcufft.so)
// and the per-step host orchestration of GPU_Wigner2D_FFT.Run /
// GPU_Wigner2D_GPitaevskii.Run (GPU_Wigner2D_FFT.py:1754-1964, GP.py:1809-2055).
// Everything propagates an ensemble; a single grid is an ensemble of one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/wignerhip.h"
#include "wh_bloch.hpp"
#include "wh_dirac.hpp"
#include "wh_kernels.hpp"
#include "wh_variants.hpp"

namespace wh {
hipError_t launch_xsweep_f64(int variant, const SweepArgs& a, int nwork, hipStream_t s);
hipError_t launch_xsweep_f32(int variant, const SweepArgs& a, int nwork, hipStream_t s);
hipError_t launch_psweep_f64(int variant, bool cl, const SweepArgs& a, int nwork, hipStream_t s);
hipError_t launch_psweep_f32(int variant, bool cl, const SweepArgs& a, int nwork, hipStream_t s);
hipError_t launch_expr_table(const Member* members, int member, int cells, int depth, const int* step,
                             int step_add, int use_arg, int ntables, double hbar_half, hipStream_t s);
hipError_t launch_gpshift(const Member* members, const unsigned* work, int nwork, hipStream_t s);
hipError_t launch_colsum(int prec, const Member* members, const unsigned* work, int nwork, hipStream_t s);
hipError_t launch_observe_noalloc(int prec, Member m, hipStream_t s);
int noalloc_max_x();
hipError_t launch_xsweep_f32w(int variant, const SweepArgs& a, int nwork, hipStream_t s);
hipError_t launch_psweep_f32w(int variant, bool cl, const SweepArgs& a, int nwork, hipStream_t s);
hipError_t launch_reduce(const ReduceArgs& a, int nmembers, int* done, hipStream_t s);
hipError_t launch_colpart_reduce(const Member* members, int nmembers, int xradix, int xblock, const int* step,
                                 hipStream_t s);
hipError_t launch_scale(int prec, const Member* members, const unsigned* work, int nwork, hipStream_t s);
hipError_t launch_set_int(int* p, int v, hipStream_t s);
hipError_t launch_bind(const BindArgs& a, hipStream_t s);
hipError_t launch_dirac_bind(DiracPlan* d, void* W, double* obs, int obs_base, int* step, int step_v, double* scale,
                             int reset_scale, hipStream_t s);
hipError_t launch_set_scale(const Member* members, int nmembers, double re, double im, hipStream_t s);
hipError_t launch_fft_axis(int prec, int columns, int N, int dir, void* d, int batch, int norm,
                           const void* tw_hi, const void* tw_lo, hipStream_t s);
int fft_cols_batch_multiple(int prec, int N);
hipError_t launch_adj_psweep(int prec, const SweepArgs& a, int nwork, hipStream_t s);
hipError_t launch_grad_reduce(const Member* members, int nmembers, double* grad, double scale, hipStream_t s);
int adj_cols(int P);
int odm_rows_per_tile(int X);
int odm_cols_per_tile(int P);
hipError_t launch_odm(int prec, const Member* members, const unsigned* work_rows, int nrows,
                      const unsigned* work_cols, int ncols, double weight, double sgn, hipStream_t s);
hipError_t launch_dirac_xsweep(int prec, int X, const DiracPlan* d, int nblocks, hipStream_t s);
hipError_t launch_dirac_psweep(int prec, int P, bool cl, const DiracPlan* d, int nblocks, hipStream_t s);
hipError_t launch_dirac_colsum(int prec, const DiracPlan* d, int X, hipStream_t s);
hipError_t launch_dirac_reduce(const DiracPlan* d, int state_add, int renorm, int anti, int inc, hipStream_t s);
hipError_t launch_dirac_expr_table(const DiracPlan* d, int cells, int depth, hipStream_t s);
hipError_t launch_dirac_expr_obs(const DiracPlan* d, int X, int depth, int state_add, hipStream_t s);
hipError_t launch_dirac_mass_table(const DiracPlan* d, int cells, int depth, hipStream_t s);
hipError_t launch_dirac_a_table(const DiracPlan* d, int cells, int depth, hipStream_t s);
hipError_t launch_dirac_a_sum(const DiracPlan* d, int X, int depth, int state_add, hipStream_t s);
hipError_t launch_dirac_finish(int prec, const DiracPlan* d, int P, hipStream_t s);
int dirac_x_blocks(int prec, int X, int P);
int dirac_gp_blocks(int X, int P);
int dirac_pg_blocks(int X, int P);
hipError_t launch_dirac_pgeneral(int prec, bool cl, const DiracPlan* d, int nblocks, hipStream_t s);
hipError_t launch_dirac_podm(int prec, const DiracPlan* d, int nblocks, hipStream_t s);
hipError_t launch_dirac_fsweep(int prec, int X, const DiracPlan* d, int nblocks, hipStream_t s);
hipError_t launch_dirac_gpshift(const DiracPlan* d, int nblocks, hipStream_t s);
int dirac_p_blocks(int X, int P);
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

#define WH_HIP(call, what)                           \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)
#define WH_TRY(expr)     \
  do {                   \
    int rc_ = (expr);    \
    if (rc_) return rc_; \
  } while (0)
#define WH_LAUNCH(expr, what) WH_HIP(expr, what)

bool size_supported(int n) { return n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048; }
bool sweep_size_supported(int n) { return n >= 128 && n <= 2048 && size_supported(n); }
// bytes of one complex element of the state, and the kernel code of its storage type
size_t state_bytes(int prec) { return prec == WH_C128 ? 16 : 8; }
int storage_code(int prec) { return prec == WH_C128 ? WH_C128 : WH_C64; }

// device allocation owned by a plan
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    if (p) (void)hipFree(p);
    bytes = n;
    hipError_t e = hipMalloc(&p, n ? n : 16);
    if (e != hipSuccess) {
      p = nullptr;
      return hip_fail(e, "hipMalloc");
    }
    return WH_SUCCESS;
  }
  int upload(const void* host, size_t n) {
    WH_TRY(alloc(n));
    WH_HIP(hipMemcpy(p, host, n, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return WH_SUCCESS;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// exp(-2 pi i n / N), n = 0..N-1, evaluated in long double; FP32 tables carry
// the rounding error of the hi part in a second (lo) table
struct TwSet {
  DevBuf hi, lo, hi64;
};
int make_twiddles(TwSet& tw, int N, int prec, bool need64) {
  const long double two_pi = 6.283185307179586476925286766559L;
  std::vector<long double> c(N), s(N);
  for (int n = 0; n < N; ++n) {
    const long double a = two_pi * (long double)n / (long double)N;
    c[n] = cosl(a);
    s[n] = -sinl(a);
  }
  if ((prec != WH_C64 || need64) && !tw.hi64.p) {
    std::vector<double> h(2 * size_t(N));
    for (int n = 0; n < N; ++n) {
      h[2 * n] = (double)c[n];
      h[2 * n + 1] = (double)s[n];
    }
    WH_TRY(tw.hi64.upload(h.data(), h.size() * sizeof(double)));
  }
  if (prec == WH_C64 && !tw.hi.p) {
    std::vector<float> hi(2 * size_t(N)), lo(2 * size_t(N));
    for (int n = 0; n < N; ++n) {
      hi[2 * n] = (float)c[n];
      hi[2 * n + 1] = (float)s[n];
      lo[2 * n] = (float)(c[n] - (long double)hi[2 * n]);
      lo[2 * n + 1] = (float)(s[n] - (long double)hi[2 * n + 1]);
    }
    WH_TRY(tw.hi.upload(hi.data(), hi.size() * sizeof(float)));
    WH_TRY(tw.lo.upload(lo.data(), lo.size() * sizeof(float)));
  }
  return WH_SUCCESS;
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

extern "C" int wh_version(void) { return (0 << 16) | 4; }

extern "C" int wh_variant_info(int table, int i, int* radix, int* lanes) {
  const wh::SweepVariant* t;
  int n;
  switch (table) {
    case 0: t = wh::kXVariants; n = wh::kNumX; break;
    case 1: t = wh::kPVariants; n = wh::kNumP; break;
    case 2: t = wh::kPCLVariants; n = wh::kNumPCL; break;
    default: return -fail(WH_INVALID_VALUE, "wh_variant_info: table %d (0 x, 1 p, 2 p+CL)", table);
  }
  if (i < 0) return n;
  if (i >= n) return -fail(WH_INVALID_VALUE, "wh_variant_info: index %d >= %d", i, n);
  if (radix) *radix = t[i].R;
  if (lanes) *lanes = t[i].block;
  return WH_SUCCESS;
}

// ============================================================================ FFT plans
// A plan is one axis transform (rank 1: batch rows, or batch interleaved columns of an
// [n][batch] array) or a 2-D transform of batch contiguous [n][n2] arrays (rank 2: the
// columns pass, then the rows pass).
struct wh_fft_plan {
  int rank = 1;
  int n = 0, n2 = 0, type = 0, batch = 0;
  bool columns = false;  // rank 1: interleaved columns of an [n][batch] array
  TwSet tw, tw2;         // rank 2: tw for the length-n columns, tw2 for the length-n2 rows
};

namespace {
int fft_plan_make(wh_fft_plan** plan, int rank, int n, int n2, bool columns, int type, int batch) {
  auto p = std::make_unique<wh_fft_plan>();
  p->rank = rank;
  p->n = n;
  p->n2 = n2;
  p->type = type;
  p->batch = batch;
  p->columns = columns;
  WH_TRY(make_twiddles(p->tw, n, type, false));
  if (rank == 2) WH_TRY(make_twiddles(p->tw2, n2, type, false));
  *plan = p.release();
  return WH_SUCCESS;
}
int check_column_batch(const char* who, int type, int n, int batch) {
  const int c = wh::fft_cols_batch_multiple(type, n);
  if (c <= 0 || batch % c != 0)
    return fail(WH_INVALID_SIZE, "%s: a column transform of length %d needs a batch that is a multiple of %d (got %d)",
                who, n, c, batch);
  return WH_SUCCESS;
}
}  // namespace

extern "C" int wh_fft_plan_many(wh_fft_plan** plan, int n, int istride, int idist, int type, int batch) {
  if (!plan) return fail(WH_INVALID_VALUE, "wh_fft_plan_many: plan is NULL");
  *plan = nullptr;
  if (type != WH_C128 && type != WH_C64)
    return fail(WH_INVALID_TYPE, "wh_fft_plan_many: type %d is neither WH_C128 nor WH_C64", type);
  if (!size_supported(n))
    return fail(WH_INVALID_SIZE, "wh_fft_plan_many: length %d is not a power of two in [64, 2048]", n);
  if (batch < 1) return fail(WH_INVALID_VALUE, "wh_fft_plan_many: batch %d < 1", batch);
  bool columns;
  if (istride == 1 && (idist == n || batch == 1)) {
    columns = false;  // one transform: the distance between transforms is never used
  } else if (istride == batch && idist == 1) {
    columns = true;
    WH_TRY(check_column_batch("wh_fft_plan_many", type, n, batch));
  } else {
    return fail(WH_INVALID_VALUE,
                "wh_fft_plan_many: layout istride=%d idist=%d is neither rows (1, n) nor columns (batch, 1)",
                istride, idist);
  }
  return fft_plan_make(plan, 1, n, 0, columns, type, batch);
}

extern "C" int wh_cufft_plan_many(wh_fft_plan** plan, int rank, const int* n, const int* inembed, int istride,
                                  int idist, const int* onembed, int ostride, int odist, int type, int batch) {
  const char* who = "wh_cufft_plan_many";
  if (!plan) return fail(WH_INVALID_VALUE, "%s: plan is NULL", who);
  *plan = nullptr;
  int prec;
  if (type == WH_CUFFT_Z2Z) {
    prec = WH_C128;
  } else if (type == WH_CUFFT_C2C) {
    prec = WH_C64;
  } else {
    return fail(WH_INVALID_TYPE, "%s: type %#x is neither CUFFT_Z2Z (0x69) nor CUFFT_C2C (0x29)", who, type);
  }
  if (rank != 1 && rank != 2) return fail(WH_INVALID_VALUE, "%s: rank %d (1 or 2 supported)", who, rank);
  if (!n) return fail(WH_INVALID_VALUE, "%s: n is NULL", who);
  if (batch < 1) return fail(WH_INVALID_VALUE, "%s: batch %d < 1", who, batch);
  for (int k = 0; k < rank; ++k)
    if (!size_supported(n[k]))
      return fail(WH_INVALID_SIZE, "%s: n[%d] = %d is not a power of two in [64, 2048]", who, k, n[k]);
  if ((inembed == nullptr) != (onembed == nullptr))
    return fail(WH_INVALID_VALUE, "%s: in-place transforms need inembed and onembed both NULL or both set", who);
  if (!inembed) {
    // basic layout: batch contiguous arrays, strides and distances ignored
    if (rank == 2) {
      WH_TRY(check_column_batch(who, prec, n[0], n[1]));
      return fft_plan_make(plan, 2, n[0], n[1], false, prec, batch);
    }
    return fft_plan_make(plan, 1, n[0], 0, false, prec, batch);
  }
  if (rank != 1)
    return fail(WH_INVALID_VALUE, "%s: a rank-2 plan with an advanced layout (inembed set) is not supported", who);
  if (ostride != istride || odist != idist)
    return fail(WH_INVALID_VALUE, "%s: output layout (%d, %d) differs from the input layout (%d, %d)", who, ostride,
                odist, istride, idist);
  // rank 1 with an advanced layout; inembed[0] (the outermost extent) is not used
  return wh_fft_plan_many(plan, n[0], istride, idist, prec, batch);
}

static int fft_exec(wh_fft_plan* plan, void* idata, void* odata, int dir, int norm, void* stream) {
  if (!plan) return fail(WH_INVALID_PLAN, "wh_fft_exec: plan is NULL");
  if (!idata || idata != odata)
    return fail(WH_INVALID_VALUE, "wh_fft_exec: only in-place transforms (idata == odata != NULL)");
  if (dir != WH_FORWARD && dir != WH_INVERSE) return fail(WH_INVALID_VALUE, "wh_fft_exec: direction %d", dir);
  const size_t align = plan->type == WH_C128 ? 16 : 8;
  if (reinterpret_cast<uintptr_t>(idata) % align)
    return fail(WH_UNALIGNED_DATA, "wh_fft_exec: data not %zu-byte aligned", align);
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const auto hi = [&](const TwSet& t) { return plan->type == WH_C128 ? t.hi64.p : t.hi.p; };
  if (plan->rank == 1) {
    WH_HIP(wh::launch_fft_axis(plan->type, plan->columns, plan->n, dir, idata, plan->batch, norm, hi(plan->tw),
                               plan->tw.lo.p, s),
           "wh_fft_exec launch");
    return WH_SUCCESS;
  }
  // rank 2, batch contiguous [n][n2] arrays: columns (length n), then rows (length n2);
  // each pass folds in its own 1/length when normalising
  const size_t bytes = size_t(plan->n) * size_t(plan->n2) * (plan->type == WH_C128 ? 16 : 8);
  for (int b = 0; b < plan->batch; ++b) {
    void* d = static_cast<char*>(idata) + size_t(b) * bytes;
    WH_HIP(wh::launch_fft_axis(plan->type, 1, plan->n, dir, d, plan->n2, norm, hi(plan->tw), plan->tw.lo.p, s),
           "wh_fft_exec launch (columns)");
    WH_HIP(wh::launch_fft_axis(plan->type, 0, plan->n2, dir, d, plan->n, norm, hi(plan->tw2), plan->tw2.lo.p, s),
           "wh_fft_exec launch (rows)");
  }
  return WH_SUCCESS;
}

extern "C" int wh_fft_exec(wh_fft_plan* plan, void* idata, void* odata, int direction, void* stream) {
  return fft_exec(plan, idata, odata, direction, 0, stream);
}
extern "C" int wh_fft_exec_normalized(wh_fft_plan* plan, void* idata, void* odata, int direction, void* stream) {
  return fft_exec(plan, idata, odata, direction, 1, stream);
}
extern "C" int wh_fft_destroy(wh_fft_plan* plan) {
  if (!plan) return fail(WH_INVALID_PLAN, "wh_fft_destroy: plan is NULL");
  delete plan;
  return WH_SUCCESS;
}

// ============================================================================ ensembles
namespace {

struct MemberBufs {
  DevBuf pot, theta_decay, lambda_decay, x_mask, sigma_p, theta_host, lambda_host;
  DevBuf x_host, p_host, dv_host, v0_host, colsums, tilepart, scale, rho, G, colpart, pot_dv;
  DevBuf odm_g, odm_x, odm_th, odm_s, pot_expr;
};

struct WorkList {
  DevBuf buf;
  int n = 0;
  int upload(const std::vector<unsigned>& items) {
    n = int(items.size());
    return buf.upload(items.data(), items.size() * sizeof(unsigned));
  }
};

struct GraphEntry {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

}  // namespace

struct wh_ensemble {
  wh_ensemble_desc d{};
  int n = 0;
  std::vector<wh::Member> mem;  // host mirror of the device table
  std::vector<std::unique_ptr<MemberBufs>> bufs;
  DevBuf mem_dev;
  std::map<int, std::unique_ptr<TwSet>> tw;  // by size
  int xvar = 0, pvar = 0, pclvar = 0;
  bool expr = false;  // a member's V is an expression program: tables before each p sweep
  // p work lists: one for the plain p variant and one for the Caldeira-Leggett p variant
  // (a tile spans block * R / P columns, which differs between the two variant tables;
  // a CL plan also runs plain p sweeps: the merged Strang first half step, Strang + GP
  // V/2, the unfused single V step)
  WorkList work_x, work_p, work_pcl, work_colsum, work_scale, work_gp, work_adj, work_odm_x, work_odm_p;
  wh::SweepVariant built_x{}, built_p{}, built_pcl{};  // the variants each list was cut for
  DevBuf step, done, gcols;
  hipStream_t capture_stream = nullptr;
  std::map<std::tuple<hipStream_t, int>, GraphEntry> graphs;  // (stream, chunk)
  int graph_chunk = 16;
  bool table_dirty = true;
  ~wh_ensemble() { drop_graphs(); if (capture_stream) (void)hipStreamDestroy(capture_stream); }
  void drop_graphs() {
    for (auto& kv : graphs) {
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
      if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    graphs.clear();
  }
  bool gp() const { return d.gp_coeff != 0.0; }
  bool cl() const { return d.damping == 1; }
  bool odm() const { return d.damping == 2; }
};

struct wh_wigner2d {
  wh_ensemble* e = nullptr;
  ~wh_wigner2d() { delete e; }
};

namespace {

// An expression program: length, op codes and the stack depth it reaches (the device
// evaluator trusts all three).
// peak (optional): the deepest stack the program reaches.
int check_expr(const double* code, int n, int member, int* peak = nullptr) {
  if (!code || n < 1 || n > WH_EXPR_MAX)
    return fail(WH_INVALID_DESCRIPTOR, "member %d: expression program %p of %d ops (1..%d)", member,
                (const void*)code, n, WH_EXPR_MAX);
  int depth = 0, top = 0;
  for (int i = 0; i < n; ++i) {
    const double opd = code[2 * i];
    // range first (NaN fails it), then integrality
    if (!(opd >= WH_EXPR_CONST && opd <= WH_EXPR_HEAVISIDE) || opd != double(int(opd)))
      return fail(WH_INVALID_DESCRIPTOR, "member %d: expression op %g at %d", member, opd, i);
    const int op = int(opd);
    if (op <= WH_EXPR_T) ++depth;
    else if (op >= WH_EXPR_ADD && op <= WH_EXPR_POW) --depth;
    if (depth < 1 || depth > WH_EXPR_STACK)
      return fail(WH_INVALID_DESCRIPTOR, "member %d: expression stack depth %d at op %d", member, depth, i);
    top = depth > top ? depth : top;
  }
  if (depth != 1) return fail(WH_INVALID_DESCRIPTOR, "member %d: expression leaves %d values", member, depth);
  if (peak) *peak = top;
  return WH_SUCCESS;
}

int validate_member(const wh_ensemble_desc& d, const wh_member_desc& m, int i) {
  if (m.struct_size != (int)sizeof(wh_member_desc))
    return fail(WH_INVALID_DESCRIPTOR, "member %d: struct_size %d != %d", i, m.struct_size, (int)sizeof(wh_member_desc));
  if (!sweep_size_supported(m.x_dim) || !sweep_size_supported(m.p_dim))
    return fail(WH_INVALID_SIZE, "member %d: grid %d x %d: both sides must be powers of two in [128, 2048]", i,
                m.p_dim, m.x_dim);
  if (m.pot_deg == WH_POT_TABLE) {
    if (!m.pot_dv) return fail(WH_INVALID_DESCRIPTOR, "member %d: pot_deg WH_POT_TABLE needs pot_dv", i);
  } else if (m.pot_deg == WH_POT_EXPR) {
    WH_TRY(check_expr(m.pot_expr, m.pot_expr_len, i));
  } else if (m.pot_deg < 0 || m.pot_deg > WH_MAX_DEG || m.pot_rows < 1 || !m.pot_coef) {
    return fail(WH_INVALID_DESCRIPTOR, "member %d: potential deg %d (max %d), rows %d, coef %p", i, m.pot_deg,
                WH_MAX_DEG, m.pot_rows, (const void*)m.pot_coef);
  }
  if (d.damping == 1 && (!m.sigma_p || !m.theta_host))
    return fail(WH_INVALID_DESCRIPTOR, "member %d: Caldeira-Leggett needs sigma_p and theta_host", i);
  if (d.damping == 2 && (!m.odm_g || !m.odm_x || !m.odm_th))
    return fail(WH_INVALID_DESCRIPTOR, "member %d: ODM needs odm_g, odm_x and odm_th", i);
  if (d.gp_coeff != 0.0 && (!m.lambda_host || !m.theta_host))
    return fail(WH_INVALID_DESCRIPTOR, "member %d: Gross-Pitaevskii needs lambda_host and theta_host", i);
  if (!m.x_host || !m.p_host) return fail(WH_INVALID_DESCRIPTOR, "member %d: observables need x_host and p_host", i);
  return WH_SUCCESS;
}

int validate_common(const wh_ensemble_desc& d) {
  if (d.struct_size != (int)sizeof(wh_ensemble_desc))
    return fail(WH_INVALID_DESCRIPTOR, "ensemble struct_size %d != %d", d.struct_size, (int)sizeof(wh_ensemble_desc));
  if (d.precision != WH_C128 && d.precision != WH_C64 && d.precision != WH_C64_WIDE)
    return fail(WH_INVALID_TYPE, "precision %d is not WH_C128, WH_C64 or WH_C64_WIDE", d.precision);
  if (d.scheme != WH_SCHEME_LIE && d.scheme != WH_SCHEME_STRANG)
    return fail(WH_INVALID_DESCRIPTOR, "scheme %d", d.scheme);
  if (d.kinetic_poly && (d.kinetic_deg < 0 || d.kinetic_deg > WH_MAX_DEG))
    return fail(WH_INVALID_DESCRIPTOR, "kinetic degree %d", d.kinetic_deg);
  if (!(d.dt == d.dt) || d.hbar == 0.0 || d.mass == 0.0)
    return fail(WH_INVALID_DESCRIPTOR, "dt/hbar/mass: %g %g %g", d.dt, d.hbar, d.mass);
  if (d.damping < 0 || d.damping > 2)
    return fail(WH_INVALID_DESCRIPTOR, "damping %d (0 none, 1 Caldeira-Leggett, 2 ODM)", d.damping);
  if (d.damping == 2 && (d.odm_channels < 0 || d.odm_channels > 2))
    return fail(WH_INVALID_DESCRIPTOR, "odm_channels %d (0 or 1: one channel, 2: two)", d.odm_channels);
  return WH_SUCCESS;
}

// variant v fits every member on the given axis (x sweep: rows of X; p sweep: columns of P;
// cl: a Caldeira-Leggett p variant)
bool fits_all(const wh_ensemble* e, const wh::SweepVariant& v, bool x_axis, bool cl = false) {
  for (const auto& m : e->mem) {
    const int N = x_axis ? m.X : m.P;
    const int other = x_axis ? m.P : m.X;
    if (!wh::variant_fits(N, v, cl)) return false;
    if (other % wh::tile_lines(N, v) != 0) return false;
  }
  return true;
}

// p tiles of one variant, grouped by grid side (ascending), members in order within a size;
// consecutive tiles of a member go to one XCD (blocks are dealt round-robin to the 8 XCDs)
void p_tiles(const wh_ensemble* e, const wh::SweepVariant& pv, std::vector<unsigned>& wp) {
  for (int N = 1; N <= (1 << 20); N <<= 1) {
    for (int i = 0; i < e->n; ++i) {
      const auto& m = e->mem[i];
      if (m.P != N) continue;
      const unsigned hi = unsigned(i) << 16;
      const int cols = wh::tile_lines(m.P, pv);
      const int nt = m.X / cols;
      for (int q = 0; q < nt; ++q) {
        const int t = (nt % 8 == 0) ? (q % 8) * (nt / 8) + q / 8 : q;
        wp.push_back(hi | unsigned(t));
      }
    }
  }
}

int build_work_lists(wh_ensemble* e) {
  const auto& xv = wh::kXVariants[e->xvar];
  std::vector<unsigned> wx, wp, wpcl, wc, ws, wg, wox, wop;
  // x items grouped by grid side (ascending), members in order within a size
  for (int N = 1; N <= (1 << 20); N <<= 1) {
    for (int i = 0; i < e->n; ++i) {
      const auto& m = e->mem[i];
      if (m.X != N) continue;
      const unsigned hi = unsigned(i) << 16;
      const int rows = wh::tile_lines(m.X, xv);
      for (int t = 0; t < m.P / rows; ++t) wx.push_back(hi | unsigned(t));
    }
  }
  p_tiles(e, wh::kPVariants[e->pvar], wp);
  if (e->cl()) p_tiles(e, wh::kPCLVariants[e->pclvar], wpcl);
  for (int i = 0; i < e->n; ++i) {
    const auto& m = e->mem[i];
    const unsigned hi = unsigned(i) << 16;
    // column-sum tiles: with a multiple of 8 tiles, block q (on XCD q % 8) takes tile
    // (q % 8) * (nt / 8) + q / 8, so that neighbouring tiles (halves of one 128-byte line)
    // are read on one XCD
    const int nc = (m.X + wh::kColsumCols - 1) / wh::kColsumCols;
    for (int q = 0; q < nc; ++q) {
      const int t = (nc % 8 == 0) ? (q % 8) * (nc / 8) + q / 8 : q;
      wc.push_back(hi | unsigned(t));
    }
    const size_t cells = size_t(m.X) * size_t(m.P);
    for (size_t t = 0; t < (cells + wh::kScaleChunk - 1) / wh::kScaleChunk; ++t) ws.push_back(hi | unsigned(t));
    if (e->gp()) {
      const int grows = wh::tile_lines(m.X, wh::SweepVariant{16, 256, 0});
      for (int t = 0; t < m.P / grows; ++t) wg.push_back(hi | unsigned(t));
    }
    if (e->odm()) {
      const int orows = wh::odm_rows_per_tile(m.X), ocols = wh::odm_cols_per_tile(m.P);
      if (orows < 1 || ocols < 1 || m.P % orows || m.X % ocols) return fail(WH_INVALID_SIZE, "ODM tiles");
      for (int t = 0; t < m.P / orows; ++t) wox.push_back(hi | unsigned(t));
      for (int t = 0; t < m.X / ocols; ++t) wop.push_back(hi | unsigned(t));
    }
  }
  if (wx.size() > 0x7fffffff || ws.size() > 0x7fffffff) return fail(WH_INVALID_SIZE, "work list too long");
  for (const auto& v : {&wx, &wp, &wpcl, &wc, &ws, &wg})
    for (unsigned w : *v)
      if ((w & 0xffffu) == 0xffffu) return fail(WH_INVALID_SIZE, "a member has more than 65535 tiles");
  WH_TRY(e->work_x.upload(wx));
  WH_TRY(e->work_p.upload(wp));
  if (e->cl()) WH_TRY(e->work_pcl.upload(wpcl));
  e->built_x = xv;
  e->built_p = wh::kPVariants[e->pvar];
  e->built_pcl = wh::kPCLVariants[e->pclvar];
  WH_TRY(e->work_colsum.upload(wc));
  WH_TRY(e->work_scale.upload(ws));
  if (e->gp()) WH_TRY(e->work_gp.upload(wg));
  if (e->odm()) {
    WH_TRY(e->work_odm_x.upload(wox));
    WH_TRY(e->work_odm_p.upload(wop));
  }
  e->drop_graphs();
  return WH_SUCCESS;
}

int ensemble_create(const wh_ensemble_desc* desc, const wh_member_desc* members, int n, wh_ensemble** out) {
  if (!out) return fail(WH_INVALID_VALUE, "out is NULL");
  *out = nullptr;
  if (!desc || !members || n < 1 || n > 65535)
    return fail(WH_INVALID_VALUE, "desc %p members %p n %d", (const void*)desc, (const void*)members, n);
  WH_TRY(validate_common(*desc));
  for (int i = 0; i < n; ++i) WH_TRY(validate_member(*desc, members[i], i));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(WH_NO_DEVICE, "no HIP device");
  auto e = std::make_unique<wh_ensemble>();
  e->d = *desc;
  e->n = n;
  e->mem.resize(n);
  const int prec = desc->precision;
  const bool gp = desc->gp_coeff != 0.0;
  for (int i = 0; i < n; ++i) {
    const auto& md = members[i];
    auto b = std::make_unique<MemberBufs>();
    const size_t X = size_t(md.x_dim), P = size_t(md.p_dim);
    std::vector<double> zeros(std::max(X, P), 0.0);
    for (int N : {md.x_dim, md.p_dim})
      if (!e->tw.count(N)) e->tw[N] = std::make_unique<TwSet>();
    WH_TRY(make_twiddles(*e->tw[md.x_dim], md.x_dim, prec, gp));
    WH_TRY(make_twiddles(*e->tw[md.p_dim], md.p_dim, prec, false));
    if (md.pot_deg == WH_POT_TABLE || md.pot_deg == WH_POT_EXPR) {
      const double zero_row[WH_MAX_DEG + 1] = {};
      WH_TRY(b->pot.upload(zero_row, sizeof(zero_row)));
      if (md.pot_deg == WH_POT_TABLE) {
        WH_TRY(b->pot_dv.upload(md.pot_dv, X * P * sizeof(double)));
      } else {
        // two [P][X] tables, filled on the device before each p sweep (time index, next index)
        WH_TRY(b->pot_expr.upload(md.pot_expr, size_t(md.pot_expr_len) * 2 * sizeof(double)));
        WH_TRY(b->pot_dv.alloc(2 * X * P * sizeof(double)));
      }
    } else {
      WH_TRY(b->pot.upload(md.pot_coef, size_t(md.pot_rows) * (WH_MAX_DEG + 1) * sizeof(double)));
    }
    WH_TRY(upload_opt(b->theta_decay, md.theta_decay, P));
    WH_TRY(upload_opt(b->lambda_decay, md.lambda_decay, X));
    WH_TRY(upload_opt(b->x_mask, md.x_mask, X));
    WH_TRY(upload_opt(b->sigma_p, md.sigma_p, P));
    WH_TRY(upload_opt(b->theta_host, md.theta_host, P));
    WH_TRY(upload_opt(b->lambda_host, md.lambda_host, X));
    WH_TRY(b->x_host.upload(md.x_host, X * sizeof(double)));
    WH_TRY(b->p_host.upload(md.p_host, P * sizeof(double)));
    WH_TRY(b->dv_host.upload(md.dv_host ? md.dv_host : zeros.data(), X * sizeof(double)));
    WH_TRY(b->v0_host.upload(md.v0_host ? md.v0_host : zeros.data(), X * sizeof(double)));
    WH_TRY(b->colsums.alloc(X * wh::kNCol * sizeof(double)));
    WH_TRY(b->tilepart.alloc(X * wh::kTileObs * sizeof(double)));  // at most X tiles (one column each)
    const double one[2] = {1.0, 0.0};
    WH_TRY(b->scale.upload(one, sizeof(one)));
    if (gp) {
      WH_TRY(b->rho.alloc(X * sizeof(double)));
      WH_TRY(b->G.alloc(X * P * sizeof(double)));
    }
    if (desc->damping == 2) {
      WH_TRY(b->odm_g.upload(md.odm_g, 4 * X * P * sizeof(double)));
      WH_TRY(b->odm_x.upload(md.odm_x, 2 * X * sizeof(double)));
      WH_TRY(b->odm_th.upload(md.odm_th, 2 * P * sizeof(double)));
      WH_TRY(b->odm_s.alloc(2 * X * P * (prec == WH_C64 ? 8 : 16)));  // arithmetic precision
    }
    wh::Member& m = e->mem[i];
    std::memset(&m, 0, sizeof(m));
    m.colsums = b->colsums.as<double>();
    m.tilepart = b->tilepart.as<double>();
    m.scale = b->scale.as<double>();
    m.rho = b->rho.as<double>();
    m.G = b->G.as<double>();
    m.pot = b->pot.as<double>();
    m.pot_dv = b->pot_dv.as<double>();
    m.theta_decay = b->theta_decay.as<double>();
    m.lambda_decay = b->lambda_decay.as<double>();
    m.xmask = b->x_mask.as<double>();
    m.sigma_p = b->sigma_p.as<double>();
    m.theta_host = b->theta_host.as<double>();
    m.lambda_host = b->lambda_host.as<double>();
    m.x_host = b->x_host.as<double>();
    m.p_host = b->p_host.as<double>();
    m.dv_host = b->dv_host.as<double>();
    m.v0_host = b->v0_host.as<double>();
    const TwSet& tx = *e->tw[md.x_dim];
    const TwSet& tp = *e->tw[md.p_dim];
    m.twX_hi = prec != WH_C64 ? tx.hi64.p : tx.hi.p;
    m.twX_lo = tx.lo.p;
    m.twP_hi = prec != WH_C64 ? tp.hi64.p : tp.hi.p;
    m.twP_lo = tp.lo.p;
    m.twX64 = tx.hi64.p;
    m.X = md.x_dim;
    m.P = md.p_dim;
    m.pot_rows = md.pot_deg < 0 ? 1 : md.pot_rows;
    m.pot_expr = b->pot_expr.as<double>();
    m.pot_expr_len = md.pot_deg == WH_POT_EXPR ? md.pot_expr_len : 0;
    m.pot_expr_depth = 0;
    m.pot_expr_kvn = md.pot_deg == WH_POT_EXPR && md.pot_expr_kvn ? 1 : 0;
    if (md.pot_deg == WH_POT_EXPR) WH_TRY(check_expr(md.pot_expr, md.pot_expr_len, i, &m.pot_expr_depth));
    m.pot_t_step = md.pot_t_step;
    m.pot_deg = md.pot_deg == WH_POT_EXPR ? WH_POT_TABLE : md.pot_deg;  // the sweeps read its tables
    m.dx = md.dx;
    m.dp = md.dp;
    m.dtheta = md.dtheta;
    m.dlambda = md.dlambda;
    m.dxdp = md.dxdp;
    m.inv_2m = md.inv_2m;
    m.dx_host = md.dx_host;
    m.rho_factor = md.gp_rho_factor;
    m.odm_g = b->odm_g.as<double>();
    m.odm_x = b->odm_x.as<double>();
    m.odm_th = b->odm_th.as<double>();
    m.odm_s = b->odm_s.p;
    m.odm_c = desc->gamma_dt * 2.0;
    m.odm_sign = md.odm_sign;
    e->bufs.push_back(std::move(b));
  }
  // default variants: the first that fits every member
  auto pick = [&](const wh::SweepVariant* t, int nt, bool x_axis, int& out_v, bool cl = false) -> int {
    for (int v = 0; v < nt; ++v)
      if (fits_all(e.get(), t[v], x_axis, cl)) {
        out_v = v;
        return WH_SUCCESS;
      }
    return fail(WH_INVALID_SIZE, "no kernel variant fits every member's grid");
  };
  WH_TRY(pick(wh::kXVariants, wh::kNumX, true, e->xvar));
  WH_TRY(pick(wh::kPVariants, wh::kNumP, false, e->pvar));
  WH_TRY(pick(wh::kPCLVariants, wh::kNumPCL, false, e->pclvar, true));
  for (int i = 0; i < n; ++i) e->expr = e->expr || e->mem[i].pot_expr_len > 0;
  if (gp) {
    const wh::SweepVariant g{16, 256, 0};
    for (int i = 0; i < n; ++i) {
      const auto& m = e->mem[i];
      if (!wh::variant_fits(m.X, g) || m.P % wh::tile_lines(m.X, g) != 0)
        return fail(WH_INVALID_SIZE, "member %d: grid does not fit the Gross-Pitaevskii shift kernel", i);
      const size_t parts = size_t(m.P / wh::tile_lines(m.X, wh::kXVariants[e->xvar])) * size_t(m.X);
      WH_TRY(e->bufs[i]->colpart.alloc(parts * sizeof(double)));
      e->mem[i].colpart = e->bufs[i]->colpart.as<double>();
    }
  }
  WH_TRY(e->mem_dev.alloc(sizeof(wh::Member) * size_t(n)));
  WH_TRY(e->step.alloc(16));
  const int zero[4] = {0, 0, 0, 0};
  WH_TRY(e->done.upload(zero, sizeof(zero)));

  WH_TRY(build_work_lists(e.get()));
  *out = e.release();
  return WH_SUCCESS;
}

// --------------------------------------------------------------- launch helpers
wh::SweepArgs sweep_args(const wh_ensemble* e, int flags, double frac, int step_add) {
  const auto& d = e->d;
  wh::SweepArgs a{};
  a.members = e->mem_dev.as<wh::Member>();
  a.step = e->step.as<int>();
  a.step_add = step_add;
  a.flags = flags;
  a.dt = d.dt;
  a.hbar = d.hbar;
  a.vphase = frac * d.dt / d.hbar;
  a.hbar_half = 0.5 * d.hbar;
  a.gp_phase = -frac * d.dt * d.gp_coeff / d.hbar;
  a.cl_a = d.gamma_dt;
  a.inv_hbar = 1.0 / d.hbar;
  a.kdeg = d.kinetic_deg;
  for (int k = 0; k <= wh::kMaxDeg; ++k) a.kcoef[k] = d.kin_coef[k];
  a.rec_step = e->step.as<int>() + 1;
  const wh::SweepVariant pv = e->cl() ? wh::kPCLVariants[e->pclvar] : wh::kPVariants[e->pvar];
  a.pt_R = pv.R;
  a.pt_block = pv.block;
  a.fin_renorm = d.renormalize;
  a.fin_record = 1;
  if (d.scheme == WH_SCHEME_LIE) {
    // expPLambda: phase = dt*lambda*p/mass (GPU_Wigner2D_FFT.py:153)
    a.kin_linear = 1;
    a.k_c0 = 0.0;
    a.k_c1 = 1.0 / d.mass;
  } else if (d.kinetic_poly && d.kinetic_deg <= 2) {
    // (K(p + h lambda/2) - K(p - h lambda/2))/h = lambda (k1 + 2 k2 p) for K = k0 + k1 p + k2 p^2
    a.kin_linear = 1;
    a.k_c0 = d.kinetic_deg >= 1 ? d.kin_coef[1] : 0.0;
    a.k_c1 = d.kinetic_deg >= 2 ? 2.0 * d.kin_coef[2] : 0.0;
  } else {
    a.kin_linear = 0;
  }
  return a;
}

bool same_variant(const wh::SweepVariant& a, const wh::SweepVariant& b) {
  return a.R == b.R && a.block == b.block && a.tp == b.tp;
}

int xsweep(wh_ensemble* e, int flags, hipStream_t s) {
  // the work list's tiles are cut for the variant launched over it
  if (!same_variant(wh::kXVariants[e->xvar], e->built_x))
    return fail(WH_INTERNAL_ERROR, "x work list cut for another variant");
  wh::SweepArgs a = sweep_args(e, flags, 1.0, 0);
  a.work = e->work_x.buf.as<unsigned>();
  WH_LAUNCH(e->d.precision == WH_C128      ? wh::launch_xsweep_f64(e->xvar, a, e->work_x.n, s)
            : e->d.precision == WH_C64_WIDE ? wh::launch_xsweep_f32w(e->xvar, a, e->work_x.n, s)
                                      : wh::launch_xsweep_f32(e->xvar, a, e->work_x.n, s),
            "xsweep");
  return WH_SUCCESS;
}

int psweep(wh_ensemble* e, int flags, double frac, int step_add, hipStream_t s) {
  wh::SweepArgs a = sweep_args(e, flags, frac, step_add);
  const bool cl = (flags & wh::kCL) != 0;
  if (cl && !e->cl()) return fail(WH_INTERNAL_ERROR, "Caldeira-Leggett p sweep on a plan without friction");
  // the plain and the Caldeira-Leggett p variants have their own work lists (their tiles
  // may span different column counts)
  const int v = cl ? e->pclvar : e->pvar;
  const wh::SweepVariant& launched = cl ? wh::kPCLVariants[v] : wh::kPVariants[v];
  const WorkList& wl = cl ? e->work_pcl : e->work_p;
  if (!same_variant(launched, cl ? e->built_pcl : e->built_p))
    return fail(WH_INTERNAL_ERROR, "p work list cut for another variant");
  a.work = wl.buf.as<unsigned>();
  if (e->expr) {
    // expression-program members: their [P][X] tables at the time index this sweep reads
    // (and the next one for the merged Strang half step)
    const bool arg = (flags & wh::kStepArg) != 0;
    for (int i = 0; i < e->n; ++i) {
      const wh::Member& m = e->mem[i];
      if (m.pot_expr_len <= 0) continue;
      WH_LAUNCH(wh::launch_expr_table(e->mem_dev.as<wh::Member>(), i, m.X * m.P, m.pot_expr_depth,
                                      e->step.as<int>(), step_add, arg ? 1 : 0, (flags & wh::kMerge) ? 2 : 1,
                                      a.hbar_half, s),
                "expr_table");
    }
  }
  WH_LAUNCH(e->d.precision == WH_C128      ? wh::launch_psweep_f64(v, cl, a, wl.n, s)
            : e->d.precision == WH_C64_WIDE ? wh::launch_psweep_f32w(v, cl, a, wl.n, s)
                                      : wh::launch_psweep_f32(v, cl, a, wl.n, s),
            "psweep");
  return WH_SUCCESS;
}

// step_arg >= -1 with use_step_arg: the time index of the record row is step_arg (the
// device counter is not read); next_step >= 0: the counter is set to it (needs inc)
int reduce(wh_ensemble* e, int renorm, int record, int rho, int update_scale, bool inc, hipStream_t s,
           bool use_step_arg = false, int step_arg = 0, int next_step = -1) {
  wh::ReduceArgs r{};
  r.use_step_arg = use_step_arg ? 1 : 0;
  r.step_arg = step_arg;
  r.next_step = next_step;
  r.members = e->mem_dev.as<wh::Member>();
  r.renorm = renorm;
  r.record = record;
  r.rho = rho;
  r.update_scale = update_scale;
  r.step = e->step.as<int>();
  r.step_inc = inc ? e->step.as<int>() : nullptr;
  WH_LAUNCH(wh::launch_reduce(r, e->n, e->done.as<int>(), s), "reduce");
  return WH_SUCCESS;
}

int gpshift(wh_ensemble* e, hipStream_t s) {
  WH_LAUNCH(wh::launch_gpshift(e->mem_dev.as<wh::Member>(), e->work_gp.buf.as<unsigned>(), e->work_gp.n, s),
            "gpshift");
  return WH_SUCCESS;
}

// One iteration after the first observable row (all members).
//  Lie    (GPU_Wigner2D_FFT.py:1843-1913): [GP terms] K (pending factor folded into the
//         load), V [+GP] [+CL] + column sums, reduce (record row, factor, step += 1).
//  Strang (GP.py:1898-1999): without GP the second half step of iteration k and the
//         first of k+1 share one p sweep, the observables taken between them; with GP
//         V/2, K + partial column sums, density, V/2 + column sums.
//  Lie and Strang without GP or ODM: no reduce launch; each p sweep writes tile partials
//  of the observables and the next x sweep turns them into the record row and the
//  factor (kTileScale).  `first`: the pending factor is the one of the initial reduce.
//  After the last iteration tiles_pending(e) asks for one closing reduce.
bool tiles_pending(const wh_ensemble* e) { return !e->gp() && !e->odm(); }
int renorm_flag(const wh_ensemble* e) { return e->d.renormalize; }

int iteration(wh_ensemble* e, bool merge, bool first, hipStream_t s) {
  const auto& d = e->d;
  const int cl = e->cl() ? wh::kCL : 0;
  const int renorm = d.renormalize;
  if (d.scheme == WH_SCHEME_LIE && e->odm()) {
    // K; V; W += L_ODM(W) (three sweeps); column sums of the result; renormalise
    if (e->gp()) WH_TRY(gpshift(e, s));
    WH_TRY(xsweep(e, wh::kScaleOnLoad, s));
    WH_TRY(psweep(e, e->gp() ? wh::kGP : 0, 1.0, 0, s));
    // one Lindblad channel (GPU_Wigner2D_FFT, weight 1, sign 1) or two applied in turn
    // (GPU_Wigner2D_ODM2: weightLindbland1 with sign 1, then weightLindbland2 with sign -1)
    for (int ch = 0; ch < (d.odm_channels == 2 ? 2 : 1); ++ch)
      WH_LAUNCH(wh::launch_odm(d.precision, e->mem_dev.as<wh::Member>(), e->work_odm_x.buf.as<unsigned>(),
                               e->work_odm_x.n, e->work_odm_p.buf.as<unsigned>(), e->work_odm_p.n,
                               ch == 0 ? d.odm_weight1 : d.odm_weight2, ch == 0 ? 1.0 : -1.0, s),
                "odm");
    WH_LAUNCH(wh::launch_colsum(storage_code(d.precision), e->mem_dev.as<wh::Member>(),
                                e->work_colsum.buf.as<unsigned>(),
                                e->work_colsum.n, s),
              "colsum");
    WH_TRY(reduce(e, renorm, 1, e->gp(), 1, true, s));
  } else if (d.scheme == WH_SCHEME_STRANG && e->odm()) {
    // V/2; K; V/2; W += L_ODM(W); column sums; renormalise (GPU_Wigner2D_GPitaevskii.Run,
    // GP.py:1942-1968): the dissipator sits between the two half steps of consecutive
    // iterations, so they are not merged
    if (e->gp()) WH_TRY(gpshift(e, s));
    WH_TRY(psweep(e, wh::kScaleOnLoad | (e->gp() ? wh::kGP : 0), 0.5, 0, s));
    if (e->gp()) {
      WH_TRY(xsweep(e, wh::kColPart, s));
      WH_LAUNCH(wh::launch_colpart_reduce(e->mem_dev.as<wh::Member>(), e->n, wh::kXVariants[e->xvar].R,
                                          wh::kXVariants[e->xvar].block, e->step.as<int>(), s),
                "colpart_reduce");
      WH_TRY(gpshift(e, s));
    } else {
      WH_TRY(xsweep(e, 0, s));
    }
    WH_TRY(psweep(e, e->gp() ? wh::kGP : 0, 0.5, 0, s));
    for (int ch = 0; ch < (d.odm_channels == 2 ? 2 : 1); ++ch)
      WH_LAUNCH(wh::launch_odm(d.precision, e->mem_dev.as<wh::Member>(), e->work_odm_x.buf.as<unsigned>(),
                               e->work_odm_x.n, e->work_odm_p.buf.as<unsigned>(), e->work_odm_p.n,
                               ch == 0 ? d.odm_weight1 : d.odm_weight2, ch == 0 ? 1.0 : -1.0, s),
                "odm");
    WH_LAUNCH(wh::launch_colsum(storage_code(d.precision), e->mem_dev.as<wh::Member>(),
                                e->work_colsum.buf.as<unsigned>(), e->work_colsum.n, s),
              "colsum");
    WH_TRY(reduce(e, renorm, 1, e->gp(), 1, true, s));
  } else if (d.scheme == WH_SCHEME_LIE && e->gp()) {
    WH_TRY(gpshift(e, s));
    WH_TRY(xsweep(e, wh::kScaleOnLoad, s));
    WH_TRY(psweep(e, wh::kColSums | cl | wh::kGP, 1.0, 0, s));
    WH_TRY(reduce(e, renorm, 1, 1, 1, true, s));
  } else if (d.scheme == WH_SCHEME_LIE) {
    WH_TRY(xsweep(e, first ? wh::kScaleOnLoad : wh::kTileScale, s));
    WH_TRY(psweep(e, wh::kColSums | wh::kTilePart | cl, 1.0, 0, s));
  } else if (!e->gp()) {
    WH_TRY(xsweep(e, first ? wh::kScaleOnLoad : wh::kTileScale, s));
    WH_TRY(psweep(e, wh::kColSums | wh::kTilePart | cl | (merge ? wh::kMerge : 0), 0.5, 0, s));
  } else {
    WH_TRY(gpshift(e, s));
    WH_TRY(psweep(e, wh::kScaleOnLoad | wh::kGP, 0.5, 0, s));
    WH_TRY(xsweep(e, wh::kColPart, s));
    WH_LAUNCH(wh::launch_colpart_reduce(e->mem_dev.as<wh::Member>(), e->n, wh::kXVariants[e->xvar].R,
                                        wh::kXVariants[e->xvar].block, e->step.as<int>(), s),
              "colpart_reduce");
    WH_TRY(gpshift(e, s));
    WH_TRY(psweep(e, wh::kGP | wh::kColSums | cl, 0.5, 0, s));
    WH_TRY(reduce(e, renorm, 1, 1, 1, true, s));
  }
  return WH_SUCCESS;
}

int iterations(wh_ensemble* e, int count, bool merge_last, bool first, hipStream_t s) {
  for (int k = 0; k < count; ++k) WH_TRY(iteration(e, (k < count - 1) || merge_last, first && k == 0, s));
  return WH_SUCCESS;
}

// Graph replay for runs with at least two chunks of middle iterations: a shorter run
// launches its iterations directly, since the host stays ahead of the GPU there and the
// first replay of a graph in a process costs ~60 us more than direct launches.
bool graph_pays(int chunk, int use_graph, int count) { return use_graph && count >= 2 * chunk; }

// The graph reads member pointers, the step counter and the record base from device
// memory, so one capture serves every later run of the plan.
// The plan's step graph for stream s (graph_chunk iterations): captured (capturing
// executes nothing), instantiated and uploaded on the first graph-enabled run whatever its
// length, so that later runs replay it without capture or upload work.
int ensure_graph(wh_ensemble* e, hipStream_t s, hipGraphExec_t* exec) {
  const int chunk = e->graph_chunk;
  auto key = std::make_tuple(s, chunk);
  auto it = e->graphs.find(key);
  if (it == e->graphs.end()) {
    if (!e->capture_stream)
      WH_HIP(hipStreamCreateWithFlags(&e->capture_stream, hipStreamNonBlocking), "hipStreamCreate");
    GraphEntry g;
    WH_HIP(hipStreamBeginCapture(e->capture_stream, hipStreamCaptureModeThreadLocal), "BeginCapture");
    const int rc = iterations(e, chunk, true, false, e->capture_stream);
    const hipError_t err = hipStreamEndCapture(e->capture_stream, &g.graph);
    if (rc) {
      if (g.graph) (void)hipGraphDestroy(g.graph);
      return rc;
    }
    if (err != hipSuccess) return hip_fail(err, "EndCapture");
    WH_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "GraphInstantiate");
    WH_HIP(hipGraphUpload(g.exec, s), "GraphUpload");
    it = e->graphs.emplace(key, g).first;
  }
  *exec = it->second.exec;
  return WH_SUCCESS;
}

// Replays `count` iterations from the step graph (whole chunks), the rest launched directly.
int run_graph(wh_ensemble* e, int count, hipStream_t s) {
  hipGraphExec_t exec = nullptr;
  WH_TRY(ensure_graph(e, s, &exec));
  while (count >= e->graph_chunk) {
    WH_HIP(hipGraphLaunch(exec, s), "GraphLaunch");
    count -= e->graph_chunk;
  }
  return iterations(e, count, true, false, s);
}

// Device member table: a structural change (creation, new potential rows, adjoint
// operands) re-uploads the whole table and waits for the copy; a change of only the
// states and record rows of a run is written by a bind kernel in stream order.
int sync_table(wh_ensemble* e, bool changed, hipStream_t s) {
  if (changed || e->table_dirty) {
    WH_HIP(hipMemcpyAsync(e->mem_dev.p, e->mem.data(), sizeof(wh::Member) * e->mem.size(), hipMemcpyHostToDevice, s),
           "member table upload");
    WH_HIP(hipStreamSynchronize(s), "member table upload");
    e->table_dirty = false;
  }
  return WH_SUCCESS;
}

int bind_states(wh_ensemble* e, hipStream_t s) {
  if (e->table_dirty) return sync_table(e, true, s);
  for (int first = 0; first < e->n; first += wh::kBindMax) {
    wh::BindArgs a{};
    a.members = e->mem_dev.as<wh::Member>();
    a.first = first;
    a.count = std::min(wh::kBindMax, e->n - first);
    a.obs_base = e->mem[first].obs_base;
    for (int i = 0; i < a.count; ++i) {
      a.W[i] = e->mem[first + i].W;
      a.obs[i] = e->mem[first + i].obs;
    }
    WH_LAUNCH(wh::launch_bind(a, s), "bind");
  }
  return WH_SUCCESS;
}

int ensemble_run(wh_ensemble* e, void* const* states, int first_step, int n_steps, double* obs,
                 int normalize_initial, int use_graph, hipStream_t s) {
  if (!states) return fail(WH_INVALID_VALUE, "states is NULL");
  if (n_steps < 0 || first_step < 0) return fail(WH_INVALID_VALUE, "n_steps %d first_step %d", n_steps, first_step);
  if (!obs) return fail(WH_INVALID_VALUE, "obs is NULL");
  const auto& d = e->d;
  const size_t align = state_bytes(d.precision);
  bool changed = false;
  for (int i = 0; i < e->n; ++i) {
    if (!states[i]) return fail(WH_INVALID_VALUE, "state %d is NULL", i);
    if (reinterpret_cast<uintptr_t>(states[i]) % align) return fail(WH_UNALIGNED_DATA, "state %d: %p", i, states[i]);
    wh::Member& m = e->mem[i];
    double* o = obs + size_t(i) * size_t(n_steps + 1) * wh::kNObs;
    // Lie rows are pre-step states: row k = state before iteration first_step + k,
    // written by the reduce that follows iteration first_step + k - 1
    const int base = d.scheme == WH_SCHEME_LIE ? first_step - 1 : first_step;
    if (m.W != states[i] || m.obs != o || m.obs_base != base) changed = true;
    m.W = states[i];
    m.obs = o;
    m.obs_base = base;
  }
  if (changed || e->table_dirty) WH_TRY(bind_states(e, s));
  const int prec = d.precision;
  const wh::Member* md = e->mem_dev.as<wh::Member>();
  // the state before the first iteration: norm, renormalisation factor, (Lie) row 0
  WH_LAUNCH(wh::launch_colsum(storage_code(prec), md, e->work_colsum.buf.as<unsigned>(), e->work_colsum.n, s),
            "colsum");
  // record row of time index first_step - 1 (Lie) or first_step; the counter then reads first_step.
  // (One launch for both, the member's last column-sum block doing its reduce, took 30 us at
  // 1024^2 instead of 13: its device-scope fences write back and invalidate the L2 per block.)
  WH_TRY(reduce(e, normalize_initial, d.scheme == WH_SCHEME_LIE, e->gp(), 1, true, s, true,
                d.scheme == WH_SCHEME_LIE ? first_step - 1 : first_step, first_step));
  const bool merged_strang = d.scheme == WH_SCHEME_STRANG && !e->gp() && !e->odm();
  if (n_steps > 0 && merged_strang) {
    // first half step of the first iteration, with the pending factor folded in
    WH_TRY(psweep(e, wh::kScaleOnLoad, 0.5, 0, s));
    WH_LAUNCH(wh::launch_set_scale(md, e->n, 1.0, 0.0, s), "set_scale");
  }
  if (use_graph) {
    hipGraphExec_t exec = nullptr;
    WH_TRY(ensure_graph(e, s, &exec));  // ready for a later, longer run
  }
  if (n_steps > 0) {
    // the first iteration takes the factor of the initial reduce; the graph replays the
    // uniform middle iterations
    WH_TRY(iterations(e, 1, merged_strang && n_steps > 1, true, s));
    if (merged_strang) {
      if (n_steps > 1) {
        WH_TRY(graph_pays(e->graph_chunk, use_graph, n_steps - 2) ? run_graph(e, n_steps - 2, s)
                                                                   : iterations(e, n_steps - 2, true, false, s));
        WH_TRY(iterations(e, 1, false, false, s));  // last iteration: no next half step
      }
    } else {
      WH_TRY(graph_pays(e->graph_chunk, use_graph, n_steps - 1) ? run_graph(e, n_steps - 1, s)
                                                                   : iterations(e, n_steps - 1, false, false, s));
    }
    if (tiles_pending(e)) WH_TRY(reduce(e, renorm_flag(e), 1, 0, 1, true, s));
  }
  if (d.renormalize || (n_steps == 0 && normalize_initial))
    WH_LAUNCH(wh::launch_scale(storage_code(prec), md, e->work_scale.buf.as<unsigned>(), e->work_scale.n, s),
              "scale");
  return WH_SUCCESS;
}

}  // namespace

extern "C" int wh_ensemble_create(const wh_ensemble_desc* desc, const wh_member_desc* members, int n_members,
                                  wh_ensemble** out) {
  return ensemble_create(desc, members, n_members, out);
}

extern "C" int wh_ensemble_destroy(wh_ensemble* e) {
  if (!e) return fail(WH_INVALID_PLAN, "wh_ensemble_destroy: NULL");
  delete e;
  return WH_SUCCESS;
}

extern "C" int wh_ensemble_run(wh_ensemble* e, void* const* states, int first_step, int n_steps, double* obs,
                               int normalize_initial, int use_graph, void* stream) {
  if (!e) return fail(WH_INVALID_PLAN, "wh_ensemble_run: NULL");
  return ensemble_run(e, states, first_step, n_steps, obs, normalize_initial, use_graph,
                      static_cast<hipStream_t>(stream));
}

extern "C" int wh_ensemble_set_variants(wh_ensemble* e, int xvar, int pvar, int pclvar) {
  if (!e) return fail(WH_INVALID_PLAN, "wh_ensemble_set_variants: NULL");
  if (xvar >= wh::kNumX || pvar >= wh::kNumP || pclvar >= wh::kNumPCL)
    return fail(WH_INVALID_VALUE, "variant index out of range (%d %d %d)", xvar, pvar, pclvar);
  if (xvar >= 0 && !fits_all(e, wh::kXVariants[xvar], true))
    return fail(WH_INVALID_VALUE, "x variant %d does not fit every member", xvar);
  if (pvar >= 0 && !fits_all(e, wh::kPVariants[pvar], false))
    return fail(WH_INVALID_VALUE, "p variant %d does not fit every member", pvar);
  if (pclvar >= 0 && !fits_all(e, wh::kPCLVariants[pclvar], false, true))
    return fail(WH_INVALID_VALUE, "p+CL variant %d does not fit every member", pclvar);
  if (e->gp() && xvar >= 0 && xvar != e->xvar)
    return fail(WH_INVALID_VALUE, "the x variant of a Gross-Pitaevskii plan is fixed at creation");
  if (xvar >= 0) e->xvar = xvar;
  if (pvar >= 0) e->pvar = pvar;
  if (pclvar >= 0) e->pclvar = pclvar;
  return build_work_lists(e);
}

// ============================================================================ single grid
extern "C" int wh_wigner2d_create(const wh_wigner2d_desc* d, wh_wigner2d** out) {
  if (!out) return fail(WH_INVALID_VALUE, "wh_wigner2d_create: out is NULL");
  *out = nullptr;
  if (!d) return fail(WH_INVALID_VALUE, "descriptor is NULL");
  if (d->struct_size != (int)sizeof(wh_wigner2d_desc))
    return fail(WH_INVALID_DESCRIPTOR, "descriptor struct_size %d != %d", d->struct_size, (int)sizeof(wh_wigner2d_desc));
  wh_ensemble_desc c{};
  c.struct_size = sizeof(c);
  c.precision = d->precision;
  c.scheme = d->scheme;
  c.dt = d->dt;
  c.hbar = d->hbar;
  c.mass = d->mass;
  c.kinetic_poly = d->kinetic_poly;
  c.kinetic_deg = d->kinetic_deg;
  std::memcpy(c.kin_coef, d->kin_coef, sizeof(c.kin_coef));
  c.damping = d->damping;
  c.gamma_dt = d->gamma_dt;
  c.renormalize = d->renormalize;
  c.gp_coeff = d->gp_coeff;
  c.odm_channels = d->odm_channels;
  c.odm_weight1 = d->odm_weight1;
  c.odm_weight2 = d->odm_weight2;
  wh_member_desc m{};
  m.struct_size = sizeof(m);
  m.x_dim = d->x_dim;
  m.p_dim = d->p_dim;
  m.dx = d->dx;
  m.dp = d->dp;
  m.dtheta = d->dtheta;
  m.dlambda = d->dlambda;
  m.pot_deg = d->pot_deg;
  m.pot_rows = d->pot_rows;
  m.pot_coef = d->pot_coef;
  m.pot_dv = d->pot_dv;
  m.pot_expr = d->pot_expr;
  m.pot_expr_len = d->pot_expr_len;
  m.pot_expr_kvn = d->pot_expr_kvn;
  m.pot_t_step = d->pot_t_step;
  m.odm_g = d->odm_g;
  m.odm_x = d->odm_x;
  m.odm_th = d->odm_th;
  m.odm_sign = d->odm_sign;
  m.theta_decay = d->theta_decay;
  m.lambda_decay = d->lambda_decay;
  m.x_mask = d->x_mask;
  m.sigma_p = d->sigma_p;
  m.theta_host = d->theta_host;
  m.lambda_host = d->lambda_host;
  m.x_host = d->x_host;
  m.p_host = d->p_host;
  m.dv_host = d->dv_host;
  m.v0_host = d->v0_host;
  m.dxdp = d->dxdp;
  m.inv_2m = d->inv_2m;
  m.dx_host = d->dx_host;
  m.gp_rho_factor = d->gp_rho_factor;
  auto p = std::make_unique<wh_wigner2d>();
  WH_TRY(ensemble_create(&c, &m, 1, &p->e));
  *out = p.release();
  return WH_SUCCESS;
}

extern "C" int wh_wigner2d_destroy(wh_wigner2d* plan) {
  if (!plan) return fail(WH_INVALID_PLAN, "wh_wigner2d_destroy: plan is NULL");
  delete plan;
  return WH_SUCCESS;
}

extern "C" wh_ensemble* wh_wigner2d_ensemble(wh_wigner2d* plan) { return plan ? plan->e : nullptr; }

extern "C" int wh_wigner2d_run(wh_wigner2d* p, void* W, int first_step, int n_steps, double* obs,
                               int normalize_initial, int use_graph, void* stream) {
  if (!p) return fail(WH_INVALID_PLAN, "wh_wigner2d_run: plan is NULL");
  void* states[1] = {W};
  return ensemble_run(p->e, states, first_step, n_steps, obs, normalize_initial, use_graph,
                      static_cast<hipStream_t>(stream));
}

namespace {
int bind_single(wh_ensemble* e, void* W, double* obs, int obs_base, hipStream_t s) {
  wh::Member& m = e->mem[0];
  const bool changed = m.W != W || m.obs != obs || m.obs_base != obs_base;
  m.W = W;
  m.obs = obs;
  m.obs_base = obs_base;
  if (changed || e->table_dirty) return bind_states(e, s);
  return WH_SUCCESS;
}
}  // namespace

extern "C" int wh_wigner2d_kinetic(wh_wigner2d* p, void* W, void* stream) {
  if (!p) return fail(WH_INVALID_PLAN, "wh_wigner2d_kinetic: plan is NULL");
  if (!W) return fail(WH_INVALID_VALUE, "wh_wigner2d_kinetic: W is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  WH_TRY(bind_single(p->e, W, p->e->mem[0].obs, p->e->mem[0].obs_base, s));
  return xsweep(p->e, 0, s);
}

extern "C" int wh_wigner2d_potential(wh_wigner2d* p, void* W, int step, int fused, void* stream) {
  if (!p) return fail(WH_INVALID_PLAN, "wh_wigner2d_potential: plan is NULL");
  if (!W || step < 0) return fail(WH_INVALID_VALUE, "wh_wigner2d_potential: W %p step %d", W, step);
  hipStream_t s = static_cast<hipStream_t>(stream);
  wh_ensemble* e = p->e;
  WH_TRY(bind_single(e, W, e->mem[0].obs, e->mem[0].obs_base, s));
  const double frac = e->d.scheme == WH_SCHEME_LIE ? 1.0 : 0.5;
  const int flags = wh::kStepArg | (fused ? (wh::kColSums | (e->cl() ? wh::kCL : 0)) : 0);
  return psweep(e, flags, frac, step, s);
}

extern "C" int wh_ensemble_set_potential(wh_ensemble* e, int member, const double* coef, int rows, int deg,
                                         void* stream) {
  if (!e) return fail(WH_INVALID_PLAN, "wh_ensemble_set_potential: NULL");
  if (member < 0 || member >= e->n || !coef || rows < 1 || deg < 0 || deg > WH_MAX_DEG)
    return fail(WH_INVALID_VALUE, "wh_ensemble_set_potential: member %d coef %p rows %d deg %d", member,
                (const void*)coef, rows, deg);
  hipStream_t s = static_cast<hipStream_t>(stream);
  wh::Member& m = e->mem[member];
  MemberBufs& b = *e->bufs[member];
  if (m.pot_deg == WH_POT_TABLE || m.pot_deg == WH_POT_EXPR)
    return fail(WH_INVALID_VALUE, "wh_ensemble_set_potential: member %d has a non-polynomial V", member);
  const size_t bytes = size_t(rows) * (WH_MAX_DEG + 1) * sizeof(double);
  bool changed = false;
  if (b.pot.bytes != bytes) {
    WH_HIP(hipStreamSynchronize(s), "sync");
    WH_TRY(b.pot.alloc(bytes));
    m.pot = b.pot.as<double>();
    changed = true;
  }
  WH_HIP(hipMemcpyAsync(b.pot.p, coef, bytes, hipMemcpyDefault, s), "potential upload");
  changed = changed || m.pot_rows != rows || m.pot_deg != deg;
  m.pot_rows = rows;
  m.pot_deg = deg;
  return sync_table(e, changed, s);
}

extern "C" int wh_wigner2d_adjoint_step(wh_wigner2d* p, void* Lambda, const void* W_next, int step, double* grad,
                                        void* stream) {
  if (!p) return fail(WH_INVALID_PLAN, "wh_wigner2d_adjoint_step: plan is NULL");
  if (!Lambda || !W_next || !grad || step < 0)
    return fail(WH_INVALID_VALUE, "wh_wigner2d_adjoint_step: Lambda %p W_next %p grad %p step %d", Lambda, W_next,
                (void*)grad, step);
  wh_ensemble* e = p->e;
  if (e->d.scheme != WH_SCHEME_LIE || e->cl() || e->odm() || e->gp())
    return fail(WH_INVALID_VALUE, "the adjoint step covers the Lie scheme without damping or Gross-Pitaevskii");
  if (e->mem[0].pot_deg < 0)
    return fail(WH_INVALID_VALUE, "the adjoint step needs V as polynomial coefficients (gradients are per coefficient)");
  if (e->d.precision == WH_C64_WIDE)
    return fail(WH_INVALID_TYPE, "the adjoint step runs on WH_C128 or WH_C64 states");
  hipStream_t s = static_cast<hipStream_t>(stream);
  wh::Member& m = e->mem[0];
  if (!e->work_adj.n) {
    const int cols = wh::adj_cols(m.P);
    if (cols < 1 || cols > 64 || m.X % cols != 0 || m.P < 128)
      return fail(WH_INVALID_SIZE, "adjoint sweep does not fit a %d x %d grid", m.P, m.X);
    std::vector<unsigned> w;
    for (int t = 0; t < m.X / cols; ++t) w.push_back(unsigned(t));
    WH_TRY(e->work_adj.upload(w));
    WH_TRY(e->gcols.alloc(size_t(m.X) * (wh::kMaxDeg + 1) * sizeof(double)));
  }
  const bool changed = m.W != Lambda || m.W2 != W_next || m.gcols != e->gcols.as<double>();
  m.W = Lambda;
  m.W2 = W_next;
  m.gcols = e->gcols.as<double>();
  WH_TRY(sync_table(e, changed, s));
  WH_LAUNCH(wh::launch_set_int(e->step.as<int>(), step, s), "set_step");
  wh::SweepArgs a = sweep_args(e, 0, 1.0, 0);
  a.work = e->work_adj.buf.as<unsigned>();
  WH_LAUNCH(wh::launch_adj_psweep(e->d.precision, a, e->work_adj.n, s), "adj_psweep");
  WH_LAUNCH(wh::launch_grad_reduce(e->mem_dev.as<wh::Member>(), 1, grad, a.vphase / double(m.P), s), "grad_reduce");
  return xsweep(e, wh::kConj, s);
}

extern "C" int wh_wigner2d_observe(wh_wigner2d* p, const void* W, double* obs_row, void* stream) {
  if (!p) return fail(WH_INVALID_PLAN, "wh_wigner2d_observe: plan is NULL");
  if (!W || !obs_row) return fail(WH_INVALID_VALUE, "wh_wigner2d_observe: W %p obs %p", W, (void*)obs_row);
  hipStream_t s = static_cast<hipStream_t>(stream);
  wh_ensemble* e = p->e;
  WH_TRY(bind_single(e, const_cast<void*>(W), obs_row, 0, s));
  WH_LAUNCH(wh::launch_colsum(storage_code(e->d.precision), e->mem_dev.as<wh::Member>(),
                              e->work_colsum.buf.as<unsigned>(),
                              e->work_colsum.n, s),
            "colsum");
  return reduce(e, 0, 1, 0, 0, false, s, true, 0);
}

extern "C" int wh_wigner2d_observe_noalloc(const wh_observe_desc* d, const void* W, double* obs_row,
                                           void* stream) {
  if (!d) return fail(WH_INVALID_VALUE, "wh_wigner2d_observe_noalloc: desc is NULL");
  if (d->struct_size != (int)sizeof(wh_observe_desc))
    return fail(WH_INVALID_DESCRIPTOR, "struct_size %d != %d", d->struct_size, (int)sizeof(wh_observe_desc));
  if (d->precision != WH_C128 && d->precision != WH_C64 && d->precision != WH_C64_WIDE)
    return fail(WH_INVALID_TYPE, "precision %d is not WH_C128, WH_C64 or WH_C64_WIDE", d->precision);
  if (d->x_dim < 1 || d->x_dim > wh::noalloc_max_x() || d->p_dim < 1)
    return fail(WH_INVALID_SIZE, "wh_wigner2d_observe_noalloc: grid %d x %d: x_dim must lie in [1, %d]",
                d->p_dim, d->x_dim, wh::noalloc_max_x());
  if (!d->x_host || !d->p_host || !d->dv_host || !d->v0_host)
    return fail(WH_INVALID_DESCRIPTOR, "wh_wigner2d_observe_noalloc: x_host, p_host, dv_host and v0_host are needed");
  if (!W || !obs_row) return fail(WH_INVALID_VALUE, "wh_wigner2d_observe_noalloc: W %p obs %p", W, (void*)obs_row);
  wh::Member m{};
  m.W = const_cast<void*>(W);
  m.obs = obs_row;
  m.x_host = d->x_host;
  m.p_host = d->p_host;
  m.dv_host = d->dv_host;
  m.v0_host = d->v0_host;
  m.X = d->x_dim;
  m.P = d->p_dim;
  m.dxdp = d->dxdp;
  m.inv_2m = d->inv_2m;
  WH_LAUNCH(wh::launch_observe_noalloc(storage_code(d->precision), m, static_cast<hipStream_t>(stream)),
            "observe_noalloc");
  return WH_SUCCESS;
}

// ============================================================================ Wigner-Dirac 4x4
struct wh_dirac {
  wh_dirac_desc d{};
  wh::DiracPlan host{};  // host mirror of the device plan
  DevBuf plan_dev, pot, pot_dv, theta_decay, xmask, x_host, p_host, theta_host, dv0, vsum, mass_dm, mass_sm;
  DevBuf colsums, rowpart, scale, step, rho, G, lambda_host, acoef_rows, pot_expr, dpot_expr, mass_expr;
  DevBuf a_expr[3], a_tab, asum;
  TwSet twx, twp;
  int xblocks = 0, pblocks = 0, gpblocks = 0, pgblocks = 0;
  hipStream_t capture_stream = nullptr;
  std::map<std::tuple<hipStream_t, int>, GraphEntry> graphs;
  int graph_chunk = 16;
  bool plan_uploaded = false;
  ~wh_dirac() {
    for (auto& kv : graphs) {
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
      if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }
};

namespace {

int dirac_validate(const wh_dirac_desc* d) {
  if (!d) return fail(WH_INVALID_VALUE, "wh_dirac_create: desc is NULL");
  if (d->struct_size != (int)sizeof(wh_dirac_desc))
    return fail(WH_INVALID_DESCRIPTOR, "struct_size %d != %d", d->struct_size, (int)sizeof(wh_dirac_desc));
  if (d->precision != WH_C128 && d->precision != WH_C64 && d->precision != WH_C64_WIDE)
    return fail(WH_INVALID_TYPE, "precision %d is not WH_C128, WH_C64 or WH_C64_WIDE", d->precision);
  if (!sweep_size_supported(d->x_dim) || !sweep_size_supported(d->p_dim))
    return fail(WH_INVALID_SIZE, "grid %d x %d: both sides must be powers of two in [128, 2048]", d->p_dim, d->x_dim);
  if (d->pot_deg == WH_POT_TABLE) {
    if (!d->pot_dv || !d->dv0 || !d->vsum)
      return fail(WH_INVALID_DESCRIPTOR, "pot_deg WH_POT_TABLE needs pot_dv, dv0 and vsum");
  } else if (d->pot_deg == WH_POT_EXPR) {
    WH_TRY(check_expr(d->pot_expr, d->pot_expr_len, 0));
    WH_TRY(check_expr(d->dpot_expr, d->dpot_expr_len, 0));
  } else if (d->pot_deg < 0 || d->pot_deg > WH_MAX_DEG || d->pot_rows < 1 || !d->pot_coef) {
    return fail(WH_INVALID_DESCRIPTOR, "potential deg %d rows %d coef %p", d->pot_deg, d->pot_rows,
                (const void*)d->pot_coef);
  }
  if (!d->x_host || !d->p_host || !d->theta_host) return fail(WH_INVALID_DESCRIPTOR, "host grids missing");
  if ((d->mass_dm == nullptr) != (d->mass_sm == nullptr))
    return fail(WH_INVALID_DESCRIPTOR, "mass_dm and mass_sm go together");
  for (int k = 0; k < 3; ++k)
    if (d->a_expr_len[k] != 0 || d->a_expr[k]) {
      WH_TRY(check_expr(d->a_expr[k], d->a_expr_len[k], k + 1));
      if (d->a_deg[k] != 0)
        return fail(WH_INVALID_DESCRIPTOR, "vector potential %d: a program needs a_deg 0", k + 1);
    }
  if (d->mass_expr || d->mass_expr_len) {
    if (d->mass_dm) return fail(WH_INVALID_DESCRIPTOR, "mass tables and a mass program exclude each other");
    if (d->kvn) return fail(WH_INVALID_DESCRIPTOR, "the KvN variant takes no mass function");
    WH_TRY(check_expr(d->mass_expr, d->mass_expr_len, 0));
  }
  for (int k = 0; k < 3; ++k)
    if (d->a_deg[k] < -1 || d->a_deg[k] > WH_MAX_DEG)
      return fail(WH_INVALID_DESCRIPTOR, "vector potential %d: degree %d", k + 1, d->a_deg[k]);
  if (d->a_rows < 0 || (d->a_rows > 0 && !d->a_coef_rows))
    return fail(WH_INVALID_DESCRIPTOR, "vector potential rows %d, coefficients %p", d->a_rows,
                (const void*)d->a_coef_rows);
  if (!(d->dt == d->dt) || d->c == 0.0) return fail(WH_INVALID_DESCRIPTOR, "dt %g c %g", d->dt, d->c);
  if (d->kvn) {
    if (d->a_deg[0] >= 0 || d->a_deg[1] >= 0 || d->a_deg[2] >= 0 || d->mass_dm || d->gp_coeff != 0.0 ||
        d->gamma_dt != 0.0 || d->odm)
      return fail(WH_INVALID_DESCRIPTOR,
                  "the KvN variant takes V0 only (no vector potential, position-dependent mass, GP or damping)");
  }
  if (d->odm) {
    if (d->gamma_dt != 0.0) return fail(WH_INVALID_DESCRIPTOR, "ODM and Caldeira-Leggett damping exclude each other");
    if (!(d->lambda_bar > 0.0)) return fail(WH_INVALID_DESCRIPTOR, "ODM damping needs lambda_bar > 0");
  }
  return WH_SUCCESS;
}

// The potential half of a step: A = 0 (diagonal factors), vector potentials (two sweeps) or
// A = 0 with ODM damping (one sweep over plane quads).
// The record row and normalisation of the state at time index step + state_add; with a V0
// program, its dV0/dx and V0 at that time first.
int dirac_reduce(wh_dirac* e, int state_add, int renorm, int anti, int inc, hipStream_t s) {
  const auto* dp = e->plan_dev.as<wh::DiracPlan>();
  if (e->host.pot_expr_len > 0) {
    const int depth = std::max(e->host.pot_expr_depth, e->host.dpot_expr_depth);
    WH_LAUNCH(wh::launch_dirac_expr_obs(dp, e->d.x_dim, depth, state_add, s), "dirac_expr_obs");
  }
  if (e->host.asum)
    WH_LAUNCH(wh::launch_dirac_a_sum(dp, e->d.x_dim, e->host.a_expr_depth, state_add, s), "dirac_a_sum");
  WH_LAUNCH(wh::launch_dirac_reduce(dp, state_add, renorm, anti, inc, s), "dirac_reduce");
  return WH_SUCCESS;
}

int dirac_potential_sweep(wh_dirac* e, hipStream_t s) {
  const auto* dp = e->plan_dev.as<wh::DiracPlan>();
  const int prec = e->d.precision;
  if (e->host.pot_expr_len > 0 && e->d.kvn) {
    // KvN: the potential factor reads dV0/dx(t, x) per column; refresh it (and V0) at this
    // sweep's time index
    const int depth = std::max(e->host.pot_expr_depth, e->host.dpot_expr_depth);
    WH_LAUNCH(wh::launch_dirac_expr_obs(dp, e->d.x_dim, depth, 0, s), "dirac_expr_obs");
  } else if (e->host.pot_expr_len > 0) {  // V0(t, x) program: the Bopp difference at this sweep's time index
    WH_LAUNCH(wh::launch_dirac_expr_table(dp, e->d.x_dim * e->d.p_dim, e->host.pot_expr_depth, s),
              "dirac_expr_table");
  }
  if (e->host.a_tab)  // vector potential programs at this sweep's time index
    WH_LAUNCH(wh::launch_dirac_a_table(dp, e->d.x_dim * e->d.p_dim, e->host.a_expr_depth, s), "dirac_a_table");
  if (e->host.mass_expr_len > 0)  // Mass(t, x) program: the mass tables at this sweep's time index
    WH_LAUNCH(wh::launch_dirac_mass_table(dp, e->d.x_dim * e->d.p_dim, e->host.mass_expr_depth, s),
              "dirac_mass_table");
  if (e->host.odm && e->host.general) {
    // A and ODM: the two factor sweeps stay in the theta picture, the ODM sweep transforms back
    WH_LAUNCH(wh::launch_dirac_pgeneral(prec, false, dp, e->pgblocks, s), "dirac_pgeneral");
    WH_LAUNCH(wh::launch_dirac_podm(prec, dp, e->pgblocks, s), "dirac_podm");
  } else if (e->host.odm)
    WH_LAUNCH(wh::launch_dirac_podm(prec, dp, e->pgblocks, s), "dirac_podm");
  else if (e->host.general)
    WH_LAUNCH(wh::launch_dirac_pgeneral(prec, e->d.gamma_dt != 0.0, dp, e->pgblocks, s), "dirac_pgeneral");
  else
    WH_LAUNCH(wh::launch_dirac_psweep(prec, e->d.p_dim, e->d.gamma_dt != 0.0, dp, e->pblocks, s), "dirac_psweep");
  return WH_SUCCESS;
}

// One iteration: kinetic sweep, potential sweep, observables of the new state + renormalisation.
int dirac_iteration(wh_dirac* e, hipStream_t s) {
  const auto* dp = e->plan_dev.as<wh::DiracPlan>();
  const int prec = e->d.precision;
  if (e->d.gp_coeff != 0.0) WH_LAUNCH(wh::launch_dirac_gpshift(dp, e->gpblocks, s), "dirac_gpshift");
  WH_LAUNCH(wh::launch_dirac_xsweep(prec, e->d.x_dim, dp, e->xblocks, s), "dirac_xsweep");
  WH_TRY(dirac_potential_sweep(e, s));
  if (e->host.filter) {
    // antiparticle step filter and the absorber after it; the record then needs the column
    // sums of the filtered state
    WH_LAUNCH(wh::launch_dirac_fsweep(prec, e->d.x_dim, dp, e->xblocks, s), "dirac_fsweep");
    WH_LAUNCH(wh::launch_dirac_colsum(storage_code(prec), dp, e->d.x_dim, s), "dirac_colsum");
  }
  WH_TRY(dirac_reduce(e, 1, 1, 1, 1, s));
  return WH_SUCCESS;
}

int dirac_iterations(wh_dirac* e, int count, hipStream_t s) {
  for (int k = 0; k < count; ++k) WH_TRY(dirac_iteration(e, s));
  return WH_SUCCESS;
}

// The plan's step graph for stream s (graph_chunk iterations): captured (capturing
// executes nothing), instantiated and uploaded on the first graph-enabled run whatever its
// length, so that later runs replay it without capture or upload work.
int dirac_ensure_graph(wh_dirac* e, hipStream_t s, hipGraphExec_t* exec) {
  const int chunk = e->graph_chunk;
  auto key = std::make_tuple(s, chunk);
  auto it = e->graphs.find(key);
  if (it == e->graphs.end()) {
    if (!e->capture_stream)
      WH_HIP(hipStreamCreateWithFlags(&e->capture_stream, hipStreamNonBlocking), "hipStreamCreate");
    GraphEntry g;
    WH_HIP(hipStreamBeginCapture(e->capture_stream, hipStreamCaptureModeThreadLocal), "BeginCapture");
    const int rc = dirac_iterations(e, chunk, e->capture_stream);
    const hipError_t err = hipStreamEndCapture(e->capture_stream, &g.graph);
    if (rc) {
      if (g.graph) (void)hipGraphDestroy(g.graph);
      return rc;
    }
    if (err != hipSuccess) return hip_fail(err, "EndCapture");
    WH_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "GraphInstantiate");
    WH_HIP(hipGraphUpload(g.exec, s), "GraphUpload");
    it = e->graphs.emplace(key, g).first;
  }
  *exec = it->second.exec;
  return WH_SUCCESS;
}

// Replays `count` iterations from the step graph (whole chunks), the rest launched directly.
int dirac_run_graph(wh_dirac* e, int count, hipStream_t s) {
  hipGraphExec_t exec = nullptr;
  WH_TRY(dirac_ensure_graph(e, s, &exec));
  while (count >= e->graph_chunk) {
    WH_HIP(hipGraphLaunch(exec, s), "GraphLaunch");
    count -= e->graph_chunk;
  }
  return dirac_iterations(e, count, s);
}

}  // namespace

extern "C" int wh_dirac_create(const wh_dirac_desc* desc, wh_dirac** out) {
  if (!out) return fail(WH_INVALID_VALUE, "wh_dirac_create: out is NULL");
  *out = nullptr;
  WH_TRY(dirac_validate(desc));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(WH_NO_DEVICE, "no HIP device");
  auto e = std::make_unique<wh_dirac>();
  e->d = *desc;
  const auto& d = *desc;
  const size_t X = size_t(d.x_dim), P = size_t(d.p_dim);
  e->xblocks = wh::dirac_x_blocks(d.precision, d.x_dim, d.p_dim);
  e->pblocks = wh::dirac_p_blocks(d.x_dim, d.p_dim);
  if (e->xblocks < 1 || e->pblocks < 1) return fail(WH_INVALID_SIZE, "grid %zu x %zu", P, X);
  WH_TRY(make_twiddles(e->twx, d.x_dim, d.precision, d.gp_coeff != 0.0));
  WH_TRY(make_twiddles(e->twp, d.p_dim, d.precision, false));
  if (d.pot_deg == WH_POT_TABLE) {
    const double zero_row[WH_MAX_DEG + 1] = {};
    WH_TRY(e->pot.upload(zero_row, sizeof(zero_row)));
    WH_TRY(e->pot_dv.upload(d.pot_dv, X * P * sizeof(double)));
    WH_TRY(e->dv0.upload(d.dv0, X * sizeof(double)));
    WH_TRY(e->vsum.upload(d.vsum, X * sizeof(double)));
  } else if (d.pot_deg == WH_POT_EXPR) {
    // the tables are filled on the device from the programs
    const double zero_row[WH_MAX_DEG + 1] = {};
    WH_TRY(e->pot.upload(zero_row, sizeof(zero_row)));
    WH_TRY(e->pot_dv.alloc(X * P * sizeof(double)));
    WH_TRY(e->dv0.alloc(X * sizeof(double)));
    WH_TRY(e->vsum.alloc(X * sizeof(double)));
    WH_TRY(e->pot_expr.upload(d.pot_expr, size_t(d.pot_expr_len) * 2 * sizeof(double)));
    WH_TRY(e->dpot_expr.upload(d.dpot_expr, size_t(d.dpot_expr_len) * 2 * sizeof(double)));
  } else {
    WH_TRY(e->pot.upload(d.pot_coef, size_t(d.pot_rows) * (WH_MAX_DEG + 1) * sizeof(double)));
  }
  WH_TRY(upload_opt(e->theta_decay, d.theta_decay, P));
  WH_TRY(upload_opt(e->xmask, d.x_mask, X));
  if (d.mass_dm) {
    WH_TRY(e->mass_dm.upload(d.mass_dm, X * P * sizeof(double)));
    WH_TRY(e->mass_sm.upload(d.mass_sm, X * P * sizeof(double)));
  } else if (d.mass_expr_len > 0) {  // tables refilled on the device from the program
    WH_TRY(e->mass_dm.alloc(X * P * sizeof(double)));
    WH_TRY(e->mass_sm.alloc(X * P * sizeof(double)));
    WH_TRY(e->mass_expr.upload(d.mass_expr, size_t(d.mass_expr_len) * 2 * sizeof(double)));
  }
  WH_TRY(e->x_host.upload(d.x_host, X * sizeof(double)));
  WH_TRY(e->p_host.upload(d.p_host, P * sizeof(double)));
  WH_TRY(e->theta_host.upload(d.theta_host, P * sizeof(double)));
  WH_TRY(e->colsums.alloc(X * wh::kDNCol * sizeof(double)));
  WH_TRY(e->rowpart.alloc(2 * P * sizeof(double)));
  const double one = 1.0;
  WH_TRY(e->scale.upload(&one, sizeof(one)));
  const int zero = 0;
  WH_TRY(e->step.upload(&zero, sizeof(zero)));
  WH_TRY(e->plan_dev.alloc(sizeof(wh::DiracPlan)));
  if (d.gp_coeff != 0.0) {
    if (!d.lambda_host) return fail(WH_INVALID_DESCRIPTOR, "Gross-Pitaevskii needs lambda_host");
    e->gpblocks = wh::dirac_gp_blocks(d.x_dim, d.p_dim);
    if (e->gpblocks < 1) return fail(WH_INVALID_SIZE, "grid %zu x %zu", P, X);
    WH_TRY(e->rho.alloc(X * sizeof(double)));
    WH_TRY(e->G.alloc(X * P * sizeof(double)));
    WH_TRY(e->lambda_host.upload(d.lambda_host, X * sizeof(double)));
  }

  wh::DiracPlan& h = e->host;
  std::memset(&h, 0, sizeof(h));
  h.colsums = e->colsums.as<double>();
  h.rowpart = e->rowpart.as<double>();
  h.scale = e->scale.as<double>();
  h.step = e->step.as<int>();
  h.pot = e->pot.as<double>();
  h.pot_dv = e->pot_dv.as<double>();
  h.theta_decay = e->theta_decay.as<double>();
  h.xmask = e->xmask.as<double>();
  h.x_host = e->x_host.as<double>();
  h.p_host = e->p_host.as<double>();
  h.theta_host = e->theta_host.as<double>();
  h.dv0 = e->dv0.as<double>();
  h.vsum = e->vsum.as<double>();
  h.mass_dm = e->mass_dm.as<double>();
  h.mass_sm = e->mass_sm.as<double>();
  if (d.a_expr_len[0] > 0 || d.a_expr_len[1] > 0 || d.a_expr_len[2] > 0) {
    // vector potential programs: A_k at x -/+ theta/2 per sweep, their sum per record row
    WH_TRY(e->a_tab.alloc(6 * X * P * sizeof(double)));
    WH_TRY(e->asum.alloc(X * sizeof(double)));
    h.a_tab = e->a_tab.as<double>();
    h.asum = e->asum.as<double>();
    h.a_expr_depth = 1;
    for (int k = 0; k < 3; ++k) {
      if (d.a_expr_len[k] <= 0) continue;
      WH_TRY(e->a_expr[k].upload(d.a_expr[k], size_t(d.a_expr_len[k]) * 2 * sizeof(double)));
      h.a_expr[k] = e->a_expr[k].as<double>();
      h.a_expr_len[k] = d.a_expr_len[k];
      int depth = 0;
      WH_TRY(check_expr(d.a_expr[k], d.a_expr_len[k], k + 1, &depth));
      h.a_expr_depth = std::max(h.a_expr_depth, depth);
    }
    h.a_t_step = d.a_t_step;
  }
  if (d.mass_expr_len > 0) {
    h.mass_expr = e->mass_expr.as<double>();
    h.mass_expr_len = d.mass_expr_len;
    WH_TRY(check_expr(d.mass_expr, d.mass_expr_len, 0, &h.mass_expr_depth));
    h.mass_c = d.c;
    h.mass_dt = d.dt;
    h.mass_t_step = d.mass_t_step;
  }
  h.rho = e->rho.as<double>();
  h.G = e->G.as<double>();
  h.lambda_host = e->lambda_host.as<double>();
  h.twX64 = e->twx.hi64.p;
  h.gp_coeff = d.gp_coeff;
  h.adeg0 = d.a_deg[0];
  h.adeg1 = d.a_deg[1];
  h.adeg2 = d.a_deg[2];
  h.general = (d.a_deg[0] >= 0 || d.a_deg[1] >= 0 || d.a_deg[2] >= 0) ? 1 : 0;
  std::memcpy(h.acoef, d.a_coef, sizeof(h.acoef));
  if (h.general && d.a_rows > 0) {
    WH_TRY(e->acoef_rows.upload(d.a_coef_rows, size_t(d.a_rows) * 3 * (WH_MAX_DEG + 1) * sizeof(double)));
    h.acoef_rows = e->acoef_rows.as<double>();
    h.a_rows = d.a_rows;
  }
  h.kvn = d.kvn ? 1 : 0;
  h.filter = d.step_filter ? 1 : 0;
  if (h.filter) {
    h.fmask = h.xmask;  // the absorber follows the filter
    h.xmask = nullptr;
  }
  h.odm = d.odm ? 1 : 0;
  h.odm_k = d.mass * d.c * d.dt * d.odm_gamma;
  h.odm_inv_lb = d.odm ? 1.0 / d.lambda_bar : 0.0;
  if (h.general || h.odm) {
    e->pgblocks = wh::dirac_pg_blocks(d.x_dim, d.p_dim);
    if (e->pgblocks < 1) return fail(WH_INVALID_SIZE, "grid %zu x %zu", P, X);
  }
  h.twX_hi = d.precision != WH_C64 ? e->twx.hi64.p : e->twx.hi.p;
  h.twX_lo = e->twx.lo.p;
  h.twP_hi = d.precision != WH_C64 ? e->twp.hi64.p : e->twp.hi.p;
  h.twP_lo = e->twp.lo.p;
  h.X = d.x_dim;
  h.P = d.p_dim;
  h.pot_rows = d.pot_deg == WH_POT_TABLE || d.pot_deg == WH_POT_EXPR ? 1 : d.pot_rows;
  h.pot_deg = d.pot_deg == WH_POT_EXPR ? WH_POT_TABLE : d.pot_deg;  // the sweeps read the tables
  if (d.pot_deg == WH_POT_EXPR) {
    h.pot_expr = e->pot_expr.as<double>();
    h.dpot_expr = e->dpot_expr.as<double>();
    h.pot_expr_len = d.pot_expr_len;
    h.dpot_expr_len = d.dpot_expr_len;
    WH_TRY(check_expr(d.pot_expr, d.pot_expr_len, 0, &h.pot_expr_depth));
    WH_TRY(check_expr(d.dpot_expr, d.dpot_expr_len, 0, &h.dpot_expr_depth));
    h.pot_t_step = d.pot_t_step;
  }
  h.energy = d.energy;
  h.antiparticle = d.antiparticle;
  h.dx = d.dx;
  h.dp = d.dp;
  h.dtheta = d.dtheta;
  h.dlambda = d.dlambda;
  h.mh_c = d.mass_half * d.c;
  h.c_dt = d.c * d.dt;
  h.mass_phase = d.mass_half * d.c * d.c * d.dt;
  h.dt = d.dt;
  h.proj_mc2 = d.mass * d.c * d.c;
  h.proj_c = d.c;
  h.cl_a = d.gamma_dt;
  h.dxdp = d.dxdp;
  h.c_host = d.c_host;
  h.mc2_host = d.mc2_host;
  h.transmission_x = d.transmission_x;
  *out = e.release();
  return WH_SUCCESS;
}

extern "C" int wh_dirac_destroy(wh_dirac* plan) {
  if (!plan) return fail(WH_INVALID_PLAN, "wh_dirac_destroy: NULL");
  delete plan;
  return WH_SUCCESS;
}

namespace {
// Points the device plan at W (and obs) in stream order and sets the step counter (a
// kernel writes the fields: no host synchronisation).  The first call uploads the plan.
int dirac_bind(wh_dirac* e, void* W, double* obs, int base, int step, bool reset_scale, hipStream_t s) {
  const size_t align = state_bytes(e->d.precision);
  if (reinterpret_cast<uintptr_t>(W) % align) return fail(WH_UNALIGNED_DATA, "W %p", W);
  e->host.W = W;
  e->host.obs = obs;
  e->host.obs_base = base;
  if (!e->plan_uploaded) {
    WH_HIP(hipMemcpyAsync(e->plan_dev.p, &e->host, sizeof(wh::DiracPlan), hipMemcpyHostToDevice, s), "plan upload");
    WH_HIP(hipStreamSynchronize(s), "plan upload");
    e->plan_uploaded = true;
  }
  WH_LAUNCH(wh::launch_dirac_bind(e->plan_dev.as<wh::DiracPlan>(), W, obs, base, e->step.as<int>(), step,
                                  e->scale.as<double>(), reset_scale ? 1 : 0, s),
            "dirac_bind");
  return WH_SUCCESS;
}
}  // namespace

extern "C" int wh_dirac_kinetic(wh_dirac* e, void* W, void* stream) {
  if (!e) return fail(WH_INVALID_PLAN, "wh_dirac_kinetic: plan is NULL");
  if (!W) return fail(WH_INVALID_VALUE, "wh_dirac_kinetic: W is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  WH_TRY(dirac_bind(e, W, e->host.obs, e->host.obs_base, 0, true, s));
  WH_LAUNCH(wh::launch_dirac_xsweep(e->d.precision, e->d.x_dim, e->plan_dev.as<wh::DiracPlan>(), e->xblocks, s), "dirac_xsweep");
  return WH_SUCCESS;
}

extern "C" int wh_dirac_potential(wh_dirac* e, void* W, int step, void* stream) {
  if (!e) return fail(WH_INVALID_PLAN, "wh_dirac_potential: plan is NULL");
  if (!W || step < 0) return fail(WH_INVALID_VALUE, "wh_dirac_potential: W %p step %d", W, step);
  if (e->d.gp_coeff != 0.0)
    return fail(WH_INVALID_VALUE, "wh_dirac_potential: the Gross-Pitaevskii density comes from wh_dirac_run");
  hipStream_t s = static_cast<hipStream_t>(stream);
  WH_TRY(dirac_bind(e, W, e->host.obs, e->host.obs_base, step, false, s));
  WH_TRY(dirac_potential_sweep(e, s));
  return WH_SUCCESS;
}

extern "C" int wh_dirac_run(wh_dirac* e, void* W, int first_step, int n_steps, double* obs, int use_graph,
                            void* stream) {
  if (!e) return fail(WH_INVALID_PLAN, "wh_dirac_run: plan is NULL");
  if (!W || !obs || n_steps < 0 || first_step < 0)
    return fail(WH_INVALID_VALUE, "wh_dirac_run: W %p obs %p n_steps %d first_step %d", W, (void*)obs, n_steps,
                first_step);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the device plan is read by every kernel (and by captured graphs): update it in stream order
  WH_TRY(dirac_bind(e, W, obs, first_step, first_step, false, s));
  const auto* dp = e->plan_dev.as<wh::DiracPlan>();
  WH_LAUNCH(wh::launch_dirac_colsum(storage_code(e->d.precision), dp, e->d.x_dim, s), "dirac_colsum");
  WH_TRY(dirac_reduce(e, 0, 0, 0, 0, s));
  if (use_graph) {
    hipGraphExec_t exec = nullptr;
    WH_TRY(dirac_ensure_graph(e, s, &exec));  // ready for a later, longer run
  }
  if (graph_pays(e->graph_chunk, use_graph, n_steps))
    WH_TRY(dirac_run_graph(e, n_steps, s));
  else
    WH_TRY(dirac_iterations(e, n_steps, s));
  WH_LAUNCH(wh::launch_dirac_finish(storage_code(e->d.precision), dp, e->d.p_dim, s), "dirac_finish");
  return WH_SUCCESS;
}

// ============================================================================ Bloch solver
struct wh_bloch {
  wh::BlochArgs a{};
  TwSet twx, twp;
  DevBuf pot, vsum, rho, G, theta_host, lambda_host, x_host, p_host, k_host, v0_host, colsums, g_colsums, ctrl;
};

extern "C" int wh_bloch_create(const wh_bloch_desc* d, wh_bloch** out) {
  if (!out) return fail(WH_INVALID_VALUE, "wh_bloch_create: out is NULL");
  *out = nullptr;
  if (!d) return fail(WH_INVALID_VALUE, "wh_bloch_create: desc is NULL");
  if (d->struct_size != (int)sizeof(wh_bloch_desc))
    return fail(WH_INVALID_DESCRIPTOR, "wh_bloch_desc.struct_size %d != %d", d->struct_size, (int)sizeof(wh_bloch_desc));
  if (!sweep_size_supported(d->x_dim) || !sweep_size_supported(d->p_dim))
    return fail(WH_INVALID_SIZE, "Bloch grid %d x %d: both sides must be powers of two in [128, 2048]", d->p_dim,
                d->x_dim);
  if (d->pot_deg == WH_POT_TABLE) {
    if (!d->pot_sum) return fail(WH_INVALID_DESCRIPTOR, "pot_deg WH_POT_TABLE needs pot_sum");
  } else if (d->pot_deg < 0 || d->pot_deg > WH_MAX_DEG || d->pot_rows < 1 || !d->pot_coef) {
    return fail(WH_INVALID_DESCRIPTOR, "Bloch potential deg %d rows %d coef %p", d->pot_deg, d->pot_rows,
                (const void*)d->pot_coef);
  }
  if (d->kinetic_deg < 0 || d->kinetic_deg > WH_MAX_DEG)
    return fail(WH_INVALID_DESCRIPTOR, "Bloch kinetic degree %d", d->kinetic_deg);
  if (!d->theta_host || !d->lambda_host || !d->x_host || !d->p_host || !d->k_host || !d->v0_host)
    return fail(WH_INVALID_DESCRIPTOR, "Bloch: the host grids, K(p) and V(0, x) tables are required");
  if (!(d->dxdp > 0.0) || !(d->hbar != 0.0)) return fail(WH_INVALID_DESCRIPTOR, "Bloch: dxdp %g hbar %g", d->dxdp, d->hbar);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(WH_NO_DEVICE, "no HIP device");
  auto b = std::make_unique<wh_bloch>();
  const size_t X = size_t(d->x_dim), P = size_t(d->p_dim);
  WH_TRY(make_twiddles(b->twx, d->x_dim, WH_C128, false));
  WH_TRY(make_twiddles(b->twp, d->p_dim, WH_C128, false));
  if (d->pot_deg == WH_POT_TABLE) {
    const double zero_row[WH_MAX_DEG + 1] = {};
    WH_TRY(b->pot.upload(zero_row, sizeof(zero_row)));
    WH_TRY(b->vsum.upload(d->pot_sum, X * P * sizeof(double)));
  } else {
    WH_TRY(b->pot.upload(d->pot_coef, size_t(d->pot_rows) * (WH_MAX_DEG + 1) * sizeof(double)));
  }
  WH_TRY(b->theta_host.upload(d->theta_host, P * sizeof(double)));
  WH_TRY(b->lambda_host.upload(d->lambda_host, X * sizeof(double)));
  WH_TRY(b->x_host.upload(d->x_host, X * sizeof(double)));
  WH_TRY(b->p_host.upload(d->p_host, P * sizeof(double)));
  WH_TRY(b->k_host.upload(d->k_host, P * sizeof(double)));
  WH_TRY(b->v0_host.upload(d->v0_host, X * sizeof(double)));
  WH_TRY(b->colsums.alloc(X * wh::kBCol * sizeof(double)));
  WH_TRY(b->g_colsums.alloc(X * wh::kBCol * sizeof(double)));
  WH_TRY(b->ctrl.alloc(wh::kBCtrlSize * sizeof(double)));
  if (d->gp_coeff != 0.0) {
    WH_TRY(b->rho.alloc(X * sizeof(double)));
    WH_TRY(b->G.alloc(X * P * sizeof(double)));
  }
  wh::BlochArgs& a = b->a;
  a.twX = b->twx.hi64.as<double>();
  a.twP = b->twp.hi64.as<double>();
  a.X = d->x_dim;
  a.P = d->p_dim;
  a.kdx = d->dx;
  a.kdp = d->dp;
  a.kdtheta = d->dtheta;
  a.kdlambda = d->dlambda;
  a.khbar = d->hbar;
  a.ka = d->gp_coeff;
  a.pot_deg = d->pot_deg == WH_POT_TABLE ? 0 : d->pot_deg;
  a.pot_rows = d->pot_deg == WH_POT_TABLE ? 1 : d->pot_rows;
  a.pot = b->pot.as<double>();
  a.vsum = b->vsum.as<double>();
  a.kdeg = d->kinetic_deg;
  a.gp = d->gp_coeff != 0.0;
  for (int k = 0; k <= WH_MAX_DEG; ++k) a.kcoef[k] = d->kin_coef[k];
  a.rho = b->rho.as<double>();
  a.G = b->G.as<double>();
  a.theta_host = b->theta_host.as<double>();
  a.lambda_host = b->lambda_host.as<double>();
  a.x_host = b->x_host.as<double>();
  a.p_host = b->p_host.as<double>();
  a.k_host = b->k_host.as<double>();
  a.v0_host = b->v0_host.as<double>();
  a.colsums = b->colsums.as<double>();
  a.g_colsums = b->g_colsums.as<double>();
  a.ctrl = b->ctrl.as<double>();
  a.dxdp = d->dxdp;
  a.dx_host = d->dx_host;
  a.dp_host = d->dp_host;
  a.hbar_host = d->hbar_host;
  a.gp_host = d->gp_coeff_host;
  a.first_test = 0;
  *out = b.release();
  return WH_SUCCESS;
}

extern "C" int wh_bloch_destroy(wh_bloch* b) {
  if (!b) return fail(WH_INVALID_PLAN, "wh_bloch_destroy: NULL");
  delete b;
  return WH_SUCCESS;
}

extern "C" int wh_bloch_run(wh_bloch* b, void* W, void* W_step, const void* W_ground, int n_iter, double dt,
                            double* hist, void* stream) {
  if (!b) return fail(WH_INVALID_PLAN, "wh_bloch_run: plan is NULL");
  if (!W || !W_step || !hist) return fail(WH_INVALID_VALUE, "wh_bloch_run: W, W_step and hist are required");
  if (W == W_step || W == W_ground || W_step == W_ground)
    return fail(WH_INVALID_VALUE, "wh_bloch_run: W, W_step and W_ground must be distinct buffers");
  if (n_iter < 0) return fail(WH_INVALID_VALUE, "wh_bloch_run: n_iter %d", n_iter);
  if (!(dt > 0.0)) return fail(WH_INVALID_VALUE, "wh_bloch_run: dt %g", dt);
  for (const void* p : {(const void*)W, (const void*)W_step, W_ground})
    if (reinterpret_cast<uintptr_t>(p) % 16) return fail(WH_UNALIGNED_DATA, "wh_bloch_run: %p not 16-byte aligned", p);
  const hipStream_t s = static_cast<hipStream_t>(stream);
  wh::BlochArgs a = b->a;
  a.W = W;
  a.Wstep = W_step;
  a.Wground = nullptr;
  a.hist = hist;
  a.hist_base = 0;
  const size_t bytes = size_t(a.X) * size_t(a.P) * 16;
  // the last kept state starts as W_init as given (GPU_Wigner2D_GPitaevskii.py:2146, 2361)
  WH_HIP(hipMemcpyAsync(W_step, W, bytes, hipMemcpyDeviceToDevice, s), "wh_bloch_run: W_step");
  WH_LAUNCH(wh::launch_bloch_set(a.ctrl, dt, 0, s), "bloch set");
  // the initial state: normalise, energy of reference (no W_0 removal yet)
  WH_LAUNCH(wh::launch_bloch_colsum(a, W, a.colsums, 0, s), "bloch colsum");
  WH_LAUNCH(wh::launch_bloch_finish(a, wh::kBInit, s), "bloch finish");
  WH_LAUNCH(wh::launch_bloch_commit(a, s), "bloch commit");
  WH_LAUNCH(wh::launch_bloch_set(a.ctrl, dt, 0, s), "bloch set");
  if (W_ground) {
    a.Wground = W_ground;
    WH_LAUNCH(wh::launch_bloch_colsum(a, W_ground, const_cast<double*>(a.g_colsums), 0, s), "bloch colsum W_0");
  }
  for (int it = 0; it < n_iter; ++it) {
    if (a.gp) {
      WH_LAUNCH(wh::launch_bloch_colsum(a, W, nullptr, 1, s), "bloch density");
      WH_LAUNCH(wh::launch_bloch_gshift(a, s), "bloch gshift");
    }
    WH_LAUNCH(wh::launch_bloch_psweep(a, 0, s), "bloch psweep");
    WH_LAUNCH(wh::launch_bloch_xsweep(a, s), "bloch xsweep");
    if (a.gp) {
      WH_LAUNCH(wh::launch_bloch_colsum(a, W, nullptr, 1, s), "bloch density");
      WH_LAUNCH(wh::launch_bloch_gshift(a, s), "bloch gshift");
    }
    WH_LAUNCH(wh::launch_bloch_psweep(a, 1, s), "bloch psweep");
    WH_LAUNCH(wh::launch_bloch_finish(a, wh::kBStep, s), "bloch finish");
    WH_LAUNCH(wh::launch_bloch_commit(a, s), "bloch commit");
  }
  return WH_SUCCESS;
}
