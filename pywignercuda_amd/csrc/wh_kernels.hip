// Hand-written gfx950 kernels of the Wigner split-operator step.
//
//   xsweep   one team per row p:   FFT_x -> kinetic multiplier(p,lambda) -> iFFT_x/X
//            (GPU_Wigner2D_FFT.py:1885-1888, expPLambda :131-158; GP.py expPLambdaKinetic :117-158)
//   psweep   one team per column x of a C-column tile:  FFT_p -> potential multiplier
//            (theta,x) [+ Gross-Pitaevskii phase] -> iFFT_p/P -> [Caldeira-Leggett order-3
//            composite] -> column sums for the observables [-> next Strang half step]
//            (GPU_Wigner2D_FFT.py:1891-1913, expPotential :80-126, theta_fp :399-431,
//             CaldeiraDissipatorOrder3 :1145-1157; GP.py :1911-1970)
//   reduce   observables + complex norm from the column sums, renormalisation factor
//            for the next load (GPU_Wigner2D_FFT.py:1857-1878, 1912-1913)
//   fft_rows / fft_cols   plain batched axis transforms replacing the cuFFT plans
//            Plan_Z2Z_2D_Axis1 / Plan_Z2Z_2D_Axis0 (cufft_wrapper.py:183-250)
//
// The state is complex, [P][X] row-major, FFT ordered (index 0 = origin).  Each
// sweep reads and writes the state once; coordinates, phases and observables are
// computed in the kernel; 1-D tables (decays, mask, sigma(p), host grids) are a few
// KiB and stay in cache.
#include <hip/hip_runtime.h>

#include "wh_kernels.hpp"
#include "wh_math.hpp"

namespace wh {

// ----------------------------------------------------------------------------- helpers
__device__ __forceinline__ double poly(const double* __restrict__ c, int deg, double x) {
  double r = c[deg];
  for (int n = deg - 1; n >= 0; --n) r = fma(r, x, c[n]);
  return r;
}

// v * r * exp(-i phi); phi in FP64.  The FP32 form reduces phi in FP64 and
// evaluates the FP32 polynomials (wh_math.hpp).
__device__ __forceinline__ cplx<double> phase_mul(cplx<double> v, double r, double phi) {
  double s, c;
  sincos_fp64(phi, &s, &c);
  c *= r;
  s *= r;
  return {v.x * c + v.y * s, v.y * c - v.x * s};
}
__device__ __forceinline__ cplx<float> phase_mul(cplx<float> v, double r, double phi) {
  float s, c;
  sincos_fp32(phi, &s, &c);
  const float rf = float(r);
  c *= rf;
  s *= rf;
  return {v.x * c + v.y * s, v.y * c - v.x * s};
}

template <typename T>
__device__ __forceinline__ cplx<T> load_scale(const double* s) {
  return {T(s[0]), T(s[1])};
}

// kernel-side FFT-ordered coordinate: step * (((k + n/2) mod n) - n/2)
__device__ __forceinline__ double fcoord(int k, int n, double step) {
  return step * double(((k + (n >> 1)) & (n - 1)) - (n >> 1));
}

// Element of a uniform row base at a per-lane 32-bit byte offset: lets the
// compiler use the SGPR-base + VGPR-offset global addressing form, so a lane keeps
// one offset register instead of one 64-bit address per strided element.
template <typename T>
__device__ __forceinline__ cplx<T>& at_off(cplx<T>* base, unsigned boff) {
  return *reinterpret_cast<cplx<T>*>(reinterpret_cast<char*>(base) + boff);
}

// a * (i s)
template <typename T>
__device__ __forceinline__ cplx<T> mul_i(cplx<T> a, T s) { return {-s * a.y, s * a.x}; }

// ----------------------------------------------------------------------------- x sweep
template <typename T, int N, int R, int ROWS>
__global__ void __launch_bounds__(ROWS*(N / R)) xsweep_kernel(XSweepArgs a) {
  constexpr int TT = N / R;
  constexpr int LSZ = Padded<N>::size;
  __shared__ cplx<T> lds[ROWS * LSZ];
  const int tid = threadIdx.x;
  const int team = tid / TT, t = tid % TT;
  const int row = blockIdx.x * ROWS + team;
  cplx<T>* __restrict__ W = reinterpret_cast<cplx<T>*>(a.W) + size_t(row) * N;
  const cplx<T>* __restrict__ tw = reinterpret_cast<const cplx<T>*>(a.tw);
  LdsRow<T> view{lds + team * LSZ};

  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = W[t + j * TT];
  if (a.scale) {
    const cplx<T> s = load_scale<T>(a.scale);
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cmul(v[j], s);
  }
  team_fft<T, N, R, -1>(v, t, view, tw);

  const double p = fcoord(row, a.P, a.dp);
  const double hh = 0.5 * a.hbar;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int m = t + j * TT;
    const double lam = fcoord(m, N, a.dlambda);
    double phi;
    if (a.kinetic_poly) {
      phi = a.dt * (poly(a.kcoef, a.kdeg, p + hh * lam) - poly(a.kcoef, a.kdeg, p - hh * lam)) *
            a.inv_hbar;
    } else {
      phi = a.dt * lam * p * a.inv_mass;
    }
    const double r = a.lambda_decay ? a.lambda_decay[m] : 1.0;
    v[j] = phase_mul(v[j], r, phi);
    __builtin_amdgcn_sched_barrier(0);
  }
  team_fft<T, N, R, +1>(v, t, view, tw);
  const T inv = T(1.0 / N);
#pragma unroll
  for (int j = 0; j < R; ++j) W[t + j * TT] = cscale(v[j], inv);

  if (a.colpart) {
    // partial column sums of Re W over this block's rows (Gross-Pitaevskii density)
    __syncthreads();
    double* red = reinterpret_cast<double*>(lds);
#pragma unroll
    for (int j = 0; j < R; ++j) red[team * N + t + j * TT] = double(v[j].x) * double(inv);
    __syncthreads();
    for (int x = tid; x < N; x += ROWS * TT) {
      double s = 0.0;
      for (int q = 0; q < ROWS; ++q) s += red[q * N + x];
      a.colpart[size_t(blockIdx.x) * N + x] = s;
    }
  }
}

// ----------------------------------------------------------------------------- p sweep
template <typename T, int N, int R, int C>
struct PSweep {
  static constexpr int TT = N / R;
  static constexpr int LSZ = Padded<N>::size;
  using View = LdsStrided<T, C>;

  // multiply by the potential step multiplier of time index `step`
  __device__ __forceinline__ static void potential(cplx<T> (&v)[R], const PSweepArgs& a, int t_in,
                                                   int col_in, int step) {
    // opaque lane indices: recomputing theta and x -/+ hbar theta/2 per call is cheaper
    // than keeping them live across the transforms between two calls
    const int t = opaque_v(t_in), col = opaque_v(col_in);
    const int X = a.X;
    const double x = fcoord(col, X, a.dx);
    int row = step < a.pot.rows ? step : a.pot.rows - 1;
    const double* __restrict__ cf = a.pot.coef + size_t(row) * (kMaxDeg + 1);
    const double mask = a.xmask ? a.xmask[col] : 1.0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = t + j * TT;
      const double th = fcoord(i, N, a.dtheta);
      const double u = a.hbar_half * th;
      double phi = a.vphase * (poly(cf, a.pot.deg, x - u) - poly(cf, a.pot.deg, x + u));
      if (a.gp_G) phi -= a.gp_phase * a.gp_G[size_t(i) * X + col];
      const double r = (a.theta_decay ? a.theta_decay[i] : 1.0) * mask;
      v[j] = phase_mul(v[j], r, phi);
      __builtin_amdgcn_sched_barrier(0);  // keep the 16 phase evaluations from interleaving
    }
  }

  // D(Y) = F_p^{-1}[ theta * F_p[ sigma(p) Y ] ]   (Theta_fp_Damping, FFT.py:1135-1143)
  __device__ __forceinline__ static void cl_operator(cplx<T> (&y)[R], const PSweepArgs& a, int t,
                                                     View view, const cplx<T>* tw) {
#pragma unroll
    for (int j = 0; j < R; ++j) y[j] = cscale(y[j], T(a.sigma_p[t + j * TT]));
    team_fft<T, N, R, -1>(y, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) y[j] = cscale(y[j], T(a.theta_host[t + j * TT]));
    team_fft<T, N, R, +1>(y, t, view, tw);
    const T inv = T(1.0 / N);
#pragma unroll
    for (int j = 0; j < R; ++j) y[j] = cscale(y[j], inv);
  }
};

template <typename T, int N, int R, int C, bool CL>
__global__ void __launch_bounds__(C*(N / R)) psweep_kernel(PSweepArgs a) {
  using S = PSweep<T, N, R, C>;
  constexpr int TT = S::TT;
  __shared__ cplx<T> lds[C * S::LSZ];
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int X = a.X;
  const int col = blockIdx.x * C + c;
  cplx<T>* __restrict__ W = reinterpret_cast<cplx<T>*>(a.W);
  const cplx<T>* __restrict__ tw = reinterpret_cast<const cplx<T>*>(a.tw);
  typename S::View view{lds + c};
  const int step = *a.step + a.step_add;

  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<T>));
  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = at_off(W + size_t(j * TT) * X, boff);
  if (a.scale) {
    const cplx<T> s = load_scale<T>(a.scale);
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cmul(v[j], s);
  }
  const T inv = T(1.0 / N);
  team_fft<T, N, R, -1>(v, t, view, tw);
  S::potential(v, a, t, col, step);
  team_fft<T, N, R, +1>(v, t, view, tw);
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = cscale(v[j], inv);

  if constexpr (CL) {
    // W + 2aD W + (10/3)a^2 D^2 W + (4/3)a^3 D^3 W, a = i dt gamma, built as
    // LW = W + (2/3)a D W; LW += a D LW; W += 2a D LW   (FFT.py:1145-1157)
    cplx<T> lw[R], d[R];
#pragma unroll
    for (int j = 0; j < R; ++j) d[j] = v[j];
    S::cl_operator(d, a, t, view, tw);
    const T c1 = T((2.0 / 3.0) * a.cl_a), c2 = T(a.cl_a), c3 = T(2.0 * a.cl_a);
#pragma unroll
    for (int j = 0; j < R; ++j) lw[j] = cadd(v[j], mul_i(d[j], c1));
#pragma unroll
    for (int j = 0; j < R; ++j) d[j] = lw[j];
    S::cl_operator(d, a, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) lw[j] = cadd(lw[j], mul_i(d[j], c2));
#pragma unroll
    for (int j = 0; j < R; ++j) d[j] = lw[j];
    S::cl_operator(d, a, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cadd(v[j], mul_i(d[j], c3));
  }

  if (a.colsums) {
    double acc[kNCol] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const double p = a.p_host[t + j * TT];
      const double w = double(v[j].x);
      acc[0] += w;
      acc[1] += double(v[j].y);
      acc[2] += p * w;
      acc[3] += p * p * w;
      acc[4] += w < 0.0 ? w : 0.0;
    }
    // lanes of one column inside a wave: c + C*t; fold t with xor shuffles
#pragma unroll
    for (int off = C; off < 64; off <<= 1) {
#pragma unroll
      for (int k = 0; k < kNCol; ++k) acc[k] += __shfl_xor(acc[k], off, 64);
    }
    constexpr int kWaves = (C * TT + 63) / 64;
    __syncthreads();
    double* red = reinterpret_cast<double*>(lds);
    const int wave = tid >> 6, lane = tid & 63;
    if (lane < C) {
#pragma unroll
      for (int k = 0; k < kNCol; ++k) red[(wave * C + lane) * kNCol + k] = acc[k];
    }
    __syncthreads();
    if (tid < C) {
      double s[kNCol] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int w = 0; w < kWaves; ++w) {
#pragma unroll
        for (int k = 0; k < kNCol; ++k) s[k] += red[(w * C + tid) * kNCol + k];
      }
#pragma unroll
      for (int k = 0; k < kNCol; ++k) a.colsums[size_t(blockIdx.x * C + tid) * kNCol + k] = s[k];
    }
  }

  if (a.merge) {
    team_fft<T, N, R, -1>(v, t, view, tw);
    S::potential(v, a, t, col, step + 1);
    team_fft<T, N, R, +1>(v, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cscale(v[j], inv);
  }

#pragma unroll
  for (int j = 0; j < R; ++j) at_off(W + size_t(j * TT) * X, boff) = v[j];
}

// ----------------------------------------------------------------------------- column sums only
template <typename T>
__global__ void __launch_bounds__(256) colsum_kernel(const void* Wv, int X, int P,
                                                     const double* p_host, double* colsums) {
  // 32 columns x 8 row groups per block
  __shared__ double red[8][32][kNCol];
  const cplx<T>* W = reinterpret_cast<const cplx<T>*>(Wv);
  const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int col = blockIdx.x * 32 + c;
  double acc[kNCol] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (col < X) {
    for (int i = g; i < P; i += 8) {
      const cplx<T> w = W[size_t(i) * X + col];
      const double p = p_host[i], wr = double(w.x);
      acc[0] += wr;
      acc[1] += double(w.y);
      acc[2] += p * wr;
      acc[3] += p * p * wr;
      acc[4] += wr < 0.0 ? wr : 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < kNCol; ++k) red[g][c][k] = acc[k];
  __syncthreads();
  if (g == 0 && col < X) {
#pragma unroll
    for (int k = 0; k < kNCol; ++k) {
      double s = 0.0;
      for (int q = 0; q < 8; ++q) s += red[q][c][k];
      colsums[size_t(col) * kNCol + k] = s;
    }
  }
}

// ----------------------------------------------------------------------------- reduce
__global__ void __launch_bounds__(256) reduce_kernel(ReduceArgs a) {
  constexpr int K = 12;
  __shared__ double red[K][256];
  __shared__ double scale_sh[2];
  const int tid = threadIdx.x;
  const int step = *a.step;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int j = tid; j < a.X; j += 256) {
    const double* cs = a.colsums + size_t(j) * kNCol;
    const double s0 = cs[0], s0i = cs[1], s1 = cs[2], s2 = cs[3], sn = cs[4];
    const double x = a.x_host[j], dv = a.dv_host[j], v0 = a.v0_host[j];
    acc[0] += s0;
    acc[1] += s0i;
    acc[2] += x * s0;
    acc[3] += x * x * s0;
    acc[4] += s1;
    acc[5] += s2;
    acc[6] += x * s1;
    acc[7] += dv * s0;
    acc[8] += dv * s1;
    acc[9] += x * dv * s0;
    acc[10] += s2 * a.inv_2m + v0 * s0;
    acc[11] += sn;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) red[k][tid] = acc[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
#pragma unroll
      for (int k = 0; k < K; ++k) red[k][tid] += red[k][tid + h];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double nr = a.dxdp * red[0][0], ni = a.dxdp * red[1][0];
    double f = 1.0, sr = 1.0, si = 0.0;
    if (a.renorm) {
      const double m2 = nr * nr + ni * ni;
      f = nr / m2;
      sr = nr / m2;
      si = -ni / m2;
    }
    scale_sh[0] = sr;
    scale_sh[1] = si;
    if (a.scale_out) {
      a.scale_out[0] = sr;
      a.scale_out[1] = si;
    }
    if (a.record) {
      double* o = a.obs + size_t(step - a.obs_base) * kNObs;
      const double d = a.dxdp * f;
      o[kObsX] = d * red[2][0];
      o[kObsX2] = d * red[3][0];
      o[kObsP] = d * red[4][0];
      o[kObsP2] = d * red[5][0];
      o[kObsXP] = d * red[6][0];
      o[kObsDV] = d * red[7][0];
      o[kObsPDV] = d * red[8][0];
      o[kObsXDV] = d * red[9][0];
      o[kObsH] = d * red[10][0];
      o[kObsNeg] = d * red[11][0];
      o[kObsNorm] = d * red[0][0];
      o[kObsRawNormRe] = nr;
      o[kObsRawNormIm] = ni;
    }
  }
  if (a.rho_out) {
    __syncthreads();
    const double sr = scale_sh[0], si = scale_sh[1];
    for (int j = tid; j < a.X; j += 256) {
      const double* cs = a.colsums + size_t(j) * kNCol;
      a.rho_out[j] = (cs[0] * sr - cs[1] * si) * a.rho_factor;
    }
  }
  if (tid == 0 && a.step_inc) {
    __threadfence();
    *a.step_inc = step + 1;
  }
}

// ----------------------------------------------------------------------------- small helpers
__global__ void set_int_kernel(int* p, int v) { *p = v; }
__global__ void set_scale_kernel(double* p, double re, double im) {
  p[0] = re;
  p[1] = im;
}

// ----------------------------------------------------------------------------- Gross-Pitaevskii
// G(theta, x) = Re[B-(theta,x) - B+(theta,x)],  B-/+ = iFFT_x[ FFT_x(rho) e^{-/+ i lambda theta/2} ]
//             = Re iFFT_x[ FFT_x(rho) * (-2i sin(lambda theta / 2)) ]
// with the host Lambda/Theta grids (phase_LambdaTheta, GPU_Wigner2D_FFT.py:594-595,
// MakeGrossPitaevskiiTerms :1687-1705).  One team per theta row recomputes FFT_x(rho).
template <int N, int R, int ROWS>
__global__ void __launch_bounds__(ROWS*(N / R))
    gp_shift_kernel(const double* __restrict__ rho, const double* __restrict__ lambda_host,
                    const double* __restrict__ theta_host, double* __restrict__ G, int P,
                    const void* twv) {
  constexpr int TT = N / R;
  constexpr int LSZ = Padded<N>::size;
  __shared__ cplx<double> lds[ROWS * LSZ];
  const int tid = threadIdx.x;
  const int team = tid / TT, t = tid % TT;
  const int row = blockIdx.x * ROWS + team;
  const cplx<double>* tw = reinterpret_cast<const cplx<double>*>(twv);
  LdsRow<double> view{lds + team * LSZ};
  cplx<double> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = {rho[t + j * TT], 0.0};
  team_fft<double, N, R, -1>(v, t, view, tw);
  const double half_theta = 0.5 * theta_host[row < P ? row : 0];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    double sn, cs;
    sincos_fp64(lambda_host[t + j * TT] * half_theta, &sn, &cs);
    const double f = 2.0 * sn;  // (-2i sin) * (a + ib) = 2 sin b - 2i sin a
    v[j] = {f * v[j].y, -f * v[j].x};
  }
  team_fft<double, N, R, +1>(v, t, view, tw);
  if (row < P) {
#pragma unroll
    for (int j = 0; j < R; ++j) G[size_t(row) * N + t + j * TT] = v[j].x * (1.0 / N);
  }
}

// rho(x) = factor * sum_b colpart[b][x];  obs ProbabilitySquare = dx * sum rho^2 (GP.py:1935-1937)
__global__ void __launch_bounds__(256)
    colpart_reduce_kernel(const double* __restrict__ colpart, int nblk, int X, double factor,
                          double* __restrict__ rho, double* obs, int obs_base, const int* step,
                          double dx) {
  __shared__ double red[256];
  double ps = 0.0;
  for (int x = threadIdx.x; x < X; x += 256) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += colpart[size_t(b) * X + x];
    s *= factor;
    rho[x] = s;
    ps += s * s;
  }
  red[threadIdx.x] = ps;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && obs) obs[size_t(*step - obs_base) * kNObs + kObsProbSq] = dx * red[0];
}

// ----------------------------------------------------------------------------- scale
template <typename T>
__global__ void __launch_bounds__(256) scale_kernel(void* Wv, size_t n, const double* s) {
  cplx<T>* W = reinterpret_cast<cplx<T>*>(Wv);
  const cplx<T> f = load_scale<T>(s);
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256)
    W[i] = cmul(W[i], f);
}

// ----------------------------------------------------------------------------- plain axis FFTs
// In-place batched transforms (cufftExecZ2Z/C2C on Plan_Z2Z_2D_Axis1 / Axis0 plans,
// cufft_wrapper.py:183-250), optionally scaled by 1/N (the callers' "*= 1./N",
// GPU_Wigner2D_FFT.py:790-801).
template <typename T, int N, int R, int ROWS, int DIR>
__global__ void __launch_bounds__(ROWS*(N / R))
    fft_rows_kernel(void* data, int rows, int normalize, const void* twv) {
  constexpr int TT = N / R;
  constexpr int LSZ = Padded<N>::size;
  __shared__ cplx<T> lds[ROWS * LSZ];
  const int tid = threadIdx.x;
  const int team = tid / TT, t = tid % TT;
  const int row = blockIdx.x * ROWS + team;
  const bool active = row < rows;
  cplx<T>* W = reinterpret_cast<cplx<T>*>(data) + size_t(active ? row : 0) * N;
  const cplx<T>* tw = reinterpret_cast<const cplx<T>*>(twv);
  LdsRow<T> view{lds + team * LSZ};
  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = active ? W[t + j * TT] : cplx<T>{T(0), T(0)};
  team_fft<T, N, R, DIR>(v, t, view, tw);
  const T f = normalize ? T(1.0 / N) : T(1);
  if (active) {
#pragma unroll
    for (int j = 0; j < R; ++j) W[t + j * TT] = cscale(v[j], f);
  }
}

template <typename T, int N, int R, int C, int DIR>
__global__ void __launch_bounds__(C*(N / R))
    fft_cols_kernel(void* data, int X, int normalize, const void* twv) {
  constexpr int TT = N / R;
  constexpr int LSZ = Padded<N>::size;
  __shared__ cplx<T> lds[C * LSZ];
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int col = blockIdx.x * C + c;
  cplx<T>* W = reinterpret_cast<cplx<T>*>(data);
  const cplx<T>* tw = reinterpret_cast<const cplx<T>*>(twv);
  LdsStrided<T, C> view{lds + c};
  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<T>));
  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = at_off(W + size_t(j * TT) * X, boff);
  team_fft<T, N, R, DIR>(v, t, view, tw);
  const T f = normalize ? T(1.0 / N) : T(1);
#pragma unroll
  for (int j = 0; j < R; ++j) at_off(W + size_t(j * TT) * X, boff) = cscale(v[j], f);
}

}  // namespace wh

// ============================================================================= launchers
namespace wh {
namespace {

// Stockham team geometry per (precision, N): radix 16, TT = N/16 lanes per transform.
constexpr int kR = 16;
template <int N>
constexpr int rows_per_block() {  // x sweep: teams per 256-lane block
  return (256 / (N / kR)) < 1 ? 1 : 256 / (N / kR);
}
// p sweep: columns per block, bounded by LDS (about 139 KiB for FP64 at N >= 512)
// and by 512 lanes per block (256 with the Caldeira-Leggett variant, whose three
// register arrays need the larger per-lane register budget).
template <typename T, int N, bool CL = false>
constexpr int cols_per_block() {
  constexpr int bytes = Padded<N>::size * int(sizeof(cplx<T>));
  constexpr int c = (144 * 1024) / bytes;
  constexpr int c2 = c >= 64 ? 64 : c >= 32 ? 32 : c >= 16 ? 16 : c >= 8 ? 8 : c >= 4 ? 4 : 2;
  constexpr int lanes = CL ? 256 : 512;
  constexpr int cmax = lanes / (N / 16) < 2 ? 2 : lanes / (N / 16);
  return c2 < cmax ? c2 : cmax;
}

template <typename T, int N>
hipError_t xsweep_n(const XSweepArgs& a, hipStream_t s) {
  constexpr int ROWS = rows_per_block<N>();
  constexpr int TT = N / kR;
  hipLaunchKernelGGL((xsweep_kernel<T, N, kR, ROWS>), dim3(a.P / ROWS), dim3(ROWS * TT), 0, s, a);
  return hipGetLastError();
}
template <typename T, int N>
hipError_t psweep_n(const PSweepArgs& a, hipStream_t s) {
  constexpr int TT = N / kR;
  if (a.cl) {
    constexpr int C = cols_per_block<T, N, true>();
    hipLaunchKernelGGL((psweep_kernel<T, N, kR, C, true>), dim3(a.X / C), dim3(C * TT), 0, s, a);
  } else {
    constexpr int C = cols_per_block<T, N, false>();
    hipLaunchKernelGGL((psweep_kernel<T, N, kR, C, false>), dim3(a.X / C), dim3(C * TT), 0, s, a);
  }
  return hipGetLastError();
}
template <typename T, int N, int DIR>
hipError_t fft_rows_n(void* d, int rows, int norm, const void* tw, hipStream_t s) {
  constexpr int ROWS = rows_per_block<N>();
  constexpr int TT = N / kR;
  hipLaunchKernelGGL((fft_rows_kernel<T, N, kR, ROWS, DIR>), dim3((rows + ROWS - 1) / ROWS),
                     dim3(ROWS * TT), 0, s, d, rows, norm, tw);
  return hipGetLastError();
}
template <typename T, int N, int DIR>
hipError_t fft_cols_n(void* d, int X, int norm, const void* tw, hipStream_t s) {
  constexpr int C = cols_per_block<T, N>();
  constexpr int TT = N / kR;
  hipLaunchKernelGGL((fft_cols_kernel<T, N, kR, C, DIR>), dim3(X / C), dim3(C * TT), 0, s, d, X,
                     norm, tw);
  return hipGetLastError();
}

#define WH_SIZE_SWITCH(N, CALL)      \
  switch (N) {                       \
    case 64: return CALL(64);        \
    case 128: return CALL(128);      \
    case 256: return CALL(256);      \
    case 512: return CALL(512);      \
    case 1024: return CALL(1024);    \
    case 2048: return CALL(2048);    \
    default: return hipErrorInvalidValue; \
  }

template <typename T>
hipError_t xsweep_t(const XSweepArgs& a, hipStream_t s) {
#define CALL(n) xsweep_n<T, n>(a, s)
  WH_SIZE_SWITCH(a.X, CALL)
#undef CALL
}
template <typename T>
hipError_t psweep_t(const PSweepArgs& a, hipStream_t s) {
#define CALL(n) psweep_n<T, n>(a, s)
  WH_SIZE_SWITCH(a.P, CALL)
#undef CALL
}
template <typename T, int DIR>
hipError_t fft_rows_t(int N, void* d, int rows, int norm, const void* tw, hipStream_t s) {
#define CALL(n) fft_rows_n<T, n, DIR>(d, rows, norm, tw, s)
  WH_SIZE_SWITCH(N, CALL)
#undef CALL
}
template <typename T, int DIR>
hipError_t fft_cols_t(int N, void* d, int X, int norm, const void* tw, hipStream_t s) {
#define CALL(n) fft_cols_n<T, n, DIR>(d, X, norm, tw, s)
  WH_SIZE_SWITCH(N, CALL)
#undef CALL
}

template <typename T, bool CL>
int cols_for(int N) {
  switch (N) {
    case 64: return cols_per_block<T, 64, CL>();
    case 128: return cols_per_block<T, 128, CL>();
    case 256: return cols_per_block<T, 256, CL>();
    case 512: return cols_per_block<T, 512, CL>();
    case 1024: return cols_per_block<T, 1024, CL>();
    case 2048: return cols_per_block<T, 2048, CL>();
    default: return 0;
  }
}
int rows_for(int N) {
  switch (N) {
    case 64: return rows_per_block<64>();
    case 128: return rows_per_block<128>();
    case 256: return rows_per_block<256>();
    case 512: return rows_per_block<512>();
    case 1024: return rows_per_block<1024>();
    case 2048: return rows_per_block<2048>();
    default: return 0;
  }
}

template <int N>
hipError_t gp_shift_n(const double* rho, const double* lh, const double* th, double* G, int P,
                      const void* tw, hipStream_t s) {
  constexpr int ROWS = rows_per_block<N>();
  constexpr int TT = N / kR;
  hipLaunchKernelGGL((gp_shift_kernel<N, kR, ROWS>), dim3((P + ROWS - 1) / ROWS), dim3(ROWS * TT), 0, s,
                     rho, lh, th, G, P, tw);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gp_shift(int X, int P, const double* rho, const double* lambda_host,
                           const double* theta_host, double* G, const void* twX64, hipStream_t s) {
#define CALL(n) gp_shift_n<n>(rho, lambda_host, theta_host, G, P, twX64, s)
  WH_SIZE_SWITCH(X, CALL)
#undef CALL
}
hipError_t launch_colpart_reduce(const double* colpart, int nblk, int X, double factor, double* rho,
                                 double* obs, int obs_base, const int* step, double dx,
                                 hipStream_t s) {
  hipLaunchKernelGGL(colpart_reduce_kernel, dim3(1), dim3(256), 0, s, colpart, nblk, X, factor, rho,
                     obs, obs_base, step, dx);
  return hipGetLastError();
}
hipError_t launch_set_int(int* p, int v, hipStream_t s) {
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, p, v);
  return hipGetLastError();
}
hipError_t launch_set_scale(double* p, double re, double im, hipStream_t s) {
  hipLaunchKernelGGL(set_scale_kernel, dim3(1), dim3(1), 0, s, p, re, im);
  return hipGetLastError();
}

bool size_supported(int n) {
  return n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 || n == 2048;
}
int xsweep_rows(int n) { return rows_for(n); }
int psweep_cols(int prec, int n, int cl) {
  if (cl) return prec == 0 ? cols_for<double, true>(n) : cols_for<float, true>(n);
  return prec == 0 ? cols_for<double, false>(n) : cols_for<float, false>(n);
}

hipError_t launch_xsweep(int prec, const XSweepArgs& a, hipStream_t s) {
  return prec == 0 ? xsweep_t<double>(a, s) : xsweep_t<float>(a, s);
}
hipError_t launch_psweep(int prec, const PSweepArgs& a, hipStream_t s) {
  return prec == 0 ? psweep_t<double>(a, s) : psweep_t<float>(a, s);
}
hipError_t launch_colsum(int prec, const void* W, int X, int P, const double* p_host,
                         double* colsums, hipStream_t s) {
  if (prec == 0)
    hipLaunchKernelGGL(colsum_kernel<double>, dim3((X + 31) / 32), dim3(256), 0, s, W, X, P, p_host,
                       colsums);
  else
    hipLaunchKernelGGL(colsum_kernel<float>, dim3((X + 31) / 32), dim3(256), 0, s, W, X, P, p_host,
                       colsums);
  return hipGetLastError();
}
hipError_t launch_reduce(const ReduceArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_scale(int prec, void* W, size_t n, const double* f, hipStream_t s) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (prec == 0)
    hipLaunchKernelGGL(scale_kernel<double>, dim3(unsigned(blocks)), dim3(256), 0, s, W, n, f);
  else
    hipLaunchKernelGGL(scale_kernel<float>, dim3(unsigned(blocks)), dim3(256), 0, s, W, n, f);
  return hipGetLastError();
}
hipError_t launch_fft_rows(int prec, int N, int dir, void* d, int rows, int norm, const void* tw,
                           hipStream_t s) {
  if (prec == 0)
    return dir < 0 ? fft_rows_t<double, -1>(N, d, rows, norm, tw, s)
                   : fft_rows_t<double, 1>(N, d, rows, norm, tw, s);
  return dir < 0 ? fft_rows_t<float, -1>(N, d, rows, norm, tw, s)
                 : fft_rows_t<float, 1>(N, d, rows, norm, tw, s);
}
hipError_t launch_fft_cols(int prec, int N, int dir, void* d, int X, int norm, const void* tw,
                           hipStream_t s) {
  if (prec == 0)
    return dir < 0 ? fft_cols_t<double, -1>(N, d, X, norm, tw, s)
                   : fft_cols_t<double, 1>(N, d, X, norm, tw, s);
  return dir < 0 ? fft_cols_t<float, -1>(N, d, X, norm, tw, s)
                 : fft_cols_t<float, 1>(N, d, X, norm, tw, s);
}

}  // namespace wh
