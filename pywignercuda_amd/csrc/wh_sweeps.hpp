// Fused sweeps of the Wigner split-operator step, ensemble form.
//
//   xsweep  one team per row p:  FFT_x -> kinetic multiplier (p, lambda) -> iFFT_x / X
//           (Fourier_X_To_Lambda_GPU, expPLambda, Fourier_Lambda_To_X_GPU:
//            GPU_Wigner2D_FFT.py:131-158, 788-793, 1885-1888; expPLambdaKinetic,
//            GPU_Wigner2D_GPitaevskii.py:117-158, 1924-1929)
//   psweep  one team per column x of a tile:  FFT_p -> potential multiplier (theta, x)
//           [* Gross-Pitaevskii phase] -> iFFT_p / P -> [Caldeira-Leggett order-3
//           composite] -> [column sums] -> [next Strang half step]
//           (expPotential GPU_Wigner2D_FFT.py:80-126, 795-801, 1891-1913;
//            expPotential_GrossPitaevskii :479-519; theta_fp + CaldeiraDissipatorOrder3
//            :399-431, 1135-1157; GPU_Wigner2D_GPitaevskii.py:60-86, 466-496, 1911-1970)
//   gpshift G(theta, x) = Re iFFT_x[ FFT_x(rho) (-2i sin(lambda theta/2)) ]
//           (MakeGrossPitaevskiiTerms, GPU_Wigner2D_FFT.py:1687-1705)
//
// Each sweep reads and writes the state once.  Phase multipliers are evaluated
// and applied in FP64 for both state precisions (an FP32 multiplier would be a
// fixed perturbation of the step operator and its error would grow linearly).
#pragma once
#include <hip/hip_runtime.h>

#include "wh_fft.hpp"
#include "wh_kernels.hpp"
#include "wh_math.hpp"

namespace wh {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ double poly(const double* __restrict__ c, int deg, double x) {
  double r = c[deg];
  for (int n = deg - 1; n >= 0; --n) r = fma(r, x, c[n]);
  return r;
}

// kernel-side FFT-ordered coordinate index: ((k + n/2) mod n) - n/2
__device__ __forceinline__ int fidx(int k, int n) { return ((k + (n >> 1)) & (n - 1)) - (n >> 1); }

// v * m with an FP64 multiplier m; FP32 states are widened for the product.
__device__ __forceinline__ cplx<double> apply(cplx<double> v, cplx<double> m) { return cmul(v, m); }
__device__ __forceinline__ cplx<float> apply(cplx<float> v, cplx<double> m) {
  const double x = v.x, y = v.y;
  return {float(x * m.x - y * m.y), float(x * m.y + y * m.x)};
}
// r * exp(-i phi)
__device__ __forceinline__ cplx<double> expmi(double r, double phi) {
  double s, c;
  sincos_fp64(phi, &s, &c);
  return {r * c, -r * s};
}

template <typename T>
__device__ __forceinline__ cplx<T> scale_load(cplx<T> v, const double* s) {
  return apply(v, cplx<double>{s[0], s[1]});
}

// element of a uniform row base at a per-lane 32-bit byte offset (SGPR base +
// VGPR offset addressing: one offset register per lane for all strided elements)
template <typename T>
__device__ __forceinline__ cplx<T>& at_off(cplx<T>* base, unsigned boff) {
  return *reinterpret_cast<cplx<T>*>(reinterpret_cast<char*>(base) + boff);
}

template <typename T>
__device__ __forceinline__ cplx<T> mul_i(cplx<T> a, T s) { return {-s * a.y, s * a.x}; }

// state storage S <-> arithmetic T (complex64 states with FP64 arithmetic convert at the
// sweep's load and store; every other combination is the identity)
template <typename T, typename S>
__device__ __forceinline__ cplx<T> to_arith(cplx<S> v) {
  return {T(v.x), T(v.y)};
}

// Largest LDS footprint over the supported sizes for a (block, radix) variant:
// ROWS*Padded(N) = (BLOCK/(N/R)) * (N + N/R + 1) elements.
template <int BLOCK, int R>
constexpr int lds_elems() {
  int m = 0;
  for (int N = 64; N <= 2048; N *= 2) {
    if (N / R > BLOCK || N < R) continue;
    const int rows = BLOCK / (N / R);
    const int e = rows * (N + N / R + 1);
    m = e > m ? e : m;
  }
  return m;
}

template <int N, int R, int BLOCK>
constexpr bool size_ok() {
  return N >= R * 4 && N / R <= BLOCK && BLOCK / (N / R) <= 64;
}

// ------------------------------------------------------------------ x sweep
// Kinetic multiplier of element m (lambda index) in row p, linear form:
// exp(-i a m_s), a = dt*dlambda*(k_c0 + k_c1 p), m_s = t + js*TT with
// js = j (j < R/2) or j - R; computed as E_t * Z^js, E_t = exp(-i a t),
// Z = exp(-i a TT): two sincos per lane instead of R.
template <typename S, typename T, int N, int R, int BLOCK>
__device__ __forceinline__ void xsweep_body(const SweepArgs& a, const Member& m, int tile,
                                            cplx<T>* lds) {
  constexpr int TT = N / R, ROWS = BLOCK / TT;
  constexpr int LSZ = Padded<N, R>::size;
  const int tid = threadIdx.x, team = tid / TT, t = tid % TT;
  const int row = tile * ROWS + team;
  cplx<S>* __restrict__ W = reinterpret_cast<cplx<S>*>(m.W) + size_t(row) * N;
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(m.twX_hi), reinterpret_cast<const cplx<T>*>(m.twX_lo)};
  LdsRow<T> view{lds + team * LSZ};

  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = to_arith<T>(W[t + j * TT]);
  if (a.flags & kScaleOnLoad) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = scale_load(v[j], m.scale);
  }
  team_fft<T, N, R, -1>(v, t, view, tw);

  const double p = m.dp * double(fidx(row, m.P));
  const double sgn = (a.flags & kConj) ? -1.0 : 1.0;  // adjoint: conj(r e^{-i phi}) = r e^{+i phi}
  if (a.kin_linear) {
    const double ak = sgn * a.dt * m.dlambda * (a.k_c0 + a.k_c1 * p);
    const cplx<double> Et = expmi(1.0, ak * double(t));
    const cplx<double> Z1 = expmi(1.0, ak * double(TT));
    cplx<double> Zp[R / 2 + 1];  // Z^0 .. Z^(R/2)
    Zp[0] = {1.0, 0.0};
    Zp[1] = Z1;
#pragma unroll
    for (int k = 2; k <= R / 2; ++k) Zp[k] = cmul(Zp[k / 2], Zp[k - k / 2]);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      cplx<double> z = j < R / 2 ? Zp[j] : cplx<double>{Zp[R - j].x, -Zp[R - j].y};
      cplx<double> mj = cmul(Et, z);
      if (m.lambda_decay) {
        const double r = m.lambda_decay[t + j * TT];
        mj = {mj.x * r, mj.y * r};
      }
      v[j] = apply(v[j], mj);
    }
  } else {
    const double hh = a.hbar_half;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int mm = t + j * TT;
      const double lam = m.dlambda * double(fidx(mm, N));
      const double phi = sgn * a.dt * (poly(a.kcoef, a.kdeg, p + hh * lam) - poly(a.kcoef, a.kdeg, p - hh * lam)) * a.inv_hbar;
      const double r = m.lambda_decay ? m.lambda_decay[mm] : 1.0;
      v[j] = apply(v[j], expmi(r, phi));
    }
  }
  team_fft<T, N, R, +1>(v, t, view, tw);
  const T inv = T(1.0 / N);
#pragma unroll
  for (int j = 0; j < R; ++j) {
    v[j] = cscale(v[j], inv);
    W[t + j * TT] = to_arith<S>(v[j]);
  }

  if (a.flags & kColPart) {
    // partial column sums of Re W over this tile's rows (Gross-Pitaevskii density after K)
    __syncthreads();
    double* red = reinterpret_cast<double*>(lds);
#pragma unroll
    for (int j = 0; j < R; ++j) red[team * N + t + j * TT] = double(v[j].x);
    __syncthreads();
    for (int x = tid; x < N; x += BLOCK) {
      double s = 0.0;
      for (int q = 0; q < ROWS; ++q) s += red[q * N + x];
      m.colpart[size_t(tile) * N + x] = s;
    }
  }
}

// ------------------------------------------------------------------ p sweep
template <typename T, int N, int R, int BLOCK>
struct PSweep {
  static constexpr int TT = N / R;
  static constexpr int C = BLOCK / TT;  // columns per tile
  static constexpr int LSZ = Padded<N, R>::size;
  using View = LdsStrided<T, C>;

  // potential step multiplier of time index `step` (expPotential; GP phase optional)
  __device__ __forceinline__ static void potential(cplx<T> (&v)[R], const SweepArgs& a,
                                                   const Member& m, int t_in, int col_in, int step) {
    // opaque lane indices: recomputing theta and x -/+ hbar theta/2 per call is cheaper
    // than keeping them live across the transforms between two calls
    const int t = opaque_v(t_in), col = opaque_v(col_in);
    const double x = m.dx * double(fidx(col, m.X));
    const int row = step < m.pot_rows ? step : m.pot_rows - 1;
    const double* __restrict__ cf = m.pot + size_t(row) * (kMaxDeg + 1);
    const double mask = m.xmask ? m.xmask[col] : 1.0;
    const bool gp = (a.flags & kGP) != 0;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = t + j * TT;
      const double th = m.dtheta * double(fidx(i, N));
      const double u = a.hbar_half * th;
      double phi = m.pot_deg >= 0 ? a.vphase * (poly(cf, m.pot_deg, x - u) - poly(cf, m.pot_deg, x + u))
                                  : a.vphase * m.pot_dv[size_t(i) * m.X + col];
      if (gp) phi -= a.gp_phase * m.G[size_t(i) * m.X + col];
      const double r = (m.theta_decay ? m.theta_decay[i] : 1.0) * mask;
      v[j] = apply(v[j], expmi(r, phi));
    }
  }

  // D(Y) = F_p^{-1}[ theta * F_p[ sigma(p) Y ] ]   (Theta_fp_Damping, FFT.py:1135-1143)
  __device__ __forceinline__ static void cl_operator(cplx<T> (&y)[R], const Member& m, int t,
                                                     View view, const Twiddles<T>& tw) {
#pragma unroll
    for (int j = 0; j < R; ++j) y[j] = cscale(y[j], T(m.sigma_p[t + j * TT]));
    team_fft<T, N, R, -1>(y, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) y[j] = cscale(y[j], T(m.theta_host[t + j * TT]));
    team_fft<T, N, R, +1>(y, t, view, tw);
    const T inv = T(1.0 / N);
#pragma unroll
    for (int j = 0; j < R; ++j) y[j] = cscale(y[j], inv);
  }
};

template <typename S, typename T, int N, int R, int BLOCK, bool CL>
__device__ __forceinline__ void psweep_body(const SweepArgs& a, const Member& m, int tile,
                                            cplx<T>* lds) {
  using PS = PSweep<T, N, R, BLOCK>;
  constexpr int TT = PS::TT, C = PS::C;
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int X = m.X;
  const int col = tile * C + c;
  cplx<S>* __restrict__ W = reinterpret_cast<cplx<S>*>(m.W);
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(m.twP_hi), reinterpret_cast<const cplx<T>*>(m.twP_lo)};
  typename PS::View view{lds + c};
  const int step = *a.step + a.step_add;

  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<S>));
  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = to_arith<T>(at_off(W + size_t(j * TT) * X, boff));
  if (a.flags & kScaleOnLoad) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = scale_load(v[j], m.scale);
  }
  const T inv = T(1.0 / N);
  team_fft<T, N, R, -1>(v, t, view, tw);
  PS::potential(v, a, m, t, col, step);
  team_fft<T, N, R, +1>(v, t, view, tw);
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = cscale(v[j], inv);

  if constexpr (CL) {
    // W + 2aD W + (10/3)a^2 D^2 W + (4/3)a^3 D^3 W, a = i dt gamma, built as
    // LW = W + (2/3)a D W; LW += a D LW; W += 2a D LW   (FFT.py:1145-1157)
    cplx<T> lw[R], d[R];
#pragma unroll
    for (int j = 0; j < R; ++j) d[j] = v[j];
    PS::cl_operator(d, m, t, view, tw);
    const T c1 = T((2.0 / 3.0) * a.cl_a), c2 = T(a.cl_a), c3 = T(2.0 * a.cl_a);
#pragma unroll
    for (int j = 0; j < R; ++j) lw[j] = cadd(v[j], mul_i(d[j], c1));
#pragma unroll
    for (int j = 0; j < R; ++j) d[j] = lw[j];
    PS::cl_operator(d, m, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) lw[j] = cadd(lw[j], mul_i(d[j], c2));
#pragma unroll
    for (int j = 0; j < R; ++j) d[j] = lw[j];
    PS::cl_operator(d, m, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cadd(v[j], mul_i(d[j], c3));
  }

  if (a.flags & kColSums) {
    double acc[kNCol] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const double p = m.p_host[t + j * TT];
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
    constexpr int kWaves = (BLOCK + 63) / 64;
    constexpr int kLanes = C < 64 ? C : 64;  // distinct columns per wave
    __syncthreads();
    double* red = reinterpret_cast<double*>(lds);
    const int wave = tid >> 6, lane = tid & 63;
    if (lane < kLanes) {
#pragma unroll
      for (int k = 0; k < kNCol; ++k) red[(wave * kLanes + lane) * kNCol + k] = acc[k];
    }
    __syncthreads();
    if (tid < C) {
      // column tid lives in waves w with (w*64) % C == tid - lane... (C <= 64: every wave holds all C)
      double s[kNCol] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int w = 0; w < kWaves; ++w) {
#pragma unroll
        for (int k = 0; k < kNCol; ++k) s[k] += red[(w * kLanes + tid) * kNCol + k];
      }
#pragma unroll
      for (int k = 0; k < kNCol; ++k) m.colsums[size_t(tile * C + tid) * kNCol + k] = s[k];
    }
  }

  if (a.flags & kMerge) {
    team_fft<T, N, R, -1>(v, t, view, tw);
    PS::potential(v, a, m, t, col, step + 1);
    team_fft<T, N, R, +1>(v, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cscale(v[j], inv);
  }

#pragma unroll
  for (int j = 0; j < R; ++j) at_off(W + size_t(j * TT) * X, boff) = to_arith<S>(v[j]);
}

// ------------------------------------------------------------------ Gross-Pitaevskii shift
template <int N, int R, int BLOCK>
__device__ __forceinline__ void gpshift_body(const Member& m, int tile, cplx<double>* lds) {
  constexpr int TT = N / R, ROWS = BLOCK / TT;
  constexpr int LSZ = Padded<N, R>::size;
  const int tid = threadIdx.x, team = tid / TT, t = tid % TT;
  const int row = tile * ROWS + team;
  const Twiddles<double> tw{reinterpret_cast<const cplx<double>*>(m.twX64), nullptr};
  LdsRow<double> view{lds + team * LSZ};
  cplx<double> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = {m.rho[t + j * TT], 0.0};
  team_fft<double, N, R, -1>(v, t, view, tw);
  const double half_theta = 0.5 * m.theta_host[row];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    double sn, cs;
    sincos_fp64(m.lambda_host[t + j * TT] * half_theta, &sn, &cs);
    const double f = 2.0 * sn;  // (-2i sin) * (a + ib) = 2 sin b - 2i sin a
    v[j] = {f * v[j].y, -f * v[j].x};
  }
  team_fft<double, N, R, +1>(v, t, view, tw);
#pragma unroll
  for (int j = 0; j < R; ++j) m.G[size_t(row) * N + t + j * TT] = v[j].x * (1.0 / N);
}

// ------------------------------------------------------------------ kernels
#define WH_DISPATCH(NVAR, BODY)                                               \
  switch (NVAR) {                                                             \
    case 128: if constexpr (size_ok<128, R, BLOCK>()) { BODY(128); } break;   \
    case 256: if constexpr (size_ok<256, R, BLOCK>()) { BODY(256); } break;   \
    case 512: if constexpr (size_ok<512, R, BLOCK>()) { BODY(512); } break;   \
    case 1024: if constexpr (size_ok<1024, R, BLOCK>()) { BODY(1024); } break; \
    case 2048: if constexpr (size_ok<2048, R, BLOCK>()) { BODY(2048); } break; \
    default: break;                                                           \
  }

template <typename S, typename T, int R, int BLOCK>
__global__ void __launch_bounds__(BLOCK) xsweep_kernel(SweepArgs a) {
  __shared__ cplx<T> lds[lds_elems<BLOCK, R>()];
  const unsigned w = a.work[blockIdx.x];
  const Member& m = a.members[w >> 16];
  const int tile = int(w & 0xffffu);
#define BODY(n) xsweep_body<S, T, n, R, BLOCK>(a, m, tile, lds)
  WH_DISPATCH(m.X, BODY)
#undef BODY
}

template <typename S, typename T, int R, int BLOCK, bool CL>
__global__ void __launch_bounds__(BLOCK) psweep_kernel(SweepArgs a) {
  __shared__ cplx<T> lds[lds_elems<BLOCK, R>()];
  const unsigned w = a.work[blockIdx.x];
  const Member& m = a.members[w >> 16];
  const int tile = int(w & 0xffffu);
#define BODY(n) psweep_body<S, T, n, R, BLOCK, CL>(a, m, tile, lds)
  WH_DISPATCH(m.P, BODY)
#undef BODY
}

template <int R, int BLOCK>
__global__ void __launch_bounds__(BLOCK) gpshift_kernel(const Member* members, const unsigned* work) {
  __shared__ cplx<double> lds[lds_elems<BLOCK, R>()];
  const unsigned w = work[blockIdx.x];
  const Member& m = members[w >> 16];
  const int tile = int(w & 0xffffu);
#define BODY(n) gpshift_body<n, R, BLOCK>(m, tile, lds)
  WH_DISPATCH(m.X, BODY)
#undef BODY
}

}  // namespace wh
