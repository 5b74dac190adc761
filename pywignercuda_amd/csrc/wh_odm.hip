// ODM Lindblad dissipator of GPU_Wigner2D_FFT (dampingFunction = 'ODM', order 1:
// W += L(W), Lindbladian_ODM_Order1 -> _Lindbladian_ODM, GPU_Wigner2D_FFT.py:1516-1634,
// 1907-1910; tables of SetA_ODM, :1196-1244).
//
// The reference sums six terms h_k(theta, x) F_p F_x^-1 [g_k(p, lambda) F_x W] (18 FFTs).
// With c+- = cos(x +- hbar theta/2), s+- = sin(x +- hbar theta/2) the products of its h_k
// split into functions of x alone and of theta alone:
//   c+c- = (cos 2x + cos hbar theta)/2,  s+s- = (cos hbar theta - cos 2x)/2,
//   c-s+ = (sin 2x + sin hbar theta)/2,  s-c+ = (sin 2x - sin hbar theta)/2,
// an x factor commutes with F_p and a theta factor with F_x, so
//   L(W)/(2 gamma dt) = cos2x/2 F_x^-1[A W^] + i s sin2x/2 F_x^-1[B W^] + F_x^-1[C W^]
//                     + F_x^-1 F_p^-1 [ cos(hbar theta)/2 F_p[E W^] + i s sin(hbar theta)/2 F_p[F W^] ]
// with W^ = F_x W, g+- = g(p +- hbar lambda/2), gs+- = sign(p +- hbar lambda/2) g+-,
//   A = g+g- - gs+gs-,  B = g+gs- - gs+g-,  C = -(g+^2 + g-^2)/2,
//   E = g+g- + gs+gs-,  F = g+gs- + gs+g-.
// Three sweeps: odm_x (rows: F_x, the three x terms, W += c Lx, E W^ and F W^ to scratch),
// odm_p (columns of the scratch: F_p, theta factors, F_p^-1) and odm_back (rows: W += F_x^-1).
#include <hip/hip_runtime.h>

#include "wh_fft.hpp"
#include "wh_kernels.hpp"
#include "wh_math.hpp"
#include "wh_sweeps.hpp"

namespace wh {

constexpr int kOdmR = 8;
constexpr int kOdmBlock = 256;

namespace {

template <typename U>
__device__ __forceinline__ cplx<double> wide(cplx<U> v) {
  return {double(v.x), double(v.y)};
}
template <typename U>
__device__ __forceinline__ cplx<U> nar(cplx<double> v) {
  return {U(v.x), U(v.y)};
}

}  // namespace

// weight, sgn: the channel's weight and sign (GPU_Wigner2D_ODM2 runs two channels,
// (weightLindbland1, +1) then (weightLindbland2, -1), GPU_Wigner2D_FFT.py:2173-2181)
template <typename S, typename T, int N, int R, int BLOCK>
__device__ __forceinline__ void odm_x_body(const Member& m, int tile, cplx<T>* lds, double weight, double sgn) {
  constexpr int TT = N / R, ROWS = BLOCK / TT;
  constexpr int LSZ = Padded<N, R>::size;
  const int tid = threadIdx.x, team = tid / TT, t = tid % TT;
  const int row = tile * ROWS + team;
  const size_t PX = size_t(m.P) * size_t(N);
  cplx<S>* __restrict__ W = reinterpret_cast<cplx<S>*>(m.W) + size_t(row) * N;
  cplx<T>* __restrict__ S1 = reinterpret_cast<cplx<T>*>(m.odm_s) + size_t(row) * N;
  cplx<T>* __restrict__ S2 = S1 + PX;
  const double* __restrict__ g = m.odm_g + size_t(row) * N;
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(m.twX_hi), reinterpret_cast<const cplx<T>*>(m.twX_lo)};
  LdsRow<T> view{lds + team * LSZ};

  cplx<T> f[R];
#pragma unroll
  for (int j = 0; j < R; ++j) f[j] = to_arith<T>(W[t + j * TT]);
  team_fft<T, N, R, -1>(f, t, view, tw);  // W^ = F_x W
  // theta terms' inputs E W^, F W^ (lambda picture)
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const size_t e = t + j * TT;
    const double gp = g[e], gm = g[PX + e], gsp = g[2 * PX + e], gsm = g[3 * PX + e];
    const cplx<double> w = wide(f[j]);
    const double E = gp * gm + gsp * gsm, F = gp * gsm + gsp * gm;
    S1[e] = nar<T>({E * w.x, E * w.y});
    S2[e] = nar<T>({F * w.x, F * w.y});
  }
  // x terms
  cplx<double> acc[R];
  const double inv = 1.0 / N;
#pragma unroll
  for (int term = 0; term < 3; ++term) {
    cplx<T> y[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const size_t e = t + j * TT;
      const double gp = g[e], gm = g[PX + e], gsp = g[2 * PX + e], gsm = g[3 * PX + e];
      const double k = term == 0 ? gp * gm - gsp * gsm : term == 1 ? gp * gsm - gsp * gm : -0.5 * (gp * gp + gm * gm);
      y[j] = nar<T>({k * double(f[j].x), k * double(f[j].y)});
    }
    team_fft<T, N, R, +1>(y, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int xi = t + j * TT;
      const cplx<double> u = wide(y[j]);
      if (term == 0) {
        const double h = 0.5 * m.odm_x[xi] * inv;  // cos 2x / 2
        acc[j] = {h * u.x, h * u.y};
      } else if (term == 1) {
        const double h = 0.5 * m.odm_sign * sgn * m.odm_x[N + xi] * inv;  // i s sin 2x / 2
        acc[j] = {acc[j].x - h * u.y, acc[j].y + h * u.x};
      } else {
        acc[j] = {acc[j].x + inv * u.x, acc[j].y + inv * u.y};
      }
    }
  }
  // W += c Lx (re-read the row: it stays in cache, and f no longer holds it)
  const double c = m.odm_c * weight;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const cplx<double> w = wide(W[t + j * TT]);
    W[t + j * TT] = nar<S>({w.x + c * acc[j].x, w.y + c * acc[j].y});
  }
}

template <typename S, typename T, int N, int R, int BLOCK>
__device__ __forceinline__ void odm_p_body(const Member& m, int tile, cplx<T>* lds, double weight, double sgn) {
  constexpr int TT = N / R;
  constexpr int C = BLOCK / TT;
  const int tid = threadIdx.x;
  const int cc = tid % C, t = tid / C;
  const int X = m.X;
  const int col = tile * C + cc;
  const size_t PX = size_t(N) * size_t(X);
  cplx<T>* __restrict__ S1 = reinterpret_cast<cplx<T>*>(m.odm_s);
  cplx<T>* __restrict__ S2 = S1 + PX;
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(m.twP_hi), reinterpret_cast<const cplx<T>*>(m.twP_lo)};
  LdsStrided<T, C> view{lds + cc};
  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<T>));
  cplx<T> a[R], b[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    a[j] = at_off(S1 + size_t(j * TT) * X, boff);
    b[j] = at_off(S2 + size_t(j * TT) * X, boff);
  }
  team_fft<T, N, R, -1>(a, t, view, tw);
  team_fft<T, N, R, -1>(b, t, view, tw);
  const double half_c = 0.5 * m.odm_c * weight, s = m.odm_sign * sgn;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int i = t + j * TT;
    const double ct = m.odm_th[i], st = s * m.odm_th[N + i];
    const cplx<double> ua = wide(a[j]), ub = wide(b[j]);
    // (c/2) [cos(hbar theta) a + i s sin(hbar theta) b] / P
    const double k = half_c / N;
    a[j] = nar<T>({k * (ct * ua.x - st * ub.y), k * (ct * ua.y + st * ub.x)});
  }
  team_fft<T, N, R, +1>(a, t, view, tw);
#pragma unroll
  for (int j = 0; j < R; ++j) at_off(S1 + size_t(j * TT) * X, boff) = a[j];
}

template <typename S, typename T, int N, int R, int BLOCK>
__device__ __forceinline__ void odm_back_body(const Member& m, int tile, cplx<T>* lds) {
  constexpr int TT = N / R, ROWS = BLOCK / TT;
  constexpr int LSZ = Padded<N, R>::size;
  const int tid = threadIdx.x, team = tid / TT, t = tid % TT;
  const int row = tile * ROWS + team;
  cplx<S>* __restrict__ W = reinterpret_cast<cplx<S>*>(m.W) + size_t(row) * N;
  const cplx<T>* __restrict__ S1 = reinterpret_cast<const cplx<T>*>(m.odm_s) + size_t(row) * N;
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(m.twX_hi), reinterpret_cast<const cplx<T>*>(m.twX_lo)};
  LdsRow<T> view{lds + team * LSZ};
  cplx<T> y[R];
#pragma unroll
  for (int j = 0; j < R; ++j) y[j] = S1[t + j * TT];
  team_fft<T, N, R, +1>(y, t, view, tw);
  const double inv = 1.0 / N;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const cplx<double> w = wide(W[t + j * TT]), u = wide(y[j]);
    W[t + j * TT] = nar<S>({w.x + inv * u.x, w.y + inv * u.y});
  }
}

template <typename S, typename T, int KIND>
__global__ void __launch_bounds__(kOdmBlock) odm_kernel(const Member* members, const unsigned* work, double weight,
                                                        double sgn) {
  __shared__ cplx<T> lds[lds_elems<kOdmBlock, kOdmR>()];
  const unsigned w = work[blockIdx.x];
  const Member& m = members[w >> 16];
  const int tile = int(w & 0xffffu);
  constexpr int R = kOdmR, BLOCK = kOdmBlock;
  if constexpr (KIND == 0) {
#define BODY(n) odm_x_body<S, T, n, R, BLOCK>(m, tile, lds, weight, sgn)
    WH_DISPATCH(m.X, BODY)
#undef BODY
  } else if constexpr (KIND == 1) {
#define BODY(n) odm_p_body<S, T, n, R, BLOCK>(m, tile, lds, weight, sgn)
    WH_DISPATCH(m.P, BODY)
#undef BODY
  } else {
#define BODY(n) odm_back_body<S, T, n, R, BLOCK>(m, tile, lds)
    WH_DISPATCH(m.X, BODY)
#undef BODY
  }
}

// Tiles of the ODM sweeps for one member: rows per x tile, columns per p tile.
int odm_rows_per_tile(int X) { return kOdmBlock / (X / kOdmR); }
int odm_cols_per_tile(int P) { return kOdmBlock / (P / kOdmR); }

hipError_t launch_odm(int prec, const Member* members, const unsigned* work_rows, int nrows,
                      const unsigned* work_cols, int ncols, double weight, double sgn, hipStream_t s) {
  const dim3 b(kOdmBlock);
  if (prec == 0) {
    hipLaunchKernelGGL((odm_kernel<double, double, 0>), dim3(nrows), b, 0, s, members, work_rows, weight, sgn);
    hipLaunchKernelGGL((odm_kernel<double, double, 1>), dim3(ncols), b, 0, s, members, work_cols, weight, sgn);
    hipLaunchKernelGGL((odm_kernel<double, double, 2>), dim3(nrows), b, 0, s, members, work_rows, weight, sgn);
  } else if (prec == 1) {
    hipLaunchKernelGGL((odm_kernel<float, float, 0>), dim3(nrows), b, 0, s, members, work_rows, weight, sgn);
    hipLaunchKernelGGL((odm_kernel<float, float, 1>), dim3(ncols), b, 0, s, members, work_cols, weight, sgn);
    hipLaunchKernelGGL((odm_kernel<float, float, 2>), dim3(nrows), b, 0, s, members, work_rows, weight, sgn);
  } else {
    hipLaunchKernelGGL((odm_kernel<float, double, 0>), dim3(nrows), b, 0, s, members, work_rows, weight, sgn);
    hipLaunchKernelGGL((odm_kernel<float, double, 1>), dim3(ncols), b, 0, s, members, work_cols, weight, sgn);
    hipLaunchKernelGGL((odm_kernel<float, double, 2>), dim3(nrows), b, 0, s, members, work_rows, weight, sgn);
  }
  return hipGetLastError();
}

}  // namespace wh
