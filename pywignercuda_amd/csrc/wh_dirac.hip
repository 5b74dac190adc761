// Wigner-Dirac 4x4 split step (GPU_WignerDirac2D_4x4.Run, GPU_WignerDirac2D_4x4.py:3409-3690),
// two fused sweeps per step:
//
//   dirac_xsweep  one team per row p and plane group:  FFT_x of the group's 4 planes ->
//                 W <- U(p + lambda/2) W U(p - lambda/2)^dagger -> iFFT_x / X
//                 (Fourier_4X4_X_To_Lambda_GPU, DiracPropagator_P_plus_Lambda /
//                  P_minus_Lambda, Fourier_4X4_Lambda_To_X_GPU; :159-351, 2622-2676, 3514-3558).
//                 The kinetic factor only couples spinor rows {1,4}, {2,3} and columns {1,4},
//                 {2,3}, so the 16 planes fall into 4 closed groups of 4 (e.g. W11 W14 W41 W44)
//                 and a team keeps one group in registers.  The lambda = 0 bin of groups
//                 {11,14,41,44} and {22,23,32,33} gives the antiparticle population of the
//                 loaded state (AntiParticlePopulation, :2436-2466: the filtered trace summed
//                 over x is Tr[Lambda_-(p) sum_x W(x, p)]).
//   dirac_psweep  one team per column x, all 16 planes in turn:  FFT_p -> e^{-dt D theta^2}
//                 e^{-i dt (V0(x - theta/2) - V0(x + theta/2))} times the constant mass phase of
//                 the plane -> iFFT_p / P -> [Caldeira-Leggett order 2] -> edge absorber ->
//                 column sums for the observables and the trace norm
//                 (Fourier_4X4_P_To_Theta_GPU, DiracPropagator_X_minus/plus_Theta for
//                  A1 = A2 = A3 = 0, Fourier_4X4_Theta_To_P_GPU, CaldeiraDissipatorOrder2_4x4,
//                  AbsorbBoundary_x; :357-670, 3011-3058, 111-155, 3560-3683).
//                 With A = 0 the potential factors are diagonal: P-(y) = e^{-iF} on spinor rows
//                 1, 2 and e^{+iF} on rows 3, 4 (F = mass_half c^2 dt), so plane W_ij picks up
//                 e^{-2iF} (i in {1,2}, j in {3,4}), e^{+2iF} (i in {3,4}, j in {1,2}) or 1.
//   dirac_reduce  one block: observables of the normalised state, trace norm -> pending
//                 factor 1/norm applied by the next x sweep's load (Wigner_4X4_Normalize,
//                 :2740-2775 and the reductions of :3429-3492).
#include <hip/hip_runtime.h>

#include "wh_dirac.hpp"
#include "wh_fft.hpp"
#include "wh_kernels.hpp"
#include "wh_math.hpp"
#include "wh_sweeps.hpp"

namespace wh {

namespace {

// plane groups closed under the kinetic factor, ordered [r1 c1, r1 c2, r2 c1, r2 c2]
__constant__ int kDiracGroups[4][4] = {{0, 3, 12, 15}, {1, 2, 13, 14}, {4, 7, 8, 11}, {5, 6, 9, 10}};

__device__ __forceinline__ cplx<double> widen(cplx<double> v) { return v; }
__device__ __forceinline__ cplx<double> widen(cplx<float> v) { return {double(v.x), double(v.y)}; }
template <typename T>
__device__ __forceinline__ cplx<T> narrow(cplx<double> v) { return {T(v.x), T(v.y)}; }

// U(q) = cos(c dt s) - i sin(c dt s) (mh_c beta + q alpha_1) / s on one {r1, r2} block:
// [[a, b], [b, d]] with a = C - i mh_c S/s, d = C + i mh_c S/s, b = -i q S/s; sgn = -1 gives
// the dt -> -dt factor of P_minus_Lambda (= U(q)^dagger).
struct Blk {
  cplx<double> a, b, d;
};
// mass_on = 0: the KvN form (GPU_WignerDiracKvN2D_4x4.py:183-189), whose diagonal is cos(c dt s)
// alone (s keeps the mass).
__device__ __forceinline__ Blk kinetic_block(double q, double mh_c, double c_dt, double sgn, double mass_on = 1.0) {
  const double s = sqrt(mh_c * mh_c + q * q);
  double sn, cs;
  sincos_fp64(c_dt * s, &sn, &cs);
  sn *= sgn;
  const double f = sn / s;
  const double m = mass_on * mh_c * f;
  return {{cs, -m}, {0.0, -q * f}, {cs, m}};
}

// Positive-energy projector Lambda_+(q) = 1/2 (1 + (m c^2 beta + c q alpha_1) / E) on one
// {r1, r2} block (DiracElectronProjector / Filter_Kernel with sign +1, :674-760).
__device__ __forceinline__ Blk projector_block(double q, double mc2, double c) {
  const double e2 = 2.0 * sqrt(mc2 * mc2 + c * c * q * q);
  return {{0.5 + mc2 / e2, 0.0}, {c * q / e2, 0.0}, {0.5 - mc2 / e2, 0.0}};
}

}  // namespace

// ------------------------------------------------------------------ x sweep
// FILTER: the antiparticle step filter instead of the kinetic factor (FilterAntiParticles,
// :2500-2540, applied at :3636-3678): W <- Lambda_+(q1) W Lambda_+(q2) in the (p, lambda)
// picture, q1,2 = p +/- lambda/2 (q1 = q2 = p in the KvN form, whose filter sets lambda = 0,
// GPU_WignerDiracKvN2D_4x4.py:775), then the edge absorber that follows it; no pending factor.
template <typename S, typename T, int N, int R, int BLOCK, bool FILTER = false>
__device__ __forceinline__ void dirac_xsweep_body(const DiracPlan& d, int group, int tile, cplx<T>* lds) {
  constexpr int TT = N / R, ROWS = BLOCK / TT;
  constexpr int LSZ = Padded<N, R>::size;
  const int tid = threadIdx.x, team = tid / TT, t = tid % TT;
  const int row = tile * ROWS + team;
  const size_t plane = size_t(d.P) * size_t(d.X);
  const gptr<cplx<S>> W = gmem(reinterpret_cast<cplx<S>*>(d.W)) + size_t(row) * N;
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(d.twX_hi), reinterpret_cast<const cplx<T>*>(d.twX_lo)};
  LdsRow<T> view{lds + team * LSZ};
  // a team inside one wave (radix 8 up to N = 512): wave-local exchanges
  using Sync = std::conditional_t<(TT <= 64), WaveSync, BlockSync>;
  const TwMem<T> twm{tw};
  const double s = FILTER ? 1.0 : *gmem(d.scale);

  cplx<T> v[4][R];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const gptr<cplx<S>> src = W + size_t(kDiracGroups[group][q]) * plane;
#pragma unroll
    for (int j = 0; j < R; ++j) v[q][j] = to_arith<T>(gload(src + t + j * TT));
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[q][j] = cscale(v[q][j], T(s));
    team_fft<T, N, R, -1, LdsRow<T>, Sync>(v[q], t, view, twm);
  }

  const double p = d.dp * double(fidx(row, d.P));
  if (!FILTER && d.antiparticle && (group == 0 || group == 3) && t == 0) {
    // lambda = 0 bin (element j = 0 of lane 0): Tr[Lambda_-(p) W^(p, 0)] over this group's planes,
    // Lambda_- = [[aa, bb], [bb, dd]] on each block (DiracElectronProjector, sign -1)
    const double e2 = 2.0 * sqrt(d.proj_mc2 * d.proj_mc2 + d.proj_c * d.proj_c * p * p);
    const double aa = 0.5 - d.proj_mc2 / e2, bb = -d.proj_c * p / e2, dd = 0.5 + d.proj_mc2 / e2;
    const double c = aa * double(v[0][0].x) + bb * (double(v[1][0].x) + double(v[2][0].x)) + dd * double(v[3][0].x);
    gmem(d.rowpart)[(group == 0 ? 0 : 1) * d.P + row] = c;
  }

#pragma unroll
  for (int j = 0; j < R; ++j) {
    const double lam = d.dlambda * double(fidx(t + j * TT, N));
    Blk K, Rt;
    if constexpr (FILTER) {
      const double h = d.kvn ? 0.0 : 0.5 * lam;
      K = projector_block(p + h, d.proj_mc2, d.proj_c);
      Rt = projector_block(p - h, d.proj_mc2, d.proj_c);
    } else if (d.kvn) {
      // KvN kinetic factors at q = +/- lambda/2 (GPU_WignerDiracKvN2D_4x4.py:183-189, 283-289)
      K = kinetic_block(0.5 * lam, d.mh_c, d.c_dt, 1.0, 0.0);
      Rt = kinetic_block(-0.5 * lam, d.mh_c, d.c_dt, -1.0, 0.0);
    } else {
      K = kinetic_block(p + 0.5 * lam, d.mh_c, d.c_dt, 1.0);
      Rt = kinetic_block(p - 0.5 * lam, d.mh_c, d.c_dt, -1.0);
    }
    const cplx<double> w0 = widen(v[0][j]), w1 = widen(v[1][j]), w2 = widen(v[2][j]), w3 = widen(v[3][j]);
    // K B
    const cplx<double> x0 = cadd(cmul(K.a, w0), cmul(K.b, w2));
    const cplx<double> x1 = cadd(cmul(K.a, w1), cmul(K.b, w3));
    const cplx<double> x2 = cadd(cmul(K.b, w0), cmul(K.d, w2));
    const cplx<double> x3 = cadd(cmul(K.b, w1), cmul(K.d, w3));
    // (K B) Rt
    v[0][j] = narrow<T>(cadd(cmul(x0, Rt.a), cmul(x1, Rt.b)));
    v[1][j] = narrow<T>(cadd(cmul(x0, Rt.b), cmul(x1, Rt.d)));
    v[2][j] = narrow<T>(cadd(cmul(x2, Rt.a), cmul(x3, Rt.b)));
    v[3][j] = narrow<T>(cadd(cmul(x2, Rt.b), cmul(x3, Rt.d)));
  }

  const T inv = T(1.0 / N);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    team_fft<T, N, R, +1, LdsRow<T>, Sync>(v[q], t, view, twm);
    const gptr<cplx<S>> dst = W + size_t(kDiracGroups[group][q]) * plane;
    if constexpr (FILTER) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const T f = d.fmask ? T(gmem(d.fmask)[t + j * TT] * (1.0 / N)) : inv;
        store_stream(dst + t + j * TT, to_arith<S>(cscale(v[q][j], f)));
      }
    } else {
#pragma unroll
      for (int j = 0; j < R; ++j) store_stream(dst + t + j * TT, to_arith<S>(cscale(v[q][j], inv)));
    }
  }
}

// ------------------------------------------------------------------ p sweep
// Per-lane accumulation of the column sums of one plane.
template <typename T, int R, int TT>
__device__ __forceinline__ void dirac_accumulate(const DiracPlan& d, int k, const cplx<T> (&v)[R], int t,
                                                 double (&acc)[kDNCol], double (&trre)[R]) {
  const bool diag = (k % 5) == 0;
  const bool alpha = k == 3 || k == 6 || k == 9 || k == 12;
  if (!diag && !alpha) return;
  const double sb = k < 8 ? 1.0 : -1.0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int i = t + j * TT;
    const double p = d.dp * double(fidx(i, d.P));
    const double wr = double(v[j].x), wi = double(v[j].y);
    if (diag) {
      acc[kDcTrRe] += wr;
      acc[kDcTrIm] += wi;
      acc[kDcPTrRe] += p * wr;
      acc[kDcPTrIm] += p * wi;
      acc[kDcP2TrRe] += p * p * wr;
      acc[kDcP2TrIm] += p * p * wi;
      acc[kDcBetaRe] += sb * wr;
      acc[kDcBetaIm] += sb * wi;
      trre[j] += wr;
    } else {
      const double ph = gmem(d.p_host)[i];
      acc[kDcA1Re] += wr;
      acc[kDcA1Im] += wi;
      acc[kDcPA1Re] += p * wr;
      acc[kDcPA1Im] += p * wi;
      acc[kDcHA1Re] += ph * wr;
      acc[kDcHA1Im] += ph * wi;
    }
  }
}

// CaldeiraDissipatorOrder2 on one plane (GPU_WignerDirac2D_4x4.py:3011-3030), host grids:
// D(Y) = F_p^-1[ theta F_p[ p Y ] ]; LW2 = W + a D W; W += 2 a D LW2, a = i dt gamma.
template <typename T, int N, int R, int C>
__device__ __forceinline__ void dirac_cl2(cplx<T> (&v)[R], const DiracPlan& d, int t, LdsStrided<T, C> view,
                                          const Twiddles<T>& tw) {
  constexpr int TT = N / R;
  const T inv = T(1.0 / N);
  cplx<T> y[R], lw[R];
#pragma unroll
  for (int j = 0; j < R; ++j) y[j] = cscale(v[j], T(gmem(d.p_host)[t + j * TT]));
  team_fft<T, N, R, -1>(y, t, view, tw);
#pragma unroll
  for (int j = 0; j < R; ++j) y[j] = cscale(y[j], T(gmem(d.theta_host)[t + j * TT]));
  team_fft<T, N, R, +1>(y, t, view, tw);
#pragma unroll
  for (int j = 0; j < R; ++j) {
    lw[j] = cadd(v[j], mul_i(cscale(y[j], inv), T(d.cl_a)));
    y[j] = cscale(lw[j], T(gmem(d.p_host)[t + j * TT]));
  }
  team_fft<T, N, R, -1>(y, t, view, tw);
#pragma unroll
  for (int j = 0; j < R; ++j) y[j] = cscale(y[j], T(gmem(d.theta_host)[t + j * TT]));
  team_fft<T, N, R, +1>(y, t, view, tw);
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = cadd(v[j], mul_i(cscale(y[j], inv), T(2.0 * d.cl_a)));
}

// Column sums of a tile: lanes of one column inside a wave are c + C*t; fold t with xor
// shuffles, then across waves through LDS.
// part: -1 all column sums; 0 the trace, beta and negativity sums; 1 the alpha_1 sums
__device__ __forceinline__ bool dirac_col_owned(int part, int q) {
  const bool alpha = q >= kDcA1Re && q <= kDcHA1Im;
  return part < 0 || (part == 1) == alpha;
}
template <int BLOCK, int C, typename L>
__device__ __forceinline__ void dirac_store_colsums(const DiracPlan& d, int tile, double (&acc)[kDNCol], L* lds,
                                                    int part = -1) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int off = C; off < 64; off <<= 1) {
#pragma unroll
    for (int q = 0; q < kDNCol; ++q) acc[q] += __shfl_xor(acc[q], off, 64);
  }
  constexpr int kWaves = (BLOCK + 63) / 64;
  constexpr int kLanes = C < 64 ? C : 64;
  __syncthreads();
  double* red = reinterpret_cast<double*>(lds);
  const int wave = tid >> 6, lane = tid & 63;
  if (lane < kLanes) {
#pragma unroll
    for (int q = 0; q < kDNCol; ++q) red[(wave * kLanes + lane) * kDNCol + q] = acc[q];
  }
  __syncthreads();
  if (tid < C) {
    double sm[kDNCol];
#pragma unroll
    for (int q = 0; q < kDNCol; ++q) sm[q] = 0.0;
    for (int w = 0; w < kWaves; ++w) {
#pragma unroll
      for (int q = 0; q < kDNCol; ++q) sm[q] += red[(w * kLanes + tid) * kDNCol + q];
    }
#pragma unroll
    for (int q = 0; q < kDNCol; ++q)
      if (dirac_col_owned(part, q)) gmem(d.colsums)[size_t(tile * C + tid) * kDNCol + q] = sm[q];
  }
}

// Four blocks per column tile, four planes each, two of each of two mass-phase classes
// (class 0: planes 0 1 4 5, class 1: 2 3 6 7, class 2: 8 9 12 13, class 3: 10 11 14 15):
//   quarter 0: 0 5 | 10 15 (every diagonal plane: the trace, beta and negativity sums)
//   quarter 1: 3 6 | 9 12  (every alpha_1 plane: the alpha sums)
//   quarter 2: 1 4 | 11 14,  quarter 3: 2 7 | 8 13 (no column sums)
// Each quarter writes only its own column sums.  With 16 planes per tile (one block) the 512^2
// sweep had one wave per SIMD on half of the CUs; two class pairs per tile filled the CUs, four
// quarters give two waves per SIMD (at two multiplier evaluations per block instead of one).
__constant__ int kDiracQuarter[4][2][2] = {{{0, 5}, {10, 15}}, {{3, 6}, {9, 12}}, {{1, 4}, {11, 14}}, {{2, 7}, {8, 13}}};
__constant__ int kDiracQuarterClass[4][2] = {{0, 3}, {1, 2}, {0, 3}, {1, 2}};
template <typename S, typename T, int N, int R, int BLOCK, bool CL>
__device__ __forceinline__ void dirac_psweep_body(const DiracPlan& d, int tile, int quarter, cplx<T>* lds) {
  constexpr int TT = N / R;
  constexpr int C = BLOCK / TT;  // columns per tile
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int X = d.X;
  const int col = tile * C + c;
  const size_t plane = size_t(d.P) * size_t(X);
  const gptr<cplx<S>> W0 = gmem(reinterpret_cast<cplx<S>*>(d.W));
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(d.twP_hi), reinterpret_cast<const cplx<T>*>(d.twP_lo)};
  LdsStrided<T, C> view{lds + c};
  const int step = *gmem(d.step);

  // Bopp phase dt (V0(x - theta/2) - V0(x + theta/2)) of this lane's elements (shared by the 16 planes)
  const double x = d.dx * double(fidx(col, X));
  const double mask = d.xmask ? gmem(d.xmask)[col] : 1.0;
  double ph[R];
  {
    const int prow = step < d.pot_rows ? step : d.pot_rows - 1;
    const BoppOdd bopp = bopp_odd(cmem(d.pot) + size_t(prow) * (kMaxDeg + 1), d.pot_deg, x, 1.0);
    // KvN: e^{i dt V0'(x) theta} for every plane (KvN_XTheta_Kernel, GPU_WignerDiracKvN2D_4x4.py:356-420)
    double dvx = 0.0;
    if (d.kvn) {
      if (d.pot_deg >= 0) {
        const cptr<double> cf = cmem(d.pot) + size_t(prow) * (kMaxDeg + 1);
        for (int n = d.pot_deg; n >= 1; --n) dvx = fma(dvx, x, double(n) * cf[n]);
      } else {
        dvx = gmem(d.dv0)[col];
      }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = t + j * TT;
      const double th = d.dtheta * double(fidx(i, N));
      if (d.kvn) {
        ph[j] = -d.dt * dvx * th;
        continue;
      }
      const double dv = d.pot_deg >= 0 ? bopp(0.5 * th) : gmem(d.pot_dv)[size_t(i) * X + col];
      ph[j] = d.dt * dv;
      if (d.gp_coeff != 0.0) ph[j] += d.dt * d.gp_coeff * gmem(d.G)[size_t(i) * X + col];
    }
  }

  double acc[kDNCol];
#pragma unroll
  for (int q = 0; q < kDNCol; ++q) acc[q] = 0.0;
  double trre[R];
#pragma unroll
  for (int j = 0; j < R; ++j) trre[j] = 0.0;

  const T inv = T(1.0 / N);
  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<S>));
  // planes by mass-phase class: spinor rows {1,2} or {3,4} x columns {1,2} or {3,4}.  Class
  // multiplier r e^{-i (ph + m)}: m = F- - F+ (lo, lo), F- + F+ (lo, hi), -(F- + F+) (hi, lo),
  // -(F- - F+) (hi, hi), F -/+ = Mass(x -/+ theta/2) c^2 dt; a constant mass gives 0, 2F, -2F, 0.
  for (int ci = 0; ci < 2; ++ci) {
    const int cls = kDiracQuarterClass[quarter][ci];
    cplx<double> M[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = t + j * TT;
      double m;
      if (d.mass_dm) {
        const size_t e = size_t(i) * X + col;
        m = cls == 0 ? gmem(d.mass_dm)[e] : cls == 1 ? gmem(d.mass_sm)[e] : cls == 2 ? -gmem(d.mass_sm)[e] : -gmem(d.mass_dm)[e];
      } else {
        m = cls == 1 ? 2.0 * d.mass_phase : cls == 2 ? -2.0 * d.mass_phase : 0.0;
      }
      if (d.kvn) m = 0.0;
      const double r = (d.theta_decay ? gmem(d.theta_decay)[i] : 1.0) * mask;
      M[j] = expmi(r, ph[j] + m);
    }
    for (int q = 0; q < 2; ++q) {
      const int k = kDiracQuarter[quarter][ci][q];
      const gptr<cplx<S>> W = W0 + size_t(k) * plane;
      cplx<T> v[R];
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = to_arith<T>(gload(at_off(W + size_t(j * TT) * X, boff)));
      team_fft<T, N, R, -1>(v, t, view, tw);
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = apply(v[j], M[j]);
      team_fft<T, N, R, +1>(v, t, view, tw);
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = cscale(v[j], inv);

      if constexpr (CL) dirac_cl2<T, N, R>(v, d, t, view, tw);

      dirac_accumulate<T, R, TT>(d, k, v, t, acc, trre);
#pragma unroll
      for (int j = 0; j < R; ++j) store_stream(at_off(W + size_t(j * TT) * X, boff), to_arith<S>(v[j]));
    }
  }
#pragma unroll
  for (int j = 0; j < R; ++j) acc[kDcNeg] += trre[j] < 0.0 ? trre[j] : 0.0;

  if (quarter < 2) dirac_store_colsums<BLOCK, C>(d, tile, acc, lds, quarter);
}


// ------------------------------------------------------------------ p sweeps with vector potentials
// With A = (A1, A2, A3) != 0 the potential factors P-(x - theta/2) (left) and P+(x + theta/2)
// (right) couple all four spinor rows (columns) at each (theta, x) cell
// (DiracPropagator_X_minus/plus_Theta, GPU_WignerDirac2D_4x4.py:357-670).  The left factor
// acts on the 4 planes of one spinor column, the right factor on the 4 planes of one row, so
// the step uses two sweeps: dirac_pleft (FFT_p, scalar phase, left factor; the state stays in
// the (theta, x) picture) and dirac_pright (right factor, iFFT_p, damping, column sums).
//
// Factor at y with sign s (-1: left at x - theta/2, +1: right at x + theta/2), F = dt
// sqrt((mh c^2)^2 + |A|^2): U11 = U22 = cos F + s i mh c^2 dt sF, U33 = U44 = conj,
// U13 = U31 = -s i dt A3 sF, U14 = U32 = -s dt (i A1 + A2) sF, U23 = U41 = -s dt (i A1 - A2) sF,
// U24 = U42 = -U13, zero elsewhere; sF = sin F / F.
struct DiracPot {
  cplx<double> u11, u33, u13, u14, u23;
};
// Coefficients [3][kMaxDeg + 1] of A1..A3 at time index `step`.
__device__ __forceinline__ const double* dirac_acoef(const DiracPlan& d, int step) {
  if (d.a_rows <= 0) return &d.acoef[0][0];
  const int row = step < d.a_rows ? step : d.a_rows - 1;
  return d.acoef_rows + size_t(row) * 3 * (kMaxDeg + 1);
}

// A1..A3 at y = x -/+ theta/2 of cell e (side 0: left, 1: right): the tables of the programs
// or the polynomial coefficients ac.
struct DiracA {
  double a1, a2, a3;
};
__device__ __forceinline__ DiracA dirac_avals(const DiracPlan& d, const double* __restrict__ ac, double y,
                                              size_t e, int side) {
  if (d.a_tab) {
    const size_t n = size_t(d.P) * size_t(d.X);
    return {d.a_tab[size_t(side) * n + e], d.a_tab[size_t(2 + side) * n + e], d.a_tab[size_t(4 + side) * n + e]};
  }
  constexpr int K = kMaxDeg + 1;
  return {d.adeg0 >= 0 ? poly(ac, d.adeg0, y) : 0.0, d.adeg1 >= 0 ? poly(ac + K, d.adeg1, y) : 0.0,
          d.adeg2 >= 0 ? poly(ac + 2 * K, d.adeg2, y) : 0.0};
}

// mdt: Mass(y) c^2 dt (mass_half c^2 dt for a constant mass)
__device__ __forceinline__ DiracPot dirac_pot_factor(const DiracPlan& d, DiracA av, double s, double mdt) {
  const double a1 = av.a1, a2 = av.a2, a3 = av.a3;
  const double F = sqrt(mdt * mdt + (a1 * a1 + a2 * a2 + a3 * a3) * d.dt * d.dt);
  double sn, cs;
  sincos_fp64(F, &sn, &cs);
  const double sf = F > 0.0 ? sn / F : 1.0;
  DiracPot u;
  u.u11 = {cs, s * mdt * sf};
  u.u33 = {cs, -s * mdt * sf};
  u.u13 = {0.0, -s * d.dt * a3 * sf};
  u.u14 = {-s * d.dt * a2 * sf, -s * d.dt * a1 * sf};
  u.u23 = {s * d.dt * a2 * sf, -s * d.dt * a1 * sf};
  return u;
}

template <typename S, typename T, int N, int R, int BLOCK>
__device__ __forceinline__ void dirac_pleft_body(const DiracPlan& d, int tile, cplx<T>* lds) {
  constexpr int TT = N / R;
  constexpr int C = BLOCK / TT;
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int X = d.X;
  const int col = tile * C + c;
  const size_t plane = size_t(d.P) * size_t(X);
  const gptr<cplx<S>> W0 = gmem(reinterpret_cast<cplx<S>*>(d.W));
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(d.twP_hi), reinterpret_cast<const cplx<T>*>(d.twP_lo)};
  LdsStrided<T, C> view{lds + c};
  const int step = *gmem(d.step);
  const double* __restrict__ ac = dirac_acoef(d, step);  // A1..A3 at this time index
  const double x = d.dx * double(fidx(col, X));
  const double mask = d.xmask ? gmem(d.xmask)[col] : 1.0;
  const int prow = step < d.pot_rows ? step : d.pot_rows - 1;
  const BoppOdd bopp = bopp_odd(cmem(d.pot) + size_t(prow) * (kMaxDeg + 1), d.pot_deg, x, 1.0);
  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<S>));
  for (int jc = 0; jc < 4; ++jc) {  // spinor column jc: planes 4 i + jc
    cplx<T> v[4][R];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const gptr<cplx<S>> W = W0 + size_t(4 * q + jc) * plane;
#pragma unroll
      for (int j = 0; j < R; ++j) v[q][j] = to_arith<T>(gload(at_off(W + size_t(j * TT) * X, boff)));
      team_fft<T, N, R, -1>(v[q], t, view, tw);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = t + j * TT;
      const double th = d.dtheta * double(fidx(i, N));
      // both scalar halves e- e+ = r e^{-i dt (V0(x - th/2) - V0(x + th/2) + a G)} applied here
      double ph = d.pot_deg >= 0 ? bopp(0.5 * th) : gmem(d.pot_dv)[size_t(i) * X + col];
      ph *= d.dt;
      if (d.gp_coeff != 0.0) ph += d.dt * d.gp_coeff * gmem(d.G)[size_t(i) * X + col];
      const double r = (d.theta_decay ? gmem(d.theta_decay)[i] : 1.0) * mask;
      const cplx<double> e = expmi(r, ph);
      // F(x - theta/2) = (sm + dm) / 2 of the mass tables (MassPotential variant)
      const size_t e_ = size_t(i) * X + col;
      const double mdt = d.mass_dm ? 0.5 * (gmem(d.mass_sm)[e_] + gmem(d.mass_dm)[e_]) : d.mass_phase;
      const DiracPot u = dirac_pot_factor(d, dirac_avals(d, ac, x - 0.5 * th, e_, 0), -1.0, mdt);
      const cplx<double> w0 = widen(v[0][j]), w1 = widen(v[1][j]), w2 = widen(v[2][j]), w3 = widen(v[3][j]);
      const cplx<double> mu13{-u.u13.x, -u.u13.y};
      const cplx<double> o0 = cadd(cmul(u.u11, w0), cadd(cmul(u.u13, w2), cmul(u.u14, w3)));
      const cplx<double> o1 = cadd(cmul(u.u11, w1), cadd(cmul(u.u23, w2), cmul(mu13, w3)));
      const cplx<double> o2 = cadd(cmul(u.u13, w0), cadd(cmul(u.u14, w1), cmul(u.u33, w2)));
      const cplx<double> o3 = cadd(cmul(u.u23, w0), cadd(cmul(mu13, w1), cmul(u.u33, w3)));
      v[0][j] = narrow<T>(cmul(e, o0));
      v[1][j] = narrow<T>(cmul(e, o1));
      v[2][j] = narrow<T>(cmul(e, o2));
      v[3][j] = narrow<T>(cmul(e, o3));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const gptr<cplx<S>> W = W0 + size_t(4 * q + jc) * plane;
#pragma unroll
      for (int j = 0; j < R; ++j) store_stream(at_off(W + size_t(j * TT) * X, boff), to_arith<S>(v[q][j]));
    }
  }
}

template <typename S, typename T, int N, int R, int BLOCK, bool CL>
__device__ __forceinline__ void dirac_pright_body(const DiracPlan& d, int tile, cplx<T>* lds) {
  constexpr int TT = N / R;
  constexpr int C = BLOCK / TT;
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int X = d.X;
  const int col = tile * C + c;
  const size_t plane = size_t(d.P) * size_t(X);
  const gptr<cplx<S>> W0 = gmem(reinterpret_cast<cplx<S>*>(d.W));
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(d.twP_hi), reinterpret_cast<const cplx<T>*>(d.twP_lo)};
  LdsStrided<T, C> view{lds + c};
  const double x = d.dx * double(fidx(col, X));
  const double* __restrict__ ac = dirac_acoef(d, *gmem(d.step));  // the time index of dirac_pleft
  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<S>));
  const T inv = T(1.0 / N);
  double acc[kDNCol];
#pragma unroll
  for (int q = 0; q < kDNCol; ++q) acc[q] = 0.0;
  double trre[R];
#pragma unroll
  for (int j = 0; j < R; ++j) trre[j] = 0.0;
  for (int ir = 0; ir < 4; ++ir) {  // spinor row ir: planes 4 ir + j
    cplx<T> v[4][R];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const gptr<cplx<S>> W = W0 + size_t(4 * ir + q) * plane;
#pragma unroll
      for (int j = 0; j < R; ++j) v[q][j] = to_arith<T>(gload(at_off(W + size_t(j * TT) * X, boff)));
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = t + j * TT;
      const double th = d.dtheta * double(fidx(i, N));
      // F(x + theta/2) = (sm - dm) / 2 of the mass tables (MassPotential variant)
      const size_t e_ = size_t(i) * X + col;
      const double mdt = d.mass_dm ? 0.5 * (gmem(d.mass_sm)[e_] - gmem(d.mass_dm)[e_]) : d.mass_phase;
      const DiracPot u = dirac_pot_factor(d, dirac_avals(d, ac, x + 0.5 * th, e_, 1), 1.0, mdt);
      const cplx<double> w0 = widen(v[0][j]), w1 = widen(v[1][j]), w2 = widen(v[2][j]), w3 = widen(v[3][j]);
      const cplx<double> mu13{-u.u13.x, -u.u13.y};
      // (W U)_{ij} = sum_k W_ik U_kj with U31 = U13, U32 = U14, U41 = U23, U42 = U24 = -U13
      v[0][j] = narrow<T>(cadd(cmul(w0, u.u11), cadd(cmul(w2, u.u13), cmul(w3, u.u23))));
      v[1][j] = narrow<T>(cadd(cmul(w1, u.u11), cadd(cmul(w2, u.u14), cmul(w3, mu13))));
      v[2][j] = narrow<T>(cadd(cmul(w0, u.u13), cadd(cmul(w1, u.u23), cmul(w2, u.u33))));
      v[3][j] = narrow<T>(cadd(cmul(w0, u.u14), cadd(cmul(w1, mu13), cmul(w3, u.u33))));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = 4 * ir + q;
      const gptr<cplx<S>> W = W0 + size_t(k) * plane;
      if (d.odm) {  // ODM damping follows in the theta picture (dirac_podm_body with A)
#pragma unroll
        for (int j = 0; j < R; ++j) store_stream(at_off(W + size_t(j * TT) * X, boff), to_arith<S>(v[q][j]));
        continue;
      }
      team_fft<T, N, R, +1>(v[q], t, view, tw);
#pragma unroll
      for (int j = 0; j < R; ++j) v[q][j] = cscale(v[q][j], inv);
      if constexpr (CL) dirac_cl2<T, N, R>(v[q], d, t, view, tw);
      dirac_accumulate<T, R, TT>(d, k, v[q], t, acc, trre);
#pragma unroll
      for (int j = 0; j < R; ++j) store_stream(at_off(W + size_t(j * TT) * X, boff), to_arith<S>(v[q][j]));
    }
  }
  if (d.odm) return;  // dirac_podm_body transforms back and writes the column sums
#pragma unroll
  for (int j = 0; j < R; ++j) acc[kDcNeg] += trre[j] < 0.0 ? trre[j] : 0.0;
  dirac_store_colsums<BLOCK, C>(d, tile, acc, lds);
}

// ------------------------------------------------------------------ p sweep with ODM damping
// DampingODM_Kernel (GPU_WignerDirac2D_4x4.py:952-1195, called at :3608-3615 between the
// potential factors and Theta_To_P) couples, at each (theta, x) cell, the four planes
// a = W_ij, d = W_(5-i)(5-j), b = W_i(5-j), c = W_(5-i)j (i, j in {1, 2}), one of each
// mass-phase class.  With u = a - d, v = a + d, w = b - c, z = b + c,
//   P = A (cx u + i sx w), Q = B (ct v + i st z), R = A (cx w + i sx u), S = B (ct z + i st v),
//   a' = (P + Q)/2, d' = (Q - P)/2, b' = (R + S)/2, c' = (S - R)/2,
// A = e^{-8 mc t g x^2 / lb}, cx + i sx = e^{4 i mc t g x}, B = e^{-2 mc t g th^2 / lb},
// ct + i st = e^{2 i mc t g th} (mc = mass c, t = dt, g = gammaDamping, lb = lambdaBar, kernel
// coordinates).  One sweep per step: each block takes the four planes of a quad through
// FFT_p, the diagonal potential factors of A = 0, the mixing and iFFT_p.
template <typename S, typename T, int N, int R, int BLOCK>
__device__ __forceinline__ void dirac_podm_body(const DiracPlan& d, int tile, cplx<T>* lds) {
  constexpr int TT = N / R;
  constexpr int C = BLOCK / TT;
  const int tid = threadIdx.x;
  const int c = tid % C, t = tid / C;
  const int X = d.X;
  const int col = tile * C + c;
  const size_t plane = size_t(d.P) * size_t(X);
  const gptr<cplx<S>> W0 = gmem(reinterpret_cast<cplx<S>*>(d.W));
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(d.twP_hi), reinterpret_cast<const cplx<T>*>(d.twP_lo)};
  LdsStrided<T, C> view{lds + c};
  const int step = *gmem(d.step);
  const double x = d.dx * double(fidx(col, X));
  const double mask = d.xmask ? gmem(d.xmask)[col] : 1.0;
  const int prow = step < d.pot_rows ? step : d.pot_rows - 1;
  const BoppOdd bopp = bopp_odd(cmem(d.pot) + size_t(prow) * (kMaxDeg + 1), d.pot_deg, x, 1.0);
  const unsigned boff = (unsigned(t) * unsigned(X) + unsigned(col)) * unsigned(sizeof(cplx<S>));
  const T inv = T(1.0 / N);
  // x factors of the mixing (per column)
  const double kx = d.odm_k * x;  // mc t g x
  double sx, cx;
  sincos_fp64(4.0 * kx, &sx, &cx);
  const double Ax = exp(-8.0 * kx * x * d.odm_inv_lb);
  double acc[kDNCol];
#pragma unroll
  for (int q = 0; q < kDNCol; ++q) acc[q] = 0.0;
  double trre[R];
#pragma unroll
  for (int j = 0; j < R; ++j) trre[j] = 0.0;
  for (int quad = 0; quad < 4; ++quad) {
    const int i0 = quad >> 1, j0 = quad & 1;
    // planes a (rows lo, cols lo: class 0), b (lo, hi: class 1), c (hi, lo: class 2), d (hi, hi: 3)
    const int kp[4] = {4 * i0 + j0, 4 * i0 + (3 - j0), 4 * (3 - i0) + j0, 4 * (3 - i0) + (3 - j0)};
    cplx<T> v[4][R];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const gptr<cplx<S>> W = W0 + size_t(kp[q]) * plane;
#pragma unroll
      for (int j = 0; j < R; ++j) v[q][j] = to_arith<T>(gload(at_off(W + size_t(j * TT) * X, boff)));
      // with A the potential sweeps left the planes in the theta picture, factors applied
      if (!d.general) team_fft<T, N, R, -1>(v[q], t, view, tw);
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int i = t + j * TT;
      const double th = d.dtheta * double(fidx(i, N));
      double ph = d.pot_deg >= 0 ? bopp(0.5 * th) : gmem(d.pot_dv)[size_t(i) * X + col];
      ph *= d.dt;
      if (d.gp_coeff != 0.0) ph += d.dt * d.gp_coeff * gmem(d.G)[size_t(i) * X + col];
      const double r = (d.theta_decay ? gmem(d.theta_decay)[i] : 1.0) * mask;
      double m[4];
      if (d.mass_dm) {
        const size_t e = size_t(i) * X + col;
        const double dm = gmem(d.mass_dm)[e], sm = gmem(d.mass_sm)[e];
        m[0] = dm;
        m[1] = sm;
        m[2] = -sm;
        m[3] = -dm;
      } else {
        m[0] = 0.0;
        m[1] = 2.0 * d.mass_phase;
        m[2] = -2.0 * d.mass_phase;
        m[3] = 0.0;
      }
      cplx<double> w[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) w[q] = d.general ? widen(v[q][j]) : cmul(widen(v[q][j]), expmi(r, ph + m[q]));
      // ODM mixing of the quad (a = w0, b = w1, c = w2, d = w3)
      const double kt = d.odm_k * th;  // mc t g theta
      double st, ct;
      sincos_fp64(2.0 * kt, &st, &ct);
      const double Bt = exp(-2.0 * kt * th * d.odm_inv_lb);
      const cplx<double> u = csub(w[0], w[3]), vv = cadd(w[0], w[3]);
      const cplx<double> ww = csub(w[1], w[2]), z = cadd(w[1], w[2]);
      const cplx<double> Pp = cscale(cadd(cscale(u, cx), mul_i(ww, sx)), 0.5 * Ax);
      const cplx<double> Qp = cscale(cadd(cscale(vv, ct), mul_i(z, st)), 0.5 * Bt);
      const cplx<double> Rp = cscale(cadd(cscale(ww, cx), mul_i(u, sx)), 0.5 * Ax);
      const cplx<double> Sp = cscale(cadd(cscale(z, ct), mul_i(vv, st)), 0.5 * Bt);
      v[0][j] = narrow<T>(cadd(Pp, Qp));
      v[3][j] = narrow<T>(csub(Qp, Pp));
      v[1][j] = narrow<T>(cadd(Rp, Sp));
      v[2][j] = narrow<T>(csub(Sp, Rp));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      team_fft<T, N, R, +1>(v[q], t, view, tw);
#pragma unroll
      for (int j = 0; j < R; ++j) v[q][j] = cscale(v[q][j], inv);
      dirac_accumulate<T, R, TT>(d, kp[q], v[q], t, acc, trre);
      const gptr<cplx<S>> W = W0 + size_t(kp[q]) * plane;
#pragma unroll
      for (int j = 0; j < R; ++j) store_stream(at_off(W + size_t(j * TT) * X, boff), to_arith<S>(v[q][j]));
    }
  }
#pragma unroll
  for (int j = 0; j < R; ++j) acc[kDcNeg] += trre[j] < 0.0 ? trre[j] : 0.0;
  dirac_store_colsums<BLOCK, C>(d, tile, acc, lds);
}

// ------------------------------------------------------------------ kernels
constexpr int kDiracXBlock = 256;
constexpr int kDiracPBlock = 256;

// radix of the x sweep (4 planes live per lane) and of the p sweep (1 plane live)
template <int N>
constexpr int dirac_rx() {
  return N >= 512 ? 8 : 4;
}
// radix of the kinetic x sweep for the state type S: at 512 a complex128 state takes radix 4
// (two waves per row, 203 VGPRs, two waves per SIMD); radix 8 needs more than the 512
// registers of a lane there and spilled 40 VGPRs to scratch (+35% memory traffic, 6% slower
// per step).  The complex64 state keeps radix 8 (no scratch, 5% faster than radix 4).
template <int N, typename S>
constexpr int dirac_rx_state() {
  return (N == 512 && sizeof(S) == 8) ? 4 : dirac_rx<N>();
}
template <int N>
constexpr int dirac_rp() {
  return N >= 512 ? 8 : 4;  // at 512 radix 4 (three waves per SIMD) is 8% slower per step
}
// radix of the vector-potential p sweeps (4 planes live per lane and two factor matrices)
template <int N>
constexpr int dirac_rv() {
  return N >= 2048 ? 8 : 4;
}

// LDS elements of a block over the supported sizes: teams * padded length.
template <int BLOCK>
constexpr int dirac_lds_elems() {
  int m = 0;
  for (int N = 128; N <= 2048; N *= 2) {
    for (int R : {N >= 2048 ? 8 : 4, N >= 512 ? 8 : 4}) {
      const int teams = BLOCK / (N / R);
      const int e = teams * (N + N / R + 1);
      m = e > m ? e : m;
    }
  }
  return m;
}

// Column tile of a p-sweep block: blocks are dealt round-robin to the 8 XCDs, and two
// neighbouring FP64 tiles (4 columns, 64 bytes) share each 128-byte line, so the tiles of
// one XCD are made contiguous: both halves of a line then pass through the same L2.
__device__ __forceinline__ int dirac_col_tile() {
  const int nt = int(gridDim.x), b = int(blockIdx.x);
  return nt % 8 == 0 ? (b % 8) * (nt / 8) + b / 8 : b;
}

#define WH_DIRAC_DISPATCH(NVAR, BODY) \
  switch (NVAR) {                     \
    case 128: BODY(128); break;       \
    case 256: BODY(256); break;       \
    case 512: BODY(512); break;       \
    case 1024: BODY(1024); break;     \
    case 2048: BODY(2048); break;     \
    default: break;                   \
  }

// One kernel per grid side: the register demand differs widely between sides (up to two
// waves per SIMD at 512 and below, one above), and one kernel for all sides would take the
// largest one's occupancy.
template <typename S, typename T, int N, bool FILTER>
__global__ void __launch_bounds__(kDiracXBlock) dirac_xsweep_kernel(const DiracPlan* __restrict__ dp) {
  __shared__ cplx<T> lds[dirac_lds_elems<kDiracXBlock>()];
  const DiracPlan& d = *dp;
  const int group = int(blockIdx.x & 3u), tile = int(blockIdx.x >> 2);
  dirac_xsweep_body<S, T, N, dirac_rx_state<N, S>(), kDiracXBlock, FILTER>(d, group, tile, lds);
}

template <typename S, typename T, int N, bool CL>
__global__ void __launch_bounds__(kDiracPBlock) dirac_psweep_kernel(const DiracPlan* __restrict__ dp) {
  __shared__ cplx<T> lds[dirac_lds_elems<kDiracPBlock>()];
  const DiracPlan& d = *dp;
  const int u = dirac_col_tile();
  dirac_psweep_body<S, T, N, dirac_rp<N>(), kDiracPBlock, CL>(d, u >> 2, u & 3, lds);
}

template <typename S, typename T>
__global__ void __launch_bounds__(kDiracXBlock) dirac_pleft_kernel(const DiracPlan* __restrict__ dp) {
  __shared__ cplx<T> lds[dirac_lds_elems<kDiracXBlock>()];
  const DiracPlan& d = *dp;
#define BODY(n) dirac_pleft_body<S, T, n, dirac_rv<n>(), kDiracXBlock>(d, dirac_col_tile(), lds)
  WH_DIRAC_DISPATCH(d.P, BODY)
#undef BODY
}

template <typename S, typename T, bool CL>
__global__ void __launch_bounds__(kDiracXBlock) dirac_pright_kernel(const DiracPlan* __restrict__ dp) {
  __shared__ cplx<T> lds[dirac_lds_elems<kDiracXBlock>()];
  const DiracPlan& d = *dp;
#define BODY(n) dirac_pright_body<S, T, n, dirac_rv<n>(), kDiracXBlock, CL>(d, dirac_col_tile(), lds)
  WH_DIRAC_DISPATCH(d.P, BODY)
#undef BODY
}

template <typename S, typename T>
__global__ void __launch_bounds__(kDiracXBlock) dirac_podm_kernel(const DiracPlan* __restrict__ dp) {
  __shared__ cplx<T> lds[dirac_lds_elems<kDiracXBlock>()];
  const DiracPlan& d = *dp;
#define BODY(n) dirac_podm_body<S, T, n, dirac_rv<n>(), kDiracXBlock>(d, dirac_col_tile(), lds)
  WH_DIRAC_DISPATCH(d.P, BODY)
#undef BODY
}

// Column sums of a state without propagating it (the record of the first state).
// Column sums of the state (no sweep): a block takes kDColsumCols consecutive columns, its
// 256 lanes 64 row groups of kDColsumCols lanes (each load instruction covers 16 rows of the
// tile's columns); the group sums are combined by a fixed tree in LDS.  One thread per column
// over all rows took 440 us at 512^2 x 16 planes (two blocks); this form fills the GPU.
constexpr int kDColsumCols = 4;
template <typename T>
__global__ void __launch_bounds__(256) dirac_colsum_kernel(const DiracPlan* __restrict__ dp) {
  constexpr int CC = kDColsumCols, G = 256 / CC;
  __shared__ double red[G][CC][kDNCol];
  const DiracPlan& d = *dp;
  const int c = threadIdx.x % CC, g = threadIdx.x / CC;
  const int col = blockIdx.x * CC + c;
  const size_t plane = size_t(d.P) * size_t(d.X);
  const cplx<T>* W = reinterpret_cast<const cplx<T>*>(d.W);
  double acc[kDNCol];
#pragma unroll
  for (int q = 0; q < kDNCol; ++q) acc[q] = 0.0;
  if (col < d.X) {
    for (int i = g; i < d.P; i += G) {
      const double p = d.dp * double(fidx(i, d.P)), ph = gmem(d.p_host)[i];
      const size_t e = size_t(i) * d.X + col;
      double trr = 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const bool diag = (k % 5) == 0, alpha = k == 3 || k == 6 || k == 9 || k == 12;
        if (!diag && !alpha) continue;
        const cplx<T> w = W[size_t(k) * plane + e];
        const double wr = double(w.x), wi = double(w.y);
        if (diag) {
          const double sb = k < 8 ? 1.0 : -1.0;
          acc[kDcTrRe] += wr;
          acc[kDcTrIm] += wi;
          acc[kDcPTrRe] += p * wr;
          acc[kDcPTrIm] += p * wi;
          acc[kDcP2TrRe] += p * p * wr;
          acc[kDcP2TrIm] += p * p * wi;
          acc[kDcBetaRe] += sb * wr;
          acc[kDcBetaIm] += sb * wi;
          trr += wr;
        } else {
          acc[kDcA1Re] += wr;
          acc[kDcA1Im] += wi;
          acc[kDcPA1Re] += p * wr;
          acc[kDcPA1Im] += p * wi;
          acc[kDcHA1Re] += ph * wr;
          acc[kDcHA1Im] += ph * wi;
        }
      }
      acc[kDcNeg] += trr < 0.0 ? trr : 0.0;
    }
  }
#pragma unroll
  for (int q = 0; q < kDNCol; ++q) red[g][c][q] = acc[q];
  __syncthreads();
  for (int h = G / 2; h > 0; h >>= 1) {
    if (g < h) {
#pragma unroll
      for (int q = 0; q < kDNCol; ++q) red[g][c][q] += red[g + h][c][q];
    }
    __syncthreads();
  }
  if (g == 0 && col < d.X) {
#pragma unroll
    for (int q = 0; q < kDNCol; ++q) gmem(d.colsums)[size_t(col) * kDNCol + q] = red[0][c][q];
  }
}

// Observables of the state whose column sums are in d.colsums (state time index = step +
// state_add), normalised by its trace norm when renorm; the pending load factor becomes
// 1/norm (renorm) or 1.  With anti, the antiparticle entry of the state at time index
// `step` is filled from the x sweep's row contributions.  inc: advance the step counter.
constexpr int kRedThreads = 256;
constexpr int kRedQ = kDoNorm + 1;  // reduced quantities
__global__ void __launch_bounds__(kRedThreads) dirac_reduce_kernel(const DiracPlan* __restrict__ dp, int state_add,
                                                                   int renorm, int anti, int inc) {
  const DiracPlan& d = *dp;
  __shared__ double red[kRedThreads][kRedQ + 1];
  const int step = *gmem(d.step);
  const int tstate = step + state_add;
  const int prow = tstate < d.pot_rows ? tstate : d.pot_rows - 1;
  const double* __restrict__ cf = d.pot + size_t(prow) * (kMaxDeg + 1);
  double s[kRedQ];
  for (int q = 0; q < kRedQ; ++q) s[q] = 0.0;
  for (int col = threadIdx.x; col < d.X; col += kRedThreads) {
    const double* cs = d.colsums + size_t(col) * kDNCol;
    const double x = d.dx * double(fidx(col, d.X));
    double dv, vs;
    if (d.pot_deg >= 0) {  // dV0/dx and V0 of this step's coefficient row
      dv = 0.0;
      for (int n = d.pot_deg; n >= 1; --n) dv = fma(dv, x, double(n) * cf[n]);
      vs = poly(cf, d.pot_deg, x);
    } else {
      dv = gmem(d.dv0)[col];
      vs = gmem(d.vsum)[col];
    }
    // DiracEnergy adds Potential_1..3 like Potential_0 (:2801-2804); a V0 table from the host
    // holds them already unless they depend on time (one from V0's program holds V0 alone)
    if (d.asum) {  // A1..A3 programs: their sum at this row's time
      vs += gmem(d.asum)[col];
    } else if (d.general && (d.pot_deg >= 0 || d.a_rows > 0 || d.pot_expr_len > 0)) {
      const double* ac = dirac_acoef(d, tstate);
      if (d.adeg0 >= 0) vs += poly(ac, d.adeg0, x);
      if (d.adeg1 >= 0) vs += poly(ac + (kMaxDeg + 1), d.adeg1, x);
      if (d.adeg2 >= 0) vs += poly(ac + 2 * (kMaxDeg + 1), d.adeg2, x);
    }
    const double tr[2] = {cs[kDcTrRe], cs[kDcTrIm]}, ptr[2] = {cs[kDcPTrRe], cs[kDcPTrIm]};
    const double a1[2] = {cs[kDcA1Re], cs[kDcA1Im]};
    for (int c = 0; c < 2; ++c) {
      s[kDoX + c] += x * tr[c];
      s[kDoP + c] += ptr[c];
      s[kDoXP + c] += x * ptr[c];
      s[kDoXX + c] += x * x * tr[c];
      s[kDoPP + c] += cs[kDcP2TrRe + c];
      s[kDoA1 + c] += a1[c];
      s[kDoDV + c] += dv * tr[c];
      s[kDoXDV + c] += x * dv * tr[c];
      s[kDoPDV + c] += dv * ptr[c];
      s[kDoPA1 + c] += cs[kDcPA1Re + c];
      s[kDoXA1 + c] += x * a1[c];
      s[kDoEnergy + c] += d.c_host * cs[kDcHA1Re + c] + d.mc2_host * cs[kDcBetaRe + c] + vs * tr[c];
    }
    s[kDoNeg] += cs[kDcNeg];
    s[kDoTrans] += x > d.transmission_x ? tr[0] : 0.0;
    s[kDoNorm] += tr[0];
  }
  // antiparticle population from the row contributions of the loaded state (normalised)
  double a = 0.0;
  if (anti && d.antiparticle) {
    for (int r = threadIdx.x; r < d.P; r += kRedThreads) a += gmem(d.rowpart)[r] + gmem(d.rowpart)[d.P + r];
  }
  // per-wave sums (fixed xor trees), then the waves in order: red[0][q] holds the totals
  constexpr int kWaves = kRedThreads / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kRedQ; ++q) s[q] = wave_sum(s[q]);
  a = wave_sum(a);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < kRedQ; ++q) red[wave][q] = s[q];
    red[wave][kRedQ] = a;
  }
  __syncthreads();
  if (threadIdx.x <= kRedQ) {
    double t = 0.0;
    for (int w = 0; w < kWaves; ++w) t += red[w][threadIdx.x];
    red[kWaves][threadIdx.x] = t;
  }
  __syncthreads();
  const double* tot = red[kWaves];
  const double anti_sum = tot[kRedQ];
  __shared__ double fshare;
  if (threadIdx.x == 0) fshare = renorm ? 1.0 / (d.dxdp * tot[kDoNorm]) : 1.0;
  __syncthreads();
  if (d.gp_coeff != 0.0) {  // density of the (normalised) state for the next Gross-Pitaevskii phase
    for (int col = threadIdx.x; col < d.X; col += kRedThreads)
      gmem(d.rho)[col] = fshare * gmem(d.colsums)[size_t(col) * kDNCol + kDcTrRe];
  }
  if (threadIdx.x == 0) {
    const double raw = d.dxdp * tot[kDoNorm];
    const double f = fshare;
    double* o = d.obs + size_t(tstate - d.obs_base) * kDNObs;
    for (int q = 0; q < kDoNeg; ++q) o[q] = f * d.dxdp * tot[q];
    if (!d.energy) o[kDoEnergy] = o[kDoEnergy + 1] = 0.0;
    o[kDoNeg] = f * d.dxdp * tot[kDoNeg];
    o[kDoTrans] = f * d.dxdp * tot[kDoTrans];
    o[kDoNorm] = f * raw;
    o[kDoRawNorm] = raw;
    if (anti && d.antiparticle) d.obs[size_t(step - d.obs_base) * kDNObs + kDoAnti] = d.dxdp * anti_sum;
    if (!anti) o[kDoAnti] = 0.0;
    *d.scale = f;
    if (inc) *const_cast<int*>(d.step) = step + 1;
  }
}

// Closing pass: W *= pending factor (the returned state is normalised) and the
// antiparticle contribution of each row (its lambda = 0 bin is the sum over x).
template <typename T>
__global__ void __launch_bounds__(256) dirac_finish_kernel(const DiracPlan* __restrict__ dp) {
  const DiracPlan& d = *dp;
  const int row = blockIdx.x;
  const size_t plane = size_t(d.P) * size_t(d.X);
  cplx<T>* W = reinterpret_cast<cplx<T>*>(d.W) + size_t(row) * d.X;
  const double s = *gmem(d.scale);
  __shared__ double red[256][4];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};  // sum_x Re of diag-a (11, 22), diag-d (33, 44), off-diagonal b
  for (int k = 0; k < 16; ++k) {
    const int slot = (k == 0 || k == 5) ? 0 : (k == 10 || k == 15) ? 1 : (k == 3 || k == 12 || k == 6 || k == 9) ? 2 : -1;
    for (int x = threadIdx.x; x < d.X; x += 256) {
      cplx<T>& w = W[size_t(k) * plane + x];
      w = cscale(w, T(s));
      if (slot >= 0) acc[slot] += double(w.x);
    }
  }
  for (int q = 0; q < 3; ++q) red[threadIdx.x][q] = acc[q];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (int(threadIdx.x) < h)
      for (int q = 0; q < 3; ++q) red[threadIdx.x][q] += red[threadIdx.x + h][q];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double p = d.dp * double(fidx(row, d.P));
    const double e2 = 2.0 * sqrt(d.proj_mc2 * d.proj_mc2 + d.proj_c * d.proj_c * p * p);
    const double aa = 0.5 - d.proj_mc2 / e2, bb = -d.proj_c * p / e2, dd = 0.5 + d.proj_mc2 / e2;
    gmem(d.rowpart)[row] = aa * red[0][0] + dd * red[0][1] + bb * red[0][2];
    gmem(d.rowpart)[d.P + row] = 0.0;
  }
}

// Antiparticle entry of the record row of the current step (after dirac_finish_kernel).
__global__ void __launch_bounds__(kRedThreads) dirac_anti_kernel(const DiracPlan* __restrict__ dp) {
  const DiracPlan& d = *dp;
  __shared__ double red[kRedThreads];
  double a = 0.0;
  for (int r = threadIdx.x; r < d.P; r += kRedThreads) a += gmem(d.rowpart)[r] + gmem(d.rowpart)[d.P + r];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int h = kRedThreads / 2; h > 0; h >>= 1) {
    if (int(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (d.antiparticle) d.obs[size_t(*d.step - d.obs_base) * kDNObs + kDoAnti] = d.dxdp * red[0];
    *d.scale = 1.0;
  }
}

// G(theta, x) = Re(B- - B+) of the Gross-Pitaevskii density (MakeGrossPitaevskiiTerms,
// GPU_WignerDirac2D_4x4.py:2695-2715), one theta row per team.
constexpr int kDiracGpR = 16;
__global__ void __launch_bounds__(256) dirac_gpshift_kernel(const DiracPlan* __restrict__ dp) {
  __shared__ cplx<double> lds[lds_elems<256, kDiracGpR>()];
  const DiracPlan& d = *dp;
  Member m{};
  m.rho = d.rho;
  m.G = d.G;
  m.theta_host = d.theta_host;
  m.lambda_host = d.lambda_host;
  m.twX64 = d.twX64;
  m.X = d.X;
  m.P = d.P;
  constexpr int R = kDiracGpR, BLOCK = 256;
#define BODY(n) gpshift_body<n, R, BLOCK>(m, int(blockIdx.x), lds)
  WH_DISPATCH(d.X, BODY)
#undef BODY
}

// ------------------------------------------------------------------ launchers
template <int N, typename S>
constexpr int xsweep_rows() {
  return kDiracXBlock / (N / dirac_rx_state<N, S>());
}
template <int N>
constexpr int psweep_cols() {
  return kDiracPBlock / (N / dirac_rp<N>());
}

// blocks of the x sweep for a state of precision code prec (0: complex128)
int dirac_x_blocks(int prec, int X, int P) {
  int rows = 0;
#define BODY(n) rows = prec == 0 ? xsweep_rows<n, double>() : xsweep_rows<n, float>()
  WH_DIRAC_DISPATCH(X, BODY)
#undef BODY
  return rows > 0 && P % rows == 0 ? 4 * (P / rows) : -1;
}

int dirac_p_blocks(int X, int P) {
  int cols = 0;
#define BODY(n) cols = psweep_cols<n>()
  WH_DIRAC_DISPATCH(P, BODY)
#undef BODY
  return cols > 0 && X % cols == 0 ? 4 * (X / cols) : -1;  // four blocks (quarters) per tile
}

template <typename S, typename T, bool FILTER>
hipError_t dirac_x_launch(int X, const DiracPlan* d, int nblocks, hipStream_t s) {
  const dim3 g(nblocks), b(kDiracXBlock);
  switch (X) {
    case 128: hipLaunchKernelGGL((dirac_xsweep_kernel<S, T, 128, FILTER>), g, b, 0, s, d); break;
    case 256: hipLaunchKernelGGL((dirac_xsweep_kernel<S, T, 256, FILTER>), g, b, 0, s, d); break;
    case 512: hipLaunchKernelGGL((dirac_xsweep_kernel<S, T, 512, FILTER>), g, b, 0, s, d); break;
    case 1024: hipLaunchKernelGGL((dirac_xsweep_kernel<S, T, 1024, FILTER>), g, b, 0, s, d); break;
    case 2048: hipLaunchKernelGGL((dirac_xsweep_kernel<S, T, 2048, FILTER>), g, b, 0, s, d); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <bool FILTER>
hipError_t dirac_x_launch_prec(int prec, int X, const DiracPlan* d, int nblocks, hipStream_t s) {
  if (prec == 0) return dirac_x_launch<double, double, FILTER>(X, d, nblocks, s);
  if (prec == 1) return dirac_x_launch<float, float, FILTER>(X, d, nblocks, s);
  return dirac_x_launch<float, double, FILTER>(X, d, nblocks, s);
}

hipError_t launch_dirac_xsweep(int prec, int X, const DiracPlan* d, int nblocks, hipStream_t s) {
  return dirac_x_launch_prec<false>(prec, X, d, nblocks, s);
}

template <typename S, typename T, bool CL>
hipError_t dirac_p_launch(int P, const DiracPlan* d, int nblocks, hipStream_t s) {
  const dim3 g(nblocks), b(kDiracPBlock);
  switch (P) {
    case 128: hipLaunchKernelGGL((dirac_psweep_kernel<S, T, 128, CL>), g, b, 0, s, d); break;
    case 256: hipLaunchKernelGGL((dirac_psweep_kernel<S, T, 256, CL>), g, b, 0, s, d); break;
    case 512: hipLaunchKernelGGL((dirac_psweep_kernel<S, T, 512, CL>), g, b, 0, s, d); break;
    case 1024: hipLaunchKernelGGL((dirac_psweep_kernel<S, T, 1024, CL>), g, b, 0, s, d); break;
    case 2048: hipLaunchKernelGGL((dirac_psweep_kernel<S, T, 2048, CL>), g, b, 0, s, d); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_dirac_psweep(int prec, int P, bool cl, const DiracPlan* d, int nblocks, hipStream_t s) {
  if (prec == 0) return cl ? dirac_p_launch<double, double, true>(P, d, nblocks, s)
                           : dirac_p_launch<double, double, false>(P, d, nblocks, s);
  if (prec == 1) return cl ? dirac_p_launch<float, float, true>(P, d, nblocks, s)
                           : dirac_p_launch<float, float, false>(P, d, nblocks, s);
  return cl ? dirac_p_launch<float, double, true>(P, d, nblocks, s) : dirac_p_launch<float, double, false>(P, d, nblocks, s);
}

int dirac_pg_blocks(int X, int P) {
  int cols = 0;
#define BODY(n) cols = kDiracXBlock / (n / dirac_rv<n>())
  WH_DIRAC_DISPATCH(P, BODY)
#undef BODY
  return cols > 0 && X % cols == 0 ? X / cols : -1;
}

hipError_t launch_dirac_pgeneral(int prec, bool cl, const DiracPlan* d, int nblocks, hipStream_t s) {
  const dim3 g(nblocks), b(kDiracXBlock);
  if (prec == 0) {
    hipLaunchKernelGGL((dirac_pleft_kernel<double, double>), g, b, 0, s, d);
    if (cl)
      hipLaunchKernelGGL((dirac_pright_kernel<double, double, true>), g, b, 0, s, d);
    else
      hipLaunchKernelGGL((dirac_pright_kernel<double, double, false>), g, b, 0, s, d);
  } else if (prec == 1) {
    hipLaunchKernelGGL((dirac_pleft_kernel<float, float>), g, b, 0, s, d);
    if (cl)
      hipLaunchKernelGGL((dirac_pright_kernel<float, float, true>), g, b, 0, s, d);
    else
      hipLaunchKernelGGL((dirac_pright_kernel<float, float, false>), g, b, 0, s, d);
  } else {
    hipLaunchKernelGGL((dirac_pleft_kernel<float, double>), g, b, 0, s, d);
    if (cl)
      hipLaunchKernelGGL((dirac_pright_kernel<float, double, true>), g, b, 0, s, d);
    else
      hipLaunchKernelGGL((dirac_pright_kernel<float, double, false>), g, b, 0, s, d);
  }
  return hipGetLastError();
}

hipError_t launch_dirac_fsweep(int prec, int X, const DiracPlan* d, int nblocks, hipStream_t s) {
  return dirac_x_launch_prec<true>(prec, X, d, nblocks, s);
}

hipError_t launch_dirac_podm(int prec, const DiracPlan* d, int nblocks, hipStream_t s) {
  const dim3 g(nblocks), b(kDiracXBlock);
  if (prec == 0)
    hipLaunchKernelGGL((dirac_podm_kernel<double, double>), g, b, 0, s, d);
  else if (prec == 1)
    hipLaunchKernelGGL((dirac_podm_kernel<float, float>), g, b, 0, s, d);
  else
    hipLaunchKernelGGL((dirac_podm_kernel<float, double>), g, b, 0, s, d);
  return hipGetLastError();
}

hipError_t launch_dirac_colsum(int prec, const DiracPlan* d, int X, hipStream_t s) {
  const int nb = (X + kDColsumCols - 1) / kDColsumCols;
  if (prec == 0)  // state storage only: complex128 or complex64
    hipLaunchKernelGGL(dirac_colsum_kernel<double>, dim3(nb), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(dirac_colsum_kernel<float>, dim3(nb), dim3(256), 0, s, d);
  return hipGetLastError();
}

int dirac_gp_blocks(int X, int P) {
  const int rows = 256 / (X / kDiracGpR);
  return rows > 0 && P % rows == 0 ? P / rows : -1;
}

hipError_t launch_dirac_gpshift(const DiracPlan* d, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(dirac_gpshift_kernel, dim3(nblocks), dim3(256), 0, s, d);
  return hipGetLastError();
}

hipError_t launch_dirac_reduce(const DiracPlan* d, int state_add, int renorm, int anti, int inc, hipStream_t s) {
  hipLaunchKernelGGL(dirac_reduce_kernel, dim3(1), dim3(kRedThreads), 0, s, d, state_add, renorm, anti, inc);
  return hipGetLastError();
}

// Per-call binding of the state and record into the device plan, in stream order (no host
// synchronisation; captured graphs read the plan): W, obs, obs_base, the step counter and,
// optionally, the pending factor reset to 1.
__global__ void dirac_bind_kernel(DiracPlan* dp, void* W, double* obs, int obs_base, int* step, int step_v,
                                  double* scale, int reset_scale) {
  dp->W = W;
  dp->obs = obs;
  dp->obs_base = obs_base;
  *step = step_v;
  if (reset_scale) *scale = 1.0;
}
hipError_t launch_dirac_bind(DiracPlan* d, void* W, double* obs, int obs_base, int* step, int step_v, double* scale,
                             int reset_scale, hipStream_t s) {
  hipLaunchKernelGGL(dirac_bind_kernel, dim3(1), dim3(1), 0, s, d, W, obs, obs_base, step, step_v, scale,
                     reset_scale);
  return hipGetLastError();
}

hipError_t launch_dirac_finish(int prec, const DiracPlan* d, int P, hipStream_t s) {
  if (prec == 0)  // state storage only
    hipLaunchKernelGGL(dirac_finish_kernel<double>, dim3(P), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(dirac_finish_kernel<float>, dim3(P), dim3(256), 0, s, d);
  if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
  hipLaunchKernelGGL(dirac_anti_kernel, dim3(1), dim3(kRedThreads), 0, s, d);
  return hipGetLastError();
}

}  // namespace wh
