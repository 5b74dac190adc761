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

#include <type_traits>

#include "wh_fft.hpp"
#include "wh_kernels.hpp"
#include "wh_math.hpp"
#include "wh_variants.hpp"

namespace wh {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ double poly(const double* __restrict__ c, int deg, double x) {
  double r = c[deg];
  for (int n = deg - 1; n >= 0; --n) r = fma(r, x, c[n]);
  return r;
}

// Bopp difference of a polynomial V of degree <= kMaxDeg as an odd polynomial in u:
// V(x - u) - V(x + u) = -2 sum_{k odd} D_k(x) u^k, with the Taylor coefficients
// D_k(x) = V^(k)(x) / k! from a Taylor shift of the coefficients c (evaluated once per
// column; each cell then costs four FMAs).  Scaled by f.
struct BoppOdd {
  double e1, e3, e5, e7;
  __device__ __forceinline__ double operator()(double u) const {
    const double u2 = u * u;
    return u * fma(u2, fma(u2, fma(u2, e7, e5), e3), e1);
  }
};
template <class CP>
__device__ __forceinline__ BoppOdd bopp_odd(CP c, int deg, double x, double f) {
  static_assert(kMaxDeg == 8, "odd Taylor terms u, u^3, u^5, u^7");
  double d[kMaxDeg + 1];
#pragma unroll
  for (int n = 0; n <= kMaxDeg; ++n) d[n] = n <= deg ? c[n] : 0.0;
#pragma unroll
  for (int i = 0; i < kMaxDeg; ++i) {
#pragma unroll
    for (int k = kMaxDeg - 1; k >= i; --k) d[k] = fma(x, d[k + 1], d[k]);
  }
  const double g = -2.0 * f;
  return {g * d[1], g * d[3], g * d[5], g * d[7]};
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
  return apply(v, cplx<double>{cmem(s)[0], cmem(s)[1]});
}

// element of a uniform row base at a per-lane 32-bit byte offset (SGPR base +
// VGPR offset addressing: one offset register per lane for all strided elements)
template <typename T>
__device__ __forceinline__ cplx<T>& at_off(cplx<T>* base, unsigned boff) {
  return *reinterpret_cast<cplx<T>*>(reinterpret_cast<char*>(base) + boff);
}
template <typename T>
__device__ __forceinline__ gptr<cplx<T>> at_off(gptr<cplx<T>> base, unsigned boff) {
  return reinterpret_cast<gptr<cplx<T>>>(reinterpret_cast<gptr<char>>(base) + boff);
}

template <typename T>
__device__ __forceinline__ cplx<T> mul_i(cplx<T> a, T s) { return {-s * a.y, s * a.x}; }

// Streaming store of a state element: the sweep writes each element once and the next
// sweep reads it from another CU, so the line is not kept dirty in this CU's L2 (less
// write-back at the end of the kernel).
template <typename S>
__device__ __forceinline__ void store_stream(gptr<cplx<S>> p, cplx<S> v) {
  typedef S vec2 __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(vec2{v.x, v.y}, reinterpret_cast<gptr<vec2>>(p));
}

// Write-back store of a state element, for sweeps whose store instruction covers half of
// each 128-byte line (the p sweep's C-column rows): the neighbouring tile, on the same XCD at
// about the same time, fills the other half in L2 before the line is written back.  With
// streaming stores the half lines went out separately (+39% WRITE_SIZE on a 64 x 1024^2
// FP64 batch; +1.2% with this store).
template <typename S>
__device__ __forceinline__ void store_cached(gptr<cplx<S>> p, cplx<S> v) {
  typedef S vec2 __attribute__((ext_vector_type(2)));
  *reinterpret_cast<gptr<vec2>>(p) = vec2{v.x, v.y};
}

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

// ------------------------------------------------------------------ observables
// Sum over the 64 lanes of a wave (fixed xor tree: deterministic).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// v from another lane by a DPP control (row = 16 lanes); lanes of rows outside ROWMASK get 0
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(unsigned(b)), CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(unsigned(b >> 32)), CTRL, ROWMASK, 0xF, false);
  return __longlong_as_double((static_cast<unsigned long long>(unsigned(hi)) << 32) | unsigned(lo));
}
// Sum over aligned groups of G lanes (G = 8, 16, 32 or 64) by VALU lane moves (no LDS
// traffic, fixed order: deterministic); the last lane of each group holds the sum.
//   quad_perm [1,0,3,2], [2,3,0,1]: quad sums;  row_half_mirror: sums of 8;  row_mirror:
//   sums of 16;  row_bcast15 into rows 1, 3: sums of 32 there;  row_bcast31 into rows 2,
//   3: the sum of 64 in row 3.
template <int G>
__device__ __forceinline__ double group_sum(double v) {
  static_assert(G == 8 || G == 16 || G == 32 || G == 64, "group of 8..64 lanes");
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  if constexpr (G >= 16) v += dpp_f64<0x140>(v);
  if constexpr (G >= 32) v += dpp_f64<0x142, 0xA>(v);
  if constexpr (G >= 64) v += dpp_f64<0x143, 0xC>(v);
  return v;
}

// Observable terms of one column from its sums s = (Re S0, Im S0, sum p W, sum p^2 W,
// sum min(W, 0)): the integrands of <x>, <x^2>, <p>, <p^2>, <xp>, <dV/dx>, <p dV/dx>,
// <x dV/dx>, energy, negative area (the per-step record of GPU_Wigner2D_FFT.py:1856-1876).
__device__ __forceinline__ void column_terms(const Member& m, int col, const double (&s)[kNCol],
                                             double (&o)[12]) {
  const double x = gmem(m.x_host)[col], dv = gmem(m.dv_host)[col], v0 = gmem(m.v0_host)[col];
  o[0] = s[0];
  o[1] = s[1];
  o[2] = x * s[0];
  o[3] = x * x * s[0];
  o[4] = s[2];
  o[5] = s[3];
  o[6] = x * s[2];
  o[7] = dv * s[0];
  o[8] = dv * s[2];
  o[9] = x * dv * s[0];
  o[10] = s[3] * m.inv_2m + v0 * s[0];
  o[11] = s[4];
}

// Record row, pending factor and step counter from the 12 observable sums of a member
// (the normalised state is W / n, n = dXdP sum W; observables Re(dXdP sum W A / n)).
__device__ __forceinline__ void write_record(const Member& m, const double (&r)[12], int renorm, int record,
                                             int step) {
  const double nr = m.dxdp * r[0], ni = m.dxdp * r[1];
  double f = 1.0, sr = 1.0, si = 0.0;
  if (renorm) {
    const double m2 = nr * nr + ni * ni;
    f = nr / m2;
    sr = nr / m2;
    si = -ni / m2;
  }
  gmem(m.scale)[0] = sr;
  gmem(m.scale)[1] = si;
  if (record) {
    const gptr<double> o = gmem(m.obs) + size_t(step - m.obs_base) * kNObs;
    const double d = m.dxdp * f;
    o[kObsX] = d * r[2];
    o[kObsX2] = d * r[3];
    o[kObsP] = d * r[4];
    o[kObsP2] = d * r[5];
    o[kObsXP] = d * r[6];
    o[kObsDV] = d * r[7];
    o[kObsPDV] = d * r[8];
    o[kObsXDV] = d * r[9];
    o[kObsH] = d * r[10];
    o[kObsNeg] = d * r[11];
    o[kObsNorm] = d * r[0];
    o[kObsRawNormRe] = nr;
    o[kObsRawNormIm] = ni;
  }
}

// kTilePart epilogue of a p-sweep tile: csum holds the kNCol sums of each of the tile's
// C columns (LDS, visible to every thread: written before a barrier); thread k < 12 forms
// observable term k of each column and sums them in column order into
// m.tilepart[k * X + tile].  Tile 0 of member 0 copies the step index to rec_step.
template <int BLOCK, int C>
__device__ __forceinline__ void write_tile_partial(const SweepArgs& a, const Member& m, int tile,
                                                   const double* csum, int step) {
  const int tid = threadIdx.x;
  if (tid < 12) {
    double t = 0.0;
    for (int c = 0; c < C; ++c) {
      double s[kNCol], o[12];
#pragma unroll
      for (int k = 0; k < kNCol; ++k) s[k] = csum[c * kNCol + k];
      column_terms(m, tile * C + c, s, o);
#pragma unroll
      for (int k = 0; k < 12; ++k)
        if (k == tid) t += o[k];
    }
    gmem(m.tilepart)[size_t(tid) * m.X + tile] = t;
  } else if (tid == 12 && tile == 0 && &m == a.members) {
    *gmem(a.rec_step) = step;  // the step index this sweep read at its start
  }
}

// kTileScale prologue of an x-sweep block: sums the tile partials of the member's last
// p sweep (every block in the same order, so all blocks derive the same factor) and
// returns the pending factor 1 / norm (or 1).  Tile 0 also forms all 12 sums, writes
// the record row and the factor; tile 0 of member 0 advances the step index.
// Contains __syncthreads(); red: kTileRedBytes of LDS outside the transform arrays.
constexpr int kTileRedBytes = 16 * 12 * 8;
template <int BLOCK>
__device__ __forceinline__ cplx<double> tile_scale(const SweepArgs& a, const Member& m, int tile, double* red) {
  constexpr int kWaves = BLOCK / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int pc = a.pt_block / (m.P / a.pt_R);  // columns per p-sweep tile
  const int ntiles = m.X / pc;
  const bool rec = tile == 0;
  const int nk = rec ? 12 : 2;
  const int rstep = rec ? *gmem(a.rec_step) : 0;  // issued with the partial loads
  double r[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) r[k] = 0.0;
  for (int i = tid; i < ntiles; i += BLOCK) {
#pragma unroll
    for (int k = 0; k < 12; ++k)
      if (k < nk) r[k] += gmem(m.tilepart)[size_t(k) * m.X + i];
  }
#pragma unroll
  for (int k = 0; k < 12; ++k)
    if (k < nk) r[k] = group_sum<64>(r[k]);
  if (lane == 63) {
#pragma unroll
    for (int k = 0; k < 12; ++k)
      if (k < nk) red[wave * 12 + k] = r[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    if (k < nk) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) t += red[w * 12 + k];
      r[k] = t;
    }
  }
  const double nr = m.dxdp * r[0], ni = m.dxdp * r[1];
  cplx<double> f{1.0, 0.0};
  if (a.fin_renorm) {
    const double m2 = nr * nr + ni * ni;
    f = {nr / m2, -ni / m2};
  }
  if (rec && tid == 0) {
    write_record(m, r, a.fin_renorm, a.fin_record, rstep);
    if (&m == a.members) *gmem(a.step) = rstep + 1;
  }
  return f;
}

// ------------------------------------------------------------------ x sweep
// Kinetic multiplier of element m (lambda index) in row p, linear form:
// exp(-i a m_s), a = dt*dlambda*(k_c0 + k_c1 p), m_s = t + js*TT with
// js = j (j < R/2) or j - R; computed as E_t * Z^js, E_t = exp(-i a t),
// Z = exp(-i a TT): two sincos per lane instead of R.
template <typename S, typename T, int N, int R, int BLOCK, bool TP, typename E>
__device__ __forceinline__ void xsweep_body(const SweepArgs& a, const Member& m, int tile,
                                            cplx<E>* lds) {
  constexpr int TT = N / R, ROWS = BLOCK / TT;
  constexpr int LSZ = Padded<N, R>::size;
  const int tid = threadIdx.x, team = tid / TT, t = tid % TT;
  const int row = tile * ROWS + team;
  // uniform member base + 32-bit per-lane byte offset (a state is at most 2048^2 x 16 B);
  // the store recomputes the offset from an opaque copy of the lane index, so that no
  // per-row address stays live (or is spilled) across the transforms
  const gptr<cplx<S>> W = gmem(reinterpret_cast<cplx<S>*>(m.W));
  const unsigned roff = (unsigned(row) * unsigned(N) + unsigned(t)) * unsigned(sizeof(cplx<S>));
  constexpr unsigned kJ = unsigned(TT * sizeof(cplx<S>));
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(m.twX_hi), reinterpret_cast<const cplx<T>*>(m.twX_lo)};
  LdsRow<E> view{lds + team * LSZ};
  // a row's team inside one wave: wave-local exchanges, twiddles loaded once into registers
  // (the throughput form reads them from the table)
  constexpr bool kWave = TT <= 64;
  using Sync = std::conditional_t<kWave, WaveSync, BlockSync>;

  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = to_arith<T>(gload(at_off(W, roff + unsigned(j) * kJ)));
  const auto twsrc = [&]() {
    if constexpr (kWave && !TP) {
      TwRegs<T, N, R> r;
      load_twiddles<T, N, R>(r, t, tw);
      return r;
    } else {
      return TwMem<T>{tw};
    }
  }();
  // The pending factor and the 1/N of the inverse transform ride on the kinetic
  // multiplier (the transforms are linear): one complex product per lane instead of 2R.
  cplx<double> g{1.0 / N, 0.0};
  if (a.flags & kTileScale) {
    const cplx<double> f = tile_scale<BLOCK>(a, m, tile, reinterpret_cast<double*>(lds + lds_elems<BLOCK, R>()));
    g = {f.x * (1.0 / N), f.y * (1.0 / N)};
  } else if (a.flags & kScaleOnLoad) {
    g = {cmem(m.scale)[0] * (1.0 / N), cmem(m.scale)[1] * (1.0 / N)};
  }
  team_fft<T, N, R, -1, LdsRow<E>, Sync>(v, t, view, twsrc);

  const double p = m.dp * double(fidx(row, m.P));
  const double sgn = (a.flags & kConj) ? -1.0 : 1.0;  // adjoint: conj(r e^{-i phi}) = r e^{+i phi}
  if (a.kin_linear) {
    const double ak = sgn * a.dt * m.dlambda * (a.k_c0 + a.k_c1 * p);
    const cplx<double> Et = cmul(expmi(1.0, ak * double(t)), g);
    const cplx<double> Z1 = expmi(1.0, ak * double(TT));
    cplx<double> Zp[R / 2 + 1];  // Z^0 .. Z^(R/2)
    Zp[0] = {1.0, 0.0};
    Zp[1] = Z1;
#pragma unroll
    for (int k = 2; k <= R / 2; ++k) Zp[k] = cmul(Zp[k / 2], Zp[k - k / 2]);
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = apply(v[j], cmul(Et, j < R / 2 ? Zp[j] : cplx<double>{Zp[R - j].x, -Zp[R - j].y}));
    if (m.lambda_decay) {
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = apply(v[j], cplx<double>{gmem(m.lambda_decay)[t + j * TT], 0.0});
    }
  } else {
    const double hh = a.hbar_half;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int mm = t + j * TT;
      const double lam = m.dlambda * double(fidx(mm, N));
      const double phi = sgn * a.dt * (poly(a.kcoef, a.kdeg, p + hh * lam) - poly(a.kcoef, a.kdeg, p - hh * lam)) * a.inv_hbar;
      const double r = m.lambda_decay ? gmem(m.lambda_decay)[mm] : 1.0;
      v[j] = apply(v[j], cmul(expmi(r, phi), g));
    }
  }
  team_fft<T, N, R, +1, LdsRow<E>, Sync>(v, t, view, twsrc);
  const int ts = opaque_v(int(threadIdx.x));  // offset recomputed from the lane index
  const unsigned soff = (unsigned(tile * ROWS + ts / TT) * unsigned(N) + unsigned(ts % TT)) * unsigned(sizeof(cplx<S>));
#pragma unroll
  for (int j = 0; j < R; ++j) store_stream(at_off(W, soff + unsigned(j) * kJ), to_arith<S>(v[j]));

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
      gmem(m.colpart)[size_t(tile) * N + x] = s;
    }
  }
}

// ------------------------------------------------------------------ p sweep
template <typename T, int N, int R, int BLOCK, typename E = T>
struct PSweep {
  static constexpr int TT = N / R;
  static constexpr int C = BLOCK / TT;  // columns per tile
  static constexpr int LSZ = Padded<N, R>::size;
  // wave teams: a column's team lies inside one wave and the transform has an exchange
  // at which the lanes can leave the load layout (then every later exchange is wave-local)
  static constexpr bool kWave = TT <= 64 && num_passes(N, R) >= 2;
  using View = std::conditional_t<kWave, LdsRow<E>, LdsStrided<E, C>>;
  using Sync = std::conditional_t<kWave, WaveSync, BlockSync>;

  // potential step multiplier of time index `step` (expPotential; GP phase optional),
  // times the real factor `scale` (the 1/N of the inverse transform that follows).
  // tab: which [P][X] table a member with an expression program reads (0: time index `step`,
  // 1: `step` + 1, the merged Strang half step; expr_table_kernel fills them before the sweep)
  __device__ __forceinline__ static void potential(cplx<T> (&v)[R], const SweepArgs& a,
                                                   const Member& m, int t_in, int col_in, int step,
                                                   double scale, int tab) {
    // opaque lane indices: recomputing theta and x -/+ hbar theta/2 per call is cheaper
    // than keeping them live across the transforms between two calls
    const int t = opaque_v(t_in), col = opaque_v(col_in);
    const double x = m.dx * double(fidx(col, m.X));
    const int row = step < m.pot_rows ? step : m.pot_rows - 1;
    const cptr<double> cf = cmem(m.pot) + size_t(row) * (kMaxDeg + 1);
    const double mask = (m.xmask ? gmem(m.xmask)[col] : 1.0) * scale;
    const bool gp = (a.flags & kGP) != 0;
    const BoppOdd bopp = bopp_odd(cf, m.pot_deg, x, a.vphase);
    // uniform conditions outside the element loops, so that each loop is one basic block
    // and no per-element array stays live across it
    const double hd = a.hbar_half * m.dtheta;
    auto mult = [&](int i) { return expmi(mask, bopp(hd * double(fidx(i, N)))); };
    if (kWave && m.pot_deg >= 0 && !gp) {
      // theta_{N-i} = -theta_i and the phase is odd in u, so the multiplier of N - i is the
      // exact conjugate of that of i (the reduction and the odd/even series negate exactly):
      // lane t evaluates j < R/2 (i < N/2) and takes j >= R/2 from lane TT - t of its team
      // (element R-1-j there); lane 0 pairs with itself (element R-j) and evaluates the
      // self-paired i = N/2 directly.
      cplx<double> M[R / 2];
#pragma unroll
      for (int j = 0; j < R / 2; ++j) M[j] = mult(t + j * TT);
      const cplx<double> nyq = mult(N / 2);
      const int lane = threadIdx.x & 63;
      const int src = lane - t + ((TT - t) & (TT - 1));
      static_for<R / 2, R>([&](auto J) {
        constexpr int j = decltype(J)::value;
        cplx<double> q{__shfl(M[R - 1 - j].x, src, 64), __shfl(M[R - 1 - j].y, src, 64)};
        if constexpr (j == R / 2) {
          if (t == 0) q = {nyq.x, -nyq.y};  // own multiplier (conjugated back below)
        } else {
          if (t == 0) q = M[R - j];
        }
        v[j] = apply(v[j], cplx<double>{q.x, -q.y});
      });
#pragma unroll
      for (int j = 0; j < R / 2; ++j) v[j] = apply(v[j], M[j]);
    } else if (m.pot_deg >= 0 && !gp) {
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = apply(v[j], mult(t + j * TT));
    } else {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int i = t + j * TT;
        double phi;
        if (m.pot_deg >= 0) {
          phi = bopp(hd * double(fidx(i, N)));
        } else {
          const size_t toff = (tab && m.pot_expr_len > 0) ? size_t(m.X) * size_t(m.P) : 0;
          phi = a.vphase * gmem(m.pot_dv)[toff + size_t(i) * m.X + col];
        }
        if (gp) phi -= a.gp_phase * gmem(m.G)[size_t(i) * m.X + col];
        v[j] = apply(v[j], expmi(mask, phi));
      }
    }
    if (m.theta_decay) {
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = apply(v[j], cplx<double>{gmem(m.theta_decay)[t + j * TT], 0.0});
    }
  }

  // D(Y) = F_p^{-1}[ theta * F_p[ sigma(p) Y ] ]   (Theta_fp_Damping, FFT.py:1135-1143);
  // the 1/N of the inverse transform is folded into theta (a power of two: exact)
  template <class Tw>
  __device__ __forceinline__ static void cl_operator(cplx<T> (&y)[R], const Member& m, int t,
                                                     View view, const Tw& tw) {
#pragma unroll
    for (int j = 0; j < R; ++j) y[j] = cscale(y[j], T(gmem(m.sigma_p)[t + j * TT]));
    team_fft<T, N, R, -1, View, Sync>(y, t, view, tw);
#pragma unroll
    for (int j = 0; j < R; ++j) y[j] = cscale(y[j], T(gmem(m.theta_host)[t + j * TT] * (1.0 / N)));
    team_fft<T, N, R, +1, View, Sync>(y, t, view, tw);
  }
};

// LDS bytes past the transform arrays of a p-sweep block: per-column sums, then the
// per-column observable terms (<= 64 columns)
constexpr int kPTailBytes = 64 * 12 * 8;

// Lanes: the load layout (column cl = tid % C, team index tl = tid / C) makes every
// global load and store instruction cover rows of C consecutive columns.  With wave
// teams the first exchange of the first transform moves each column to one team of
// TT consecutive lanes (column cc = tid / TT, team index tc = tid % TT) and the store
// moves it back through LDS; otherwise the load layout is kept throughout.
template <typename S, typename T, int N, int R, int BLOCK, bool CL, bool TP, typename E>
__device__ __forceinline__ void psweep_body(const SweepArgs& a, const Member& m, int tile,
                                            cplx<E>* lds) {
  using PS = PSweep<T, N, R, BLOCK, E>;
  using View = typename PS::View;
  using Sync = typename PS::Sync;
  constexpr int TT = PS::TT, C = PS::C, LSZ = PS::LSZ;
  constexpr bool kWave = PS::kWave;
  const int tid = threadIdx.x;
  const int cl = tid % C, tl = tid / C;
  const int cc = kWave ? tid / TT : cl, tc = kWave ? tid % TT : tl;
  const int X = m.X;
  const int col = tile * C + cc;
  const gptr<cplx<S>> W = gmem(reinterpret_cast<cplx<S>*>(m.W));
  const Twiddles<T> tw{reinterpret_cast<const cplx<T>*>(m.twP_hi), reinterpret_cast<const cplx<T>*>(m.twP_lo)};
  const View view_c{kWave ? lds + cc * LSZ : lds + cl};
  const int step = (a.flags & kStepArg) ? a.step_add : *gmem(a.step) + a.step_add;

  // 32-bit byte offsets from the member's base (a state is at most 2048^2 x 16 B); the
  // store recomputes them from an opaque copy of the lane index, so that no address
  // stays live (or is spilled) across the transforms
  const unsigned boff = (unsigned(tl) * unsigned(X) + unsigned(tile * C + cl)) * unsigned(sizeof(cplx<S>));
  const unsigned jstride = unsigned(TT) * unsigned(X) * unsigned(sizeof(cplx<S>));
  cplx<T> v[R];
#pragma unroll
  for (int j = 0; j < R; ++j) v[j] = to_arith<T>(gload(at_off(W, boff + unsigned(j) * jstride)));
  const auto twsrc = [&]() {
    if constexpr (kWave && !TP) {
      TwRegs<T, N, R> r;
      load_twiddles<T, N, R>(r, tc, tw);
      return r;
    } else {
      return TwMem<T>{tw};
    }
  }();
  if (a.flags & kScaleOnLoad) {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = scale_load(v[j], m.scale);
  }
  if constexpr (kWave) {
    team_fft_remap<T, N, R, -1>(v, tl, LdsRow<E>{lds + cl * LSZ}, tc, view_c, twsrc);
  } else {
    team_fft<T, N, R, -1, View, Sync>(v, tc, view_c, twsrc);
  }
  PS::potential(v, a, m, tc, col, step, 1.0 / N, 0);
  team_fft<T, N, R, +1, View, Sync>(v, tc, view_c, twsrc);

  if constexpr (CL) {
    // W + 2aD W + (10/3)a^2 D^2 W + (4/3)a^3 D^3 W, a = i dt gamma, built as
    // LW = W + (2/3)a D W; LW += a D LW; W += 2a D LW   (FFT.py:1145-1157)
    cplx<T> lw[R], d[R];
#pragma unroll
    for (int j = 0; j < R; ++j) d[j] = v[j];
    PS::cl_operator(d, m, tc, view_c, twsrc);
    const T c1 = T((2.0 / 3.0) * a.cl_a), c2 = T(a.cl_a), c3 = T(2.0 * a.cl_a);
#pragma unroll
    for (int j = 0; j < R; ++j) lw[j] = cadd(v[j], mul_i(d[j], c1));
#pragma unroll
    for (int j = 0; j < R; ++j) d[j] = lw[j];
    PS::cl_operator(d, m, tc, view_c, twsrc);
#pragma unroll
    for (int j = 0; j < R; ++j) lw[j] = cadd(lw[j], mul_i(d[j], c2));
#pragma unroll
    for (int j = 0; j < R; ++j) d[j] = lw[j];
    PS::cl_operator(d, m, tc, view_c, twsrc);
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = cadd(v[j], mul_i(d[j], c3));
  }

  // per-column sums of the tile in the LDS tail: csum[c * kNCol + k] for column tile*C + c
  double* csum = reinterpret_cast<double*>(reinterpret_cast<char*>(lds) + lds_elems<BLOCK, R>() * sizeof(cplx<E>));
  // column sums to global memory, after a block barrier that follows their LDS write
  const auto publish_colsums = [&]() {
    if (tid < C) {
#pragma unroll
      for (int k = 0; k < kNCol; ++k) gmem(m.colsums)[size_t(tile * C + tid) * kNCol + k] = csum[tid * kNCol + k];
    }
  };
  bool colsums_pending = false;
  if (a.flags & kColSums) {
    double acc[kNCol] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const double p = gmem(m.p_host)[tc + j * TT];
      const double w = double(v[j].x);
      acc[0] += w;
      acc[1] += double(v[j].y);
      acc[2] += p * w;
      acc[3] += p * p * w;
      acc[4] += w < 0.0 ? w : 0.0;
    }
    if constexpr (kWave) {
      // a column's lanes are TT consecutive lanes of one wave; the last one gets the sums
#pragma unroll
      for (int k = 0; k < kNCol; ++k) acc[k] = group_sum<TT>(acc[k]);
      if (tc == TT - 1) {
#pragma unroll
        for (int k = 0; k < kNCol; ++k) csum[cc * kNCol + k] = acc[k];
      }
      if (a.flags & kMerge) {
        __syncthreads();
        publish_colsums();
      } else {
        colsums_pending = true;  // published after the store's block barrier
      }
    } else {
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
        // C <= 64: every wave holds all C columns
        double sc[kNCol] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < kWaves; ++w) {
#pragma unroll
          for (int k = 0; k < kNCol; ++k) sc[k] += red[(w * kLanes + tid) * kNCol + k];
        }
#pragma unroll
        for (int k = 0; k < kNCol; ++k) {
          gmem(m.colsums)[size_t(tile * C + tid) * kNCol + k] = sc[k];
          csum[tid * kNCol + k] = sc[k];
        }
      }
      __syncthreads();  // csum complete for the tile partials; red free for the transforms
    }
  }

  if (a.flags & kMerge) {
    team_fft<T, N, R, -1, View, Sync>(v, tc, view_c, twsrc);
    PS::potential(v, a, m, tc, col, step + 1, 1.0 / N, 1);
    team_fft<T, N, R, +1, View, Sync>(v, tc, view_c, twsrc);
  }

  if constexpr (kWave) {
    // back to the load layout: natural order through the team's LDS array
    Sync::sync();
#pragma unroll
    for (int j = 0; j < R; ++j) *view_c.ptr(padr<R>(tc + j * TT)) = ccvt<E>(v[j]);
    __syncthreads();
    if (colsums_pending) publish_colsums();
    const LdsRow<E> view_l{lds + cl * LSZ};
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = ccvt<T>(*view_l.ptr(padr<R>(tl + j * TT)));
  }
  const int ts = opaque_v(int(threadIdx.x));  // offset recomputed from the lane index
  const unsigned soff = (unsigned(ts / C) * unsigned(X) + unsigned(tile * C + ts % C)) * unsigned(sizeof(cplx<S>));
#pragma unroll
  for (int j = 0; j < R; ++j) store_cached(at_off(W, soff + unsigned(j) * jstride), to_arith<S>(v[j]));
  if (a.flags & kTilePart) write_tile_partial<BLOCK, C>(a, m, tile, csum, step - a.step_add);
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
  for (int j = 0; j < R; ++j) v[j] = {gmem(m.rho)[t + j * TT], 0.0};
  team_fft<double, N, R, -1>(v, t, view, tw);
  const double half_theta = 0.5 * cmem(m.theta_host)[row];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    double sn, cs;
    sincos_fp64(gmem(m.lambda_host)[t + j * TT] * half_theta, &sn, &cs);
    const double f = 2.0 * sn;  // (-2i sin) * (a + ib) = 2 sin b - 2i sin a
    v[j] = {f * v[j].y, -f * v[j].x};
  }
  team_fft<double, N, R, +1>(v, t, view, tw);
#pragma unroll
  for (int j = 0; j < R; ++j) gmem(m.G)[size_t(row) * N + t + j * TT] = v[j].x * (1.0 / N);
}

// ------------------------------------------------------------------ kernels
// sizes up to MAXN.  One launch covers members of every size: the throughput forms (sides
// up to 1024) spill a few VGPRs at three waves per SIMD, in the general-polynomial kinetic
// branch, and still run an uneven 128..1024 ensemble 2% faster than one launch per size
// (and a uniform batch within 0.5%)
#define WH_DISPATCH_TO(NVAR, BODY, MAXN)                                                     \
  switch (NVAR) {                                                                            \
    case 128: if constexpr (size_ok<128, R, BLOCK>() && 128 <= (MAXN)) { BODY(128); } break;   \
    case 256: if constexpr (size_ok<256, R, BLOCK>() && 256 <= (MAXN)) { BODY(256); } break;   \
    case 512: if constexpr (size_ok<512, R, BLOCK>() && 512 <= (MAXN)) { BODY(512); } break;   \
    case 1024: if constexpr (size_ok<1024, R, BLOCK>() && 1024 <= (MAXN)) { BODY(1024); } break; \
    case 2048: if constexpr (size_ok<2048, R, BLOCK>() && 2048 <= (MAXN)) { BODY(2048); } break; \
    default: break;                                                                          \
  }
#define WH_DISPATCH(NVAR, BODY) WH_DISPATCH_TO(NVAR, BODY, 2048)

// Two waves per SIMD (<= 256 VGPRs) for the plain sweeps: a batch of grids overlaps one
// wave's loads and stores with the other's transforms; three for the throughput form
// (TP: table twiddles, exchanges in the storage precision).  The Caldeira-Leggett p sweep
// holds three state copies and keeps the full register file of one wave.
template <typename S, typename T, bool TP>
using LdsElem = std::conditional_t<TP, S, T>;
template <typename S, typename T, int R, int BLOCK, bool TP>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(TP ? 3 : 2))) xsweep_kernel(SweepArgs a) {
  using E = LdsElem<S, T, TP>;
  __shared__ cplx<E> lds[lds_elems<BLOCK, R>() + kTileRedBytes / sizeof(cplx<E>)];
  const unsigned w = a.work[blockIdx.x];
  const Member& m = a.members[w >> 16];
  const int tile = int(w & 0xffffu);
#define BODY(n) xsweep_body<S, T, n, R, BLOCK, TP, E>(a, m, tile, lds)
  WH_DISPATCH_TO(m.X, BODY, TP ? kTpMaxN : 2048)
#undef BODY
}

template <typename S, typename T, int R, int BLOCK, bool CL, bool TP>
__device__ __forceinline__ void psweep_entry(const SweepArgs& a) {
  using E = LdsElem<S, T, TP>;
  __shared__ cplx<E> lds[lds_elems<BLOCK, R>() + kPTailBytes / sizeof(cplx<E>)];
  const unsigned w = a.work[blockIdx.x];
  const Member& m = a.members[w >> 16];
  const int tile = int(w & 0xffffu);
#define BODY(n) psweep_body<S, T, n, R, BLOCK, CL, TP, E>(a, m, tile, lds)
  WH_DISPATCH_TO(m.P, BODY, TP && !CL ? kTpMaxN : 2048)
#undef BODY
}
template <typename S, typename T, int R, int BLOCK, bool CL, bool TP>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(CL ? 1 : TP ? 3 : 2))) psweep_kernel(SweepArgs a) {
  psweep_entry<S, T, R, BLOCK, CL, TP>(a);
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
