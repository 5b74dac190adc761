// Device-side FFT building blocks for gfx950 (CDNA4), wave64.
//
// A "team" of TT = N/R lanes computes one length-N complex transform.  Lane t
// holds R elements in registers in the strided layout  v[j] = a[t + j*TT].
// The transform is a Stockham autosort sequence of radix passes [R, R, ..., r]
// (r = N / R^k).  The first pass works on the registers directly, every later
// pass exchanges through LDS, and the last pass writes back into the same
// strided register layout.  So a forward transform, a pointwise multiply and an
// inverse transform chain in registers, and global loads and stores of the
// strided layout are coalesced for row teams and column teams alike.
//
// Pass with radix r after the span Ns = prod(previous radices):
//   butterfly b in [0, N/r) reads a[b + j*N/r], twiddles by w^(j*(b mod Ns)),
//   w = exp(dir*2*pi*i/(Ns*r)), applies an r-point DFT and writes
//   a'[(b - b mod Ns)*r + (b mod Ns) + q*Ns].
// Twiddles come from a table tw[n] = exp(-2*pi*i*n/N) computed in FP64 on the
// host.  Forward = exponent -i (dir = -1), inverse = +i and unnormalised, as
// cufftExecZ2Z (cufft_wrapper.py:11-12, 81-103).
//
// FP32: every constant the transform multiplies by (table twiddles and the
// in-register radix-16 roots) is carried as a hi + lo float pair, so the
// transform's operator is exact to ~2^-48 and only value-dependent rounding
// remains: a fixed perturbation of the step operator would accumulate linearly
// over thousands of unitary steps, rounding noise only as a random walk.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace wh {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

// Address-space qualified pointers.  Pointers that reach device code through memory
// (tables, member records) are generic, for which the compiler emits flat_ instructions;
// those also count against the LDS counter, so every LDS wait of an exchange would
// wait for outstanding memory loads as well.  gptr: global memory (global_ loads and
// stores).  cptr: read-only for the whole kernel (uniform addresses load through the
// scalar cache; never written through).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> gmem(T* p) {
  return (gptr<T>)p;
}
template <typename T>
__device__ __forceinline__ cptr<T> cmem(const T* p) {
  return (cptr<T>)p;
}
// complex element through a global pointer (one 2-vector load)
template <typename T>
__device__ __forceinline__ cplx<T> gload(gptr<const cplx<T>> p) {
  typedef T vec2 __attribute__((ext_vector_type(2)));
  const vec2 r = *reinterpret_cast<gptr<const vec2>>(p);
  return {r.x, r.y};
}
template <typename T>
__device__ __forceinline__ cplx<T> gload(gptr<cplx<T>> p) {
  return gload<T>(gptr<const cplx<T>>(p));
}

template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return {a.x * s, a.y * s}; }

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }
// number of Stockham passes for N with maximal radix R
constexpr int num_passes(int N, int R) { return N <= R ? 1 : 1 + num_passes(N / R, R); }
// radix of pass p (all R except possibly the last)
constexpr int pass_radix(int N, int R, int p) {
  return (p == num_passes(N, R) - 1) ? N / ipow(R, p) : R;
}
constexpr int bitrev(int v, int bits) {
  return bits == 0 ? 0 : ((v & 1) << (bits - 1)) | bitrev(v >> 1, bits - 1);
}

// cos(2*pi*k/16), k = 0..15
constexpr double cos16(int k) {
  constexpr double c1 = 0.92387953251128675613;  // cos(pi/8)
  constexpr double c2 = 0.70710678118654752440;  // cos(pi/4)
  constexpr double c3 = 0.38268343236508977173;  // cos(3pi/8)
  const double t[16] = {1.0, c1, c2, c3, 0.0, -c3, -c2, -c1, -1.0, -c1, -c2, -c3, 0.0, c3, c2, c1};
  return t[k & 15];
}

// a * (c + i s) with constants known at compile time; FP32 adds the products
// with the low parts of the constants.
template <typename T>
struct ConstMul;
template <>
struct ConstMul<double> {
  template <int K, int DIR>
  __device__ __forceinline__ static cplx<double> mul(cplx<double> a) {
    constexpr double C = cos16(K), S = (DIR < 0 ? -1.0 : 1.0) * cos16(K + 12);
    return {a.x * C - a.y * S, a.x * S + a.y * C};
  }
};
template <>
struct ConstMul<float> {
  template <int K, int DIR>
  __device__ __forceinline__ static cplx<float> mul(cplx<float> a) {
    constexpr double C = cos16(K), S = (DIR < 0 ? -1.0 : 1.0) * cos16(K + 12);
    constexpr float ch = float(C), sh = float(S);
    constexpr float cl = float(C - double(ch)), sl = float(S - double(sh));
    const float rx = fmaf(a.x, ch, -a.y * sh), ry = fmaf(a.x, sh, a.y * ch);
    return {rx + fmaf(a.x, cl, -a.y * sl), ry + fmaf(a.x, sl, a.y * cl)};
  }
};

// a * exp(DIR * 2*pi*i * K / M), M | 16; the special angles are exact.
template <int DIR, int K, int M, typename T>
__device__ __forceinline__ cplx<T> rot(cplx<T> a) {
  constexpr int k = ((K % M) * (16 / M)) & 15;
  if constexpr (k == 0) {
    return a;
  } else if constexpr (k == 8) {
    return {-a.x, -a.y};
  } else if constexpr ((k == 4 && DIR < 0) || (k == 12 && DIR > 0)) {
    return {a.y, -a.x};  // * (-i)
  } else if constexpr (k == 4 || k == 12) {
    return {-a.y, a.x};  // * (+i)
  } else {
    return ConstMul<T>::template mul<k, DIR>(a);
  }
}

// Radix-2 decimation-in-frequency butterflies of one stage of length L at offset OFF.
template <int DIR, int L, int OFF, int I, typename T>
struct DifStage {
  __device__ __forceinline__ static void run(cplx<T>* v) {
    if constexpr (I < L / 2) {
      cplx<T> a = v[OFF + I], b = v[OFF + I + L / 2];
      v[OFF + I] = cadd(a, b);
      v[OFF + I + L / 2] = rot<DIR, I, L>(csub(a, b));
      DifStage<DIR, L, OFF, I + 1, T>::run(v);
    }
  }
};
template <int DIR, int L, int OFF, typename T>
struct Dif {
  __device__ __forceinline__ static void run(cplx<T>* v) {
    if constexpr (L > 1) {
      DifStage<DIR, L, OFF, 0, T>::run(v);
      Dif<DIR, L / 2, OFF, T>::run(v);
      Dif<DIR, L / 2, OFF + L / 2, T>::run(v);
    }
  }
};

// In-register r-point DFT:  y_q = sum_j x_j exp(DIR*2*pi*i*j*q/r), r | 16.
template <int DIR, int r, typename T>
__device__ __forceinline__ void dft(cplx<T>* x) {
  Dif<DIR, r, 0, T>::run(x);
  cplx<T> y[r];
#pragma unroll
  for (int q = 0; q < r; ++q) y[q] = x[bitrev(q, ilog2(r))];
#pragma unroll
  for (int q = 0; q < r; ++q) x[q] = y[q];
}

// LDS index with one pad slot every PR entries (PR = the team's radix): spreads
// the radix-strided writes of a pass over the banks.
template <int PR>
__device__ __forceinline__ int padr(int i) { return i + i / PR; }
template <int N, int PR>
struct Padded {
  static constexpr int size = N + N / PR + 1;
};

// LDS views of one team's working array.  ptr(pb) is the address of padded
// index pb; S is the element stride between consecutive padded indices.  The element
// type E may be narrower than the arithmetic type: exchanged values are then rounded
// to E (complex64 exchanges of an FP64 transform halve the LDS bytes).
template <typename E>
struct LdsRow {  // contiguous: a[i] at base[padr(i)]
  using elem = E;
  static constexpr int S = 1;
  cplx<E>* base;
  __device__ __forceinline__ cplx<E>* ptr(int pb) const { return base + pb; }
};
template <typename E, int LS>
struct LdsStrided {  // interleaved columns: a[i] at base[padr(i) * LS]
  using elem = E;
  static constexpr int S = LS;
  cplx<E>* base;
  __device__ __forceinline__ cplx<E>* ptr(int pb) const { return base + pb * LS; }
};

// complex conversion between arithmetic and LDS element types (identity when equal)
template <typename D, typename T>
__device__ __forceinline__ cplx<D> ccvt(cplx<T> v) {
  if constexpr (std::is_same_v<D, T>) {
    return v;
  } else {
    return {D(v.x), D(v.y)};
  }
}

// Hides the value of an expression from the optimiser, so that per-lane
// addresses and indices of one transform are not merged with (and kept live
// until) those of the next transform in the same kernel.
template <typename P>
__device__ __forceinline__ P opaque_s(P p) {
  asm volatile("" : "+s"(p));
  return p;
}
__device__ __forceinline__ int opaque_v(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Twiddle tables: FP64 hi only; FP32 hi and lo (the rounding error of hi).
template <typename T>
struct Twiddles {
  const cplx<T>* hi;
  const cplx<T>* lo;
};

// x * w (DIR < 0) or x * conj(w) (DIR > 0); FP32 adds the product with the low part l.
template <int DIR>
__device__ __forceinline__ cplx<double> tw_apply(cplx<double> x, cplx<double> w) {
  if (DIR > 0) w.y = -w.y;
  return cmul(x, w);
}
template <int DIR>
__device__ __forceinline__ cplx<float> tw_apply(cplx<float> x, cplx<float> w, cplx<float> l) {
  if (DIR > 0) {
    w.y = -w.y;
    l.y = -l.y;
  }
  const float rx = fmaf(x.x, w.x, -x.y * w.y), ry = fmaf(x.x, w.y, x.y * w.x);
  return {rx + fmaf(x.x, l.x, -x.y * l.y), ry + fmaf(x.x, l.y, x.y * l.x)};
}

template <int DIR>
__device__ __forceinline__ cplx<double> tw_mul(cplx<double> x, const Twiddles<double>& tw, int idx) {
  return tw_apply<DIR>(x, gload(gmem(tw.hi) + idx));
}
template <int DIR>
__device__ __forceinline__ cplx<float> tw_mul(cplx<float> x, const Twiddles<float>& tw, int idx) {
  return tw_apply<DIR>(x, gload(gmem(tw.hi) + idx), gload(gmem(tw.lo) + idx));
}

// Twiddles of every pass of one lane, held in registers for the whole kernel (a lane's
// twiddles depend only on its team index t, so all transforms of a sweep share them and
// the inverse uses their conjugates).  Only butterfly mm = 0 of a pass is stored: the
// passes before the last have one butterfly per lane (r = R), and in the last pass
// (span Ns = N / r, twiddle w^(j b), b = t + mm TT) butterfly mm differs from mm = 0 by
// the exact constant w^(j mm TT) = exp(-2 pi i j mm / R).  Slot of (pass P, input j >= 1):
// off(P) + j - 1.
template <int N, int R>
struct TwSlots {
  static constexpr int NP = num_passes(N, R);
  static constexpr int off(int P) {
    int o = 0;
    for (int q = 1; q < P; ++q) o += pass_radix(N, R, q) - 1;
    return o;
  }
  static constexpr int count = off(NP) > 0 ? off(NP) : 1;
};

template <typename T, int N, int R>
struct TwRegs;
template <int N, int R>
struct TwRegs<double, N, R> {
  cplx<double> hi[TwSlots<N, R>::count];
  template <int DIR>
  __device__ __forceinline__ cplx<double> mul(cplx<double> x, int slot, int) const {
    return tw_apply<DIR>(x, hi[slot]);
  }
  __device__ __forceinline__ void set(int slot, const Twiddles<double>& tw, int idx) { hi[slot] = gload(gmem(tw.hi) + idx); }
};
template <int N, int R>
struct TwRegs<float, N, R> {
  cplx<float> hi[TwSlots<N, R>::count], lo[TwSlots<N, R>::count];
  template <int DIR>
  __device__ __forceinline__ cplx<float> mul(cplx<float> x, int slot, int) const {
    return tw_apply<DIR>(x, hi[slot], lo[slot]);
  }
  __device__ __forceinline__ void set(int slot, const Twiddles<float>& tw, int idx) {
    hi[slot] = gload(gmem(tw.hi) + idx);
    lo[slot] = gload(gmem(tw.lo) + idx);
  }
};

// Table-backed source with the same interface (twiddle index idx, slot unused).
template <typename T>
struct TwMem {
  Twiddles<T> tw;
  template <int DIR>
  __device__ __forceinline__ cplx<T> mul(cplx<T> x, int, int idx) const {
    return tw_mul<DIR>(x, tw, idx);
  }
};

// Fills the registers of team index t from the table.
template <typename T, int N, int R, int P = 1>
__device__ __forceinline__ void load_twiddles(TwRegs<T, N, R>& w, int t, const Twiddles<T>& tw) {
  if constexpr (P < num_passes(N, R)) {
    constexpr int r = pass_radix(N, R, P), Ns = ipow(R, P);
    const int tstep = (t % Ns) * (N / (Ns * r));
#pragma unroll
    for (int j = 1; j < r; ++j) w.set(TwSlots<N, R>::off(P) + j - 1, tw, (j * tstep) & (N - 1));
    load_twiddles<T, N, R, P + 1>(w, t, tw);
  }
}

// Calls f(std::integral_constant<int, i>) for i = B .. E-1.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Ordering of a pass's LDS write and read.  BlockSync: the team spans several waves.
// WaveSync: every lane of the team is in one wave, whose LDS operations execute in
// program order, so only the compiler has to keep the order (no s_barrier); the
// fences name the LDS address space only, so global loads (tables, twiddles) may
// still be scheduled across an exchange.
struct BlockSync {
  __device__ __forceinline__ static void sync() { __syncthreads(); }
};
struct WaveSync {
  __device__ __forceinline__ static void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  }
};

// One Stockham pass P >= 1 (LDS-sourced) through the team's LDS view.
//
// Every LDS address is a per-lane base plus a compile-time offset: spans after
// the first pass are multiples of R, so padr(i + R m) = padr(i) + (R+1) m, and
// the first pass (radix R, span 1) writes the R consecutive indices of one
// aligned group of R, so padr(R b + q) = padr(R b) + q.
template <typename T, int N, int R, int P>
struct PassIO {
  static constexpr int NP = num_passes(N, R);
  static constexpr int TT = N / R;
  // write the outputs of pass P-1 (team index t, strided register layout of that pass)
  template <class Lds>
  __device__ __forceinline__ static void write(const cplx<T> (&v)[R], int tl, Lds lds) {
    constexpr int rp = pass_radix(N, R, P - 1);  // radix of the previous pass
    constexpr int Nsp = ipow(R, P - 1);          // span before the previous pass
    constexpr int nbp = R / rp;
    constexpr int S = Lds::S;
    using E = typename Lds::elem;
    static_assert(Nsp % R == 0 || (Nsp == 1 && rp == R), "separable LDS offsets");
#pragma unroll
    for (int mm = 0; mm < nbp; ++mm) {
      const int b = tl + mm * TT;
      const int k = b % Nsp;
      cplx<E>* p = lds.ptr(padr<R>((b - k) * rp + k));
#pragma unroll
      for (int q = 0; q < rp; ++q) {
        constexpr int unit = Nsp == 1 ? 1 : Nsp / R * (R + 1);
        p[S * q * unit] = ccvt<E>(v[mm * rp + q]);
      }
    }
  }
  // read the inputs of pass P, twiddle, r-point DFTs; outputs in v
  template <int DIR, class Lds, class Tw>
  __device__ __forceinline__ static void read(cplx<T> (&v)[R], int tl, Lds lds, const Tw& tw) {
    constexpr int r = pass_radix(N, R, P);
    constexpr int Ns = ipow(R, P);
    constexpr int nb = R / r;
    constexpr int S = Lds::S;
    static_assert((N / r) % R == 0, "separable LDS offsets");
    static_assert(nb == 1 || P == NP - 1, "several butterflies per lane only in the last pass");
    static_assert(nb == 1 || Ns == N / r, "the last pass spans N / r");
    cplx<T> o[R];
    // twiddle of input j: w^(j k s), k = b mod Ns, s = N / (Ns r); for mm > 0 (last pass,
    // k = t + mm TT, s = 1) it is w^(j t) * exp(DIR 2 pi i j mm / R)
    const int tstep = (tl % Ns) * (N / (Ns * r));
    static_for<0, nb>([&](auto MM) {
      constexpr int mm = decltype(MM)::value;
      const auto* p = lds.ptr(padr<R>(tl + mm * TT));
      cplx<T> x[r];
#pragma unroll
      for (int j = 0; j < r; ++j) x[j] = ccvt<T>(p[S * j * ((N / r) / R * (R + 1))]);
      static_for<1, r>([&](auto J) {
        constexpr int j = decltype(J)::value;
        x[j] = tw.template mul<DIR>(x[j], TwSlots<N, R>::off(P) + j - 1, (j * tstep) & (N - 1));
        x[j] = rot<DIR, (j * mm) % R, R>(x[j]);
      });
      dft<DIR, r>(x);
      if constexpr (P == NP - 1) {
        // last pass: outputs land on t + TT*(mm + q*nb) -> strided register layout
#pragma unroll
        for (int q = 0; q < r; ++q) o[mm + q * nb] = x[q];
      } else {
#pragma unroll
        for (int q = 0; q < r; ++q) o[mm * r + q] = x[q];
      }
    });
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = o[j];
  }
};

template <typename T, int N, int R, int DIR, int P, class Sync, class Lds, class Tw>
struct TeamPass {
  __device__ __forceinline__ static void run(cplx<T> (&v)[R], int t, Lds lds, const Tw& tw) {
    if constexpr (P < num_passes(N, R)) {
      const int tl = opaque_v(t);
      Sync::sync();  // every lane has read the previous pass's inputs
      PassIO<T, N, R, P>::write(v, tl, lds);
      Sync::sync();
      PassIO<T, N, R, P>::template read<DIR>(v, tl, lds, tw);
      TeamPass<T, N, R, DIR, P + 1, Sync, Lds, Tw>::run(v, t, lds, tw);
    }
  }
};

// Full length-N transform by a team.  In/out: v[j] = a[t + j*TT] (natural order).
// lds: view of the team's private LDS array (Padded<N,R>::size entries).
// BlockSync contains __syncthreads(): every lane of the block must call it.
template <typename T, int N, int R, int DIR, class Lds, class Sync = BlockSync, class Tw = TwMem<T>>
__device__ __forceinline__ void team_fft(cplx<T> (&v)[R], int t, Lds lds, const Tw& tw) {
  static_assert(N % R == 0 && R <= 16 && (R & (R - 1)) == 0 && N >= R, "radix");
  dft<DIR, R>(v);  // pass 0: butterfly b = t reads v[j] = a[t + j*TT], no twiddle
  TeamPass<T, N, R, DIR, 1, Sync, Lds, Tw>::run(v, t, lds, tw);
}
template <typename T, int N, int R, int DIR, class Lds>
__device__ __forceinline__ void team_fft(cplx<T> (&v)[R], int t, Lds lds, Twiddles<T> tw) {
  team_fft<T, N, R, DIR, Lds, BlockSync, TwMem<T>>(v, t, lds, TwMem<T>{tw});
}

// Transform whose lanes change teams at the first exchange: the input is held in a
// load layout (team index tw_, LDS view lw of its team), the output in a compute layout
// (team index tc, view lc) whose teams each sit in one wave.  The first exchange is
// block-wide, every later one wave-local.  Every lane of the block must call it.
template <typename T, int N, int R, int DIR, class Lds, class Tw>
__device__ __forceinline__ void team_fft_remap(cplx<T> (&v)[R], int tl, Lds lw, int tc, Lds lc, const Tw& tw) {
  static_assert(num_passes(N, R) >= 2, "the remap happens at the first exchange");
  dft<DIR, R>(v);
  PassIO<T, N, R, 1>::write(v, opaque_v(tl), lw);
  __syncthreads();
  const int t = opaque_v(tc);
  PassIO<T, N, R, 1>::template read<DIR>(v, t, lc, tw);
  TeamPass<T, N, R, DIR, 2, WaveSync, Lds, Tw>::run(v, tc, lc, tw);
}

}  // namespace wh
