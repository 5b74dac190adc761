// Tables of a time-dependent potential given as an expression program (pot_deg WH_POT_EXPR
// on the host): before each p sweep, the Bopp difference V(x - hbar theta/2, t) -
// V(x + hbar theta/2, t) of every cell at the sweep's time index (and the next one, for the
// merged Strang half step) -- the per-cell evaluation of expPotential (GPU_Wigner2D_FFT.py:
// 88-94) as a kernel of its own, so that the p sweep reads a table as for a time-independent
// non-polynomial V and keeps its register budget.  With pot_expr_kvn the program is dV/dx and
// the table holds the classical KvN phase -hbar theta dV/dx(t, x) (exp_PotentialKvN, :27-75).
#include <hip/hip_runtime.h>

#include "wh_dirac.hpp"
#include "wh_sweeps.hpp"

namespace wh {

constexpr int kExprBlock = 256;

// V(x1, t) - V(x2, t) of a program (op codes WH_EXPR_* of include/wignerhip.h; the library
// checks length, op codes and stack depth when the plan is created) in one pass: both stacks
// advance together (the same ops), so each op is dispatched once and the two evaluations
// overlap.  Ops are read through the constant address space (uniform: scalar loads).  The tops of the
// stacks live in registers; slot s below them: st[s * kExprBlock] (x1) and
// st[off + s * kExprBlock] (x2), off = (depth - 1) * kExprBlock.
struct ExprPair {
  double a, b;  // V(x1, t), V(x2, t)
};
__device__ __forceinline__ ExprPair eval_expr_pair(cptr<double> code, int n, double x1, double x2, double t,
                                                   double* st, int off) {
  constexpr int LS = kExprBlock;
  const int OFF = off;
  double a = 0.0, b = 0.0;  // tops of the two stacks
  int sp = 0;
  for (int i = 0; i < n; ++i) {
    const int op = int(code[2 * i]);
    const double v = code[2 * i + 1];
    if (op <= 2) {
      if (sp > 0) {
        st[(sp - 1) * LS] = a;
        st[OFF + (sp - 1) * LS] = b;
      }
      a = op == 0 ? v : op == 1 ? x1 : t;
      b = op == 0 ? v : op == 1 ? x2 : t;
      ++sp;
    } else if (op == 3) {
      a = -a;
      b = -b;
    } else if (op <= 8) {
      const double p = st[(sp - 2) * LS], q = st[OFF + (sp - 2) * LS];
      --sp;
      switch (op) {
        case 4: a = p + a; b = q + b; break;
        case 5: a = p - a; b = q - b; break;
        case 6: a = p * a; b = q * b; break;
        case 7: a = p / a; b = q / b; break;
        default: a = pow(p, a); b = pow(q, b); break;
      }
    } else {
      switch (op) {
        case 9: a = exp(a); b = exp(b); break;
        case 10: a = sin(a); b = sin(b); break;
        case 11: a = cos(a); b = cos(b); break;
        case 12: a = tan(a); b = tan(b); break;
        case 13: a = sqrt(a); b = sqrt(b); break;
        case 14: a = fabs(a); b = fabs(b); break;
        case 15: a = tanh(a); b = tanh(b); break;
        case 16: a = sinh(a); b = sinh(b); break;
        case 17: a = cosh(a); b = cosh(b); break;
        case 18: a = log(a); b = log(b); break;
        case 19: a = atan(a); b = atan(b); break;
        default: a = a < 0.0 ? 0.0 : 1.0; b = b < 0.0 ? 0.0 : 1.0; break;  // Heaviside
      }
    }
  }
  return {a, b};
}

// single: return V(x1, t) alone.
__device__ __forceinline__ double eval_expr_diff(cptr<double> code, int n, double x1, double x2, double t,
                                                 double* st, int off, bool single) {
  const ExprPair r = eval_expr_pair(code, n, x1, x2, t, st, off);
  return single ? r.a : r.a - r.b;
}

// Table k (k < ntables) of member `member` at time index step + k, where step is step_add
// (use_arg) or the device counter plus step_add, as the p sweep reads it.  One thread per cell,
// grid-stride; the stack slots of the threads interleave in LDS.
// depth: the deepest stack of the program (the LDS holds 2 (depth - 1) slots per thread).
__global__ void __launch_bounds__(kExprBlock) expr_table_kernel(const Member* __restrict__ members, int member,
                                                                int depth, const int* __restrict__ step,
                                                                int step_add, int use_arg, int ntables,
                                                                double hbar_half) {
  extern __shared__ double st[];
  const int off = (depth > 1 ? depth - 1 : 1) * kExprBlock;
  const Member& m = members[member];
  const int X = m.X, P = m.P;
  const size_t n = size_t(X) * size_t(P);
  const int s0 = use_arg ? step_add : *step + step_add;
  const double hd = hbar_half * m.dtheta;
  double* tab = const_cast<double*>(m.pot_dv);
  const cptr<double> code = cmem(m.pot_expr);
  const int len = m.pot_expr_len;
  const bool kvn = m.pot_expr_kvn != 0;
  const int lx = __ffs(X) - 1;  // X is a power of two
  const int cells = int(n);     // at most 2048^2
  for (int c = int(blockIdx.x) * kExprBlock + int(threadIdx.x); c < cells; c += int(gridDim.x) * kExprBlock) {
    const int i = c >> lx, col = c & (X - 1);
    const double x = m.dx * double(fidx(col, X)), u = hd * double(fidx(i, P));
    for (int k = 0; k < ntables; ++k) {
      const double t = double(s0 + k) * m.pot_t_step;
      // Bopp difference V(x - u) - V(x + u), or the KvN phase -2 u dV/dx(x) (u = hbar theta / 2)
      tab[size_t(k) * n + size_t(c)] =
          kvn ? -2.0 * u * eval_expr_diff(code, len, x, x, t, st + threadIdx.x, off, true)
              : eval_expr_diff(code, len, x - u, x + u, t, st + threadIdx.x, off, false);
    }
  }
}

// Wigner-Dirac V0(t, x) programs: the potential step's Bopp difference V0(x - theta/2) -
// V0(x + theta/2) (no hbar: GPU_WignerDirac2D_4x4.py:408, 568) at the sweep's time index.
__global__ void __launch_bounds__(kExprBlock) dirac_expr_table_kernel(const DiracPlan* __restrict__ dp) {
  extern __shared__ double st[];
  const DiracPlan& d = *dp;
  const int depth = d.pot_expr_depth;
  const int off = (depth > 1 ? depth - 1 : 1) * kExprBlock;
  const int X = d.X, P = d.P;
  const int lx = __ffs(X) - 1;
  const int cells = X * P;
  const double t = double(*d.step) * d.pot_t_step;
  double* tab = const_cast<double*>(d.pot_dv);
  const cptr<double> code = cmem(d.pot_expr);
  for (int c = int(blockIdx.x) * kExprBlock + int(threadIdx.x); c < cells; c += int(gridDim.x) * kExprBlock) {
    const int i = c >> lx, col = c & (X - 1);
    const double x = d.dx * double(fidx(col, X)), u = 0.5 * d.dtheta * double(fidx(i, P));
    tab[c] = eval_expr_diff(code, d.pot_expr_len, x - u, x + u, t, st + threadIdx.x, off, false);
  }
}

// dV0/dx and V0 of the record row the next reduce writes (time index step + state_add).
__global__ void __launch_bounds__(kExprBlock) dirac_expr_obs_kernel(const DiracPlan* __restrict__ dp,
                                                                    int state_add) {
  extern __shared__ double st[];
  const DiracPlan& d = *dp;
  const int depth = d.pot_expr_depth > d.dpot_expr_depth ? d.pot_expr_depth : d.dpot_expr_depth;
  const int off = (depth > 1 ? depth - 1 : 1) * kExprBlock;
  const double t = double(*d.step + state_add) * d.pot_t_step;
  for (int col = int(blockIdx.x) * kExprBlock + int(threadIdx.x); col < d.X; col += int(gridDim.x) * kExprBlock) {
    const double x = d.dx * double(fidx(col, d.X));
    const_cast<double*>(d.dv0)[col] =
        eval_expr_diff(cmem(d.dpot_expr), d.dpot_expr_len, x, x, t, st + threadIdx.x, off, true);
    const_cast<double*>(d.vsum)[col] =
        eval_expr_diff(cmem(d.pot_expr), d.pot_expr_len, x, x, t, st + threadIdx.x, off, true);
  }
}

// Mass(t, x) program of the MassPotential variant: the tables F(x - theta/2) -/+ F(x + theta/2),
// F = Mass c^2 dt (kernel constants, in the host's order of products), at the sweep's time index.
__global__ void __launch_bounds__(kExprBlock) dirac_mass_table_kernel(const DiracPlan* __restrict__ dp) {
  extern __shared__ double st[];
  const DiracPlan& d = *dp;
  const int depth = d.mass_expr_depth;
  const int off = (depth > 1 ? depth - 1 : 1) * kExprBlock;
  const int X = d.X, P = d.P;
  const int lx = __ffs(X) - 1;
  const int cells = X * P;
  const double t = double(*d.step) * d.mass_t_step;
  const cptr<double> code = cmem(d.mass_expr);
  for (int c = int(blockIdx.x) * kExprBlock + int(threadIdx.x); c < cells; c += int(gridDim.x) * kExprBlock) {
    const int i = c >> lx, col = c & (X - 1);
    const double x = d.dx * double(fidx(col, X)), u = 0.5 * d.dtheta * double(fidx(i, P));
    const ExprPair m = eval_expr_pair(code, d.mass_expr_len, x - u, x + u, t, st + threadIdx.x, off);
    const double fm = m.a * d.mass_c * d.mass_c * d.mass_dt, fp = m.b * d.mass_c * d.mass_c * d.mass_dt;
    const_cast<double*>(d.mass_dm)[c] = fm - fp;
    const_cast<double*>(d.mass_sm)[c] = fm + fp;
  }
}

// Potential_1..3 programs (vector potentials that are not polynomials in x): A_k(x -/+ theta/2)
// of every cell at the potential sweep's time index, for the two factor sweeps.
__global__ void __launch_bounds__(kExprBlock) dirac_a_table_kernel(const DiracPlan* __restrict__ dp) {
  extern __shared__ double st[];
  const DiracPlan& d = *dp;
  const int depth = d.a_expr_depth;
  const int off = (depth > 1 ? depth - 1 : 1) * kExprBlock;
  const int X = d.X, P = d.P;
  const int lx = __ffs(X) - 1;
  const int cells = X * P;
  const size_t n = size_t(cells);
  const double t = double(*d.step) * d.a_t_step;
  double* tab = const_cast<double*>(d.a_tab);
  for (int c = int(blockIdx.x) * kExprBlock + int(threadIdx.x); c < cells; c += int(gridDim.x) * kExprBlock) {
    const int i = c >> lx, col = c & (X - 1);
    const double x = d.dx * double(fidx(col, X)), u = 0.5 * d.dtheta * double(fidx(i, P));
    for (int k = 0; k < 3; ++k) {
      ExprPair a{0.0, 0.0};
      if (d.a_expr_len[k] > 0) a = eval_expr_pair(cmem(d.a_expr[k]), d.a_expr_len[k], x - u, x + u, t, st + threadIdx.x, off);
      tab[size_t(2 * k) * n + size_t(c)] = a.a;
      tab[size_t(2 * k + 1) * n + size_t(c)] = a.b;
    }
  }
}

// sum_k A_k(x) of the record row the next reduce writes (time index step + state_add).
__global__ void __launch_bounds__(kExprBlock) dirac_a_sum_kernel(const DiracPlan* __restrict__ dp, int state_add) {
  extern __shared__ double st[];
  const DiracPlan& d = *dp;
  const int depth = d.a_expr_depth;
  const int off = (depth > 1 ? depth - 1 : 1) * kExprBlock;
  const double t = double(*d.step + state_add) * d.a_t_step;
  for (int col = int(blockIdx.x) * kExprBlock + int(threadIdx.x); col < d.X; col += int(gridDim.x) * kExprBlock) {
    const double x = d.dx * double(fidx(col, d.X));
    double s = 0.0;
    for (int k = 0; k < 3; ++k)
      if (d.a_expr_len[k] > 0) s += eval_expr_pair(cmem(d.a_expr[k]), d.a_expr_len[k], x, x, t, st + threadIdx.x, off).a;
    const_cast<double*>(d.asum)[col] = s;
  }
}

static size_t expr_lds(int depth) {
  return size_t(2) * size_t(depth > 1 ? depth - 1 : 1) * kExprBlock * sizeof(double);
}

hipError_t launch_dirac_expr_table(const DiracPlan* d, int cells, int depth, hipStream_t s) {
  if (depth < 1 || depth > kExprStack) return hipErrorInvalidValue;
  const int blocks = (cells + kExprBlock - 1) / kExprBlock;
  hipLaunchKernelGGL(dirac_expr_table_kernel, dim3(blocks < 8192 ? blocks : 8192), dim3(kExprBlock),
                     expr_lds(depth), s, d);
  return hipGetLastError();
}

hipError_t launch_dirac_mass_table(const DiracPlan* d, int cells, int depth, hipStream_t s) {
  if (depth < 1 || depth > kExprStack) return hipErrorInvalidValue;
  const int blocks = (cells + kExprBlock - 1) / kExprBlock;
  hipLaunchKernelGGL(dirac_mass_table_kernel, dim3(blocks < 8192 ? blocks : 8192), dim3(kExprBlock),
                     expr_lds(depth), s, d);
  return hipGetLastError();
}

hipError_t launch_dirac_a_table(const DiracPlan* d, int cells, int depth, hipStream_t s) {
  if (depth < 1 || depth > kExprStack) return hipErrorInvalidValue;
  const int blocks = (cells + kExprBlock - 1) / kExprBlock;
  hipLaunchKernelGGL(dirac_a_table_kernel, dim3(blocks < 8192 ? blocks : 8192), dim3(kExprBlock), expr_lds(depth),
                     s, d);
  return hipGetLastError();
}

hipError_t launch_dirac_a_sum(const DiracPlan* d, int X, int depth, int state_add, hipStream_t s) {
  if (depth < 1 || depth > kExprStack) return hipErrorInvalidValue;
  hipLaunchKernelGGL(dirac_a_sum_kernel, dim3((X + kExprBlock - 1) / kExprBlock), dim3(kExprBlock),
                     expr_lds(depth), s, d, state_add);
  return hipGetLastError();
}

hipError_t launch_dirac_expr_obs(const DiracPlan* d, int X, int depth, int state_add, hipStream_t s) {
  if (depth < 1 || depth > kExprStack) return hipErrorInvalidValue;
  hipLaunchKernelGGL(dirac_expr_obs_kernel, dim3((X + kExprBlock - 1) / kExprBlock), dim3(kExprBlock),
                     expr_lds(depth), s, d, state_add);
  return hipGetLastError();
}

hipError_t launch_expr_table(const Member* members, int member, int cells, int depth, const int* step,
                             int step_add, int use_arg, int ntables, double hbar_half, hipStream_t s) {
  if (depth < 1 || depth > kExprStack) return hipErrorInvalidValue;
  const int blocks = (cells + kExprBlock - 1) / kExprBlock;
  const size_t lds = size_t(2) * size_t(depth > 1 ? depth - 1 : 1) * kExprBlock * sizeof(double);
  hipLaunchKernelGGL(expr_table_kernel, dim3(blocks < 8192 ? blocks : 8192), dim3(kExprBlock), lds, s, members,
                     member, depth, step, step_add, use_arg, ntables, hbar_half);
  return hipGetLastError();
}

}  // namespace wh
