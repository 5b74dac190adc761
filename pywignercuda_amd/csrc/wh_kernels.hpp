// Kernel argument blocks shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include "wh_fft.hpp"

namespace wh {

constexpr int kMaxDeg = 8;       // polynomial degree of V(t,x) and K(p)
constexpr int kNCol = 5;         // column sums: Re S0, Im S0, sum p W, sum p^2 W, sum min(W,0)
constexpr int kNObs = 16;        // per-step observable record

// Observable record layout (doubles), one row per recorded step.
enum ObsIndex {
  kObsX = 0, kObsX2, kObsP, kObsP2, kObsXP, kObsDV, kObsPDV, kObsXDV, kObsH, kObsNeg,
  kObsNorm,        // dXdP * sum Re W of the recorded (normalised) state
  kObsRawNormRe,   // dXdP * sum W before renormalisation (complex)
  kObsRawNormIm,
  kObsProbSq,      // dX * sum rho(x)^2 (Gross-Pitaevskii record)
  kObsReserved0, kObsReserved1
};

// Potential V(t,x) = sum_n c_n(t) x^n, coefficient rows indexed by step.
struct PotentialArgs {
  const double* coef;   // [rows][kMaxDeg+1]
  int rows;             // rows in the table (1 = time independent)
  int deg;
};

struct XSweepArgs {
  void* W;                       // complex [P][X]
  int X, P;
  const void* tw;                // twiddles for length X (cplx<T>)
  const double* scale;           // complex factor applied on load (2 doubles) or null
  int kinetic_poly;              // 0: dt*lambda*p/m (expPLambda); 1: polynomial K
  double kcoef[kMaxDeg + 1];
  int kdeg;
  double dt, inv_mass, hbar, inv_hbar;
  double dp, dlambda;            // kernel coordinate steps
  const double* lambda_decay;    // [X] exp(-D_lambda*lambda^2) or null
  double* colpart;               // [P/ROWS][X] partial column sums (Re W) or null
};

struct PSweepArgs {
  void* W;                       // complex [P][X]
  int X, P;
  const void* tw;                // twiddles for length P
  const double* scale;           // complex factor applied on load or null
  const int* step;               // device step counter (time index of this V step)
  int step_add;                  // added to *step
  PotentialArgs pot;
  double vphase;                 // frac*dt/hbar
  double hbar_half;              // hbar/2
  double dx, dtheta;             // kernel coordinate steps
  const double* theta_decay;     // [P] exp(-frac*D_theta*theta^2*dt) or null
  const double* xmask;           // [X] absorbing mask or null
  const double* gp_G;            // [P][X] Re(B- - B+) or null
  double gp_phase;               // -frac*dt*a_GP/hbar
  // Caldeira-Leggett order-3 composite (after the inverse transform)
  int cl;
  const double* sigma_p;         // [P] sign(p) f(p) at kernel p
  const double* theta_host;      // [P] host theta grid
  double cl_a;                   // dt*gamma (a = i*cl_a)
  // observables: column sums of the state after the V step (and damping)
  double* colsums;               // [X][kNCol] or null
  const double* p_host;          // [P]
  // Strang merge: after the column sums apply the next half step (step+1)
  int merge;
};

struct ReduceArgs {
  const double* colsums;         // [X][kNCol]
  int X;
  const double* x_host;          // [X]
  const double* dv_host;         // [X] dV/dx(0,x)
  const double* v0_host;         // [X] V(0,x)
  double dxdp, inv_2m;
  int renorm;                    // divide by the complex norm
  double* obs;                   // [rows][kNObs] record, row = *step - obs_base
  int obs_base;
  int record;                    // write the record row
  const int* step;               // device step counter
  int* step_inc;                 // incremented by one when non-null
  double* scale_out;             // complex factor for the next load (2 doubles) or null
  // Gross-Pitaevskii density rho(x) = Re(S0 * scale) * rho_factor
  double* rho_out;               // [X] or null
  double rho_factor;
  double dx;
};

}  // namespace wh
