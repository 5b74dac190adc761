// Argument blocks shared by the HIP kernels and the host runtime.
//
// The unit of work is an ensemble of Wigner functions (members); one plan
// propagates all members with one launch per sweep.  A single grid is an
// ensemble of one.  Each launch walks a work list of (member, tile) items; all
// items of a kernel variant have the same lane count and LDS footprint whatever
// the member's grid size, so members of uneven sizes share one launch.
#pragma once
#include <hip/hip_runtime.h>

namespace wh {

constexpr int kMaxDeg = 8;  // polynomial degree of V(t,x) and K(p)
constexpr int kPotTable = -1;  // pot_deg: V given as a [P][X] Bopp-difference table
constexpr int kPotExpr = -2;   // pot_deg: V(t, x) given as an expression program
constexpr int kExprMax = 64, kExprStack = 16;
constexpr int kNCol = 5;    // column sums: Re S0, Im S0, sum p W, sum p^2 W, sum min(W,0)
constexpr int kNObs = 16;   // per-step observable record
constexpr int kColsumCols = 4;     // columns per column-sum tile (64 row groups of 4 lanes)
constexpr int kScaleChunk = 512;   // elements per block of the scale kernel

enum ObsIndex {
  kObsX = 0, kObsX2, kObsP, kObsP2, kObsXP, kObsDV, kObsPDV, kObsXDV, kObsH, kObsNeg,
  kObsNorm,       // dXdP * sum Re W of the recorded (normalised) state
  kObsRawNormRe,  // dXdP * sum W before renormalisation (complex)
  kObsRawNormIm,
  kObsProbSq,     // dX * sum rho(x)^2 (Gross-Pitaevskii record)
  kObsReserved0, kObsReserved1
};

// Sweep feature flags (uniform per launch).
enum SweepFlags {
  kScaleOnLoad = 1,  // multiply by the member's pending complex factor on load
  kColSums = 2,      // p sweep: write the member's column sums (observables, norm, rho)
  kCL = 4,           // p sweep: Caldeira-Leggett order-3 composite after the V step
  kMerge = 8,        // p sweep: apply the next half step after the column sums (Strang)
  kGP = 16,          // p sweep: add the Gross-Pitaevskii phase from the member's G grid
  kColPart = 32,     // x sweep: partial column sums (Gross-Pitaevskii density after K)
  kConj = 64,        // x sweep: conjugated kinetic multiplier (adjoint step)
  kTilePart = 128,   // p sweep (requires kColSums): each tile writes its 12 observable partial sums
  kTileScale = 256,  // x sweep: the pending factor from the previous p sweep's tile partials;
                     // the member's tile 0 writes the record row and the factor (replaces
                     // the reduce launch between the sweeps)
  kStepArg = 512,    // p sweep: the time index is step_add itself (single-sweep entry points),
                     // not the device counter
};

constexpr int kTileObs = 12;  // observable partials per tile

// One ensemble member, resident in device memory.
struct Member {
  void* W;                     // complex [P][X], FFT ordered
  double* colsums;             // [X][kNCol]
  double* scale;               // [2] pending renormalisation factor
  double* obs;                 // record rows [..][kNObs]; row = step - obs_base
  double* rho;                 // [X] Gross-Pitaevskii density
  double* G;                   // [P][X] Re(B- - B+)
  double* colpart;             // [P/rows][X]
  double* tilepart;            // [12][X] per-tile observable partials (kTilePart): k * X + tile
  const double* pot;           // [pot_rows][kMaxDeg+1] coefficients of V(t,x)
  const double* pot_dv;        // [P][X] V(x - hbar theta/2) - V(x + hbar theta/2) when pot_deg < 0
  const double* theta_decay;   // [P] or null
  const double* lambda_decay;  // [X] or null
  const double* xmask;         // [X] or null
  const double* sigma_p;       // [P] sign(p) f(p) (Caldeira-Leggett)
  const double* theta_host;    // [P]
  const double* lambda_host;   // [X]
  const double* x_host;        // [X]
  const double* p_host;        // [P]
  const double* dv_host;       // [X]
  const double* v0_host;       // [X]
  const void* twX_hi;          // twiddles for X in the state precision (+ lo for FP32)
  const void* twX_lo;
  const void* twP_hi;
  const void* twP_lo;
  const void* twX64;           // FP64 twiddles for X (Gross-Pitaevskii shift)
  int X, P, pot_rows, pot_deg;
  double dx, dp, dtheta, dlambda;  // kernel coordinate steps
  double dxdp, inv_2m, dx_host, rho_factor;
  // ODM Lindblad dissipator (dampingFunction 'ODM'), order 1
  const double* odm_g;         // [4][P][X] g(p + hbar l/2), g(p - hbar l/2) and their signed forms
  const double* odm_x;         // [2][X] cos 2x, sin 2x (host x)
  const double* odm_th;        // [2][P] cos(hbar theta), sin(hbar theta) (host theta)
  void* odm_s;                 // [2][P][X] scratch in the state precision (lambda picture)
  double odm_c, odm_sign;      // 2 gamma dt; +1 damping, -1 pumping
  // adjoint step (single-grid plans): forward state after the step, gradient column sums
  const void* W2;              // [P][X] W_{n+1}
  double* gcols;               // [X][kMaxDeg+1]
  // V(t, x) as an expression program (pot_deg = kPotExpr): (op, value) pairs, t = step * pot_t_step
  const double* pot_expr;
  int pot_expr_len, pot_expr_depth;  // ops; the deepest stack the program reaches
  double pot_t_step;
  int obs_base;
  int pot_expr_kvn;            // the program is dV/dx: the table holds the KvN phase -hbar theta dV/dx
  int pad1, pad2;
};

struct SweepArgs {
  const Member* members;
  const unsigned* work;  // (member << 16) | tile, one per block
  int* step;             // device step counter (time index)
  int step_add;
  int flags;
  // uniform physics
  double dt, hbar, vphase, hbar_half, gp_phase, cl_a;
  // kinetic: phase = dt*lambda*(k_c0 + k_c1*p) when kin_linear, else
  // dt*(K(p + hbar lambda/2) - K(p - hbar lambda/2))/hbar with K = sum kcoef[n] p^n
  int kin_linear, kdeg;
  double k_c0, k_c1, inv_hbar;
  double kcoef[kMaxDeg + 1];
  // Step counters.  step: the time index, read by p sweeps; rec_step: the index of the
  // last p sweep, read by the x sweeps that record it.  A p sweep with kTilePart copies
  // step to rec_step, an x sweep with kTileScale sets step = rec_step + 1: each kernel
  // writes only the counter the other one reads.
  int* rec_step;
  int pt_R, pt_block;  // kTileScale: launch shape of the p sweep that wrote the partials
  int fin_renorm, fin_record;
};

// Per-run binding of member states and record rows, written into the device member table
// by a kernel in stream order (no host synchronisation; captured graphs read the table)
constexpr int kBindMax = 60;  // members per bind launch (kernel arguments stay under 1 KiB)
struct BindArgs {
  Member* members;
  int first, count, obs_base, pad;
  void* W[kBindMax];
  double* obs[kBindMax];
};

struct ReduceArgs {
  const Member* members;
  int renorm;     // divide by the complex norm
  int record;     // write the record row
  int rho;        // write rho(x) from the column sums (Gross-Pitaevskii)
  int update_scale;
  const int* step;
  int* step_inc;  // set to step + 1 (or next_step) by the last block to finish
  int use_step_arg;  // 1: the time index is step_arg, the device counter is not read
  int step_arg;
  int next_step;     // >= 0: the value written to *step_inc instead of step + 1
  int pad;
};

}  // namespace wh
