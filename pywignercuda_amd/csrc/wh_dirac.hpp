// Device-resident set-up of one Wigner-Dirac 4x4 propagator (shared by the HIP
// kernels and the host runtime).
//
// State: 16 complex planes W_ij[P][X] (plane k = 4 i + j), FFT ordered, contiguous
// (the reference's W_init[4, 4, P, X] and its W11_GPU .. W44_GPU,
// GPU_WignerDirac2D_4x4.py:3311-3330).
#pragma once
#include <hip/hip_runtime.h>

#include "wh_kernels.hpp"

namespace wh {

constexpr int kDNCol = 16;  // per-column sums of one p sweep (see DiracCol)
constexpr int kDNObs = 32;  // per-step record (see DiracObs)

// Column sums over p of the state after the p sweep (unnormalised), per x column.
enum DiracCol {
  kDcTrRe = 0, kDcTrIm,      // sum_p tr W
  kDcPTrRe, kDcPTrIm,        // sum_p p tr W            (kernel p)
  kDcP2TrRe, kDcP2TrIm,      // sum_p p^2 tr W
  kDcA1Re, kDcA1Im,          // sum_p (W14 + W23 + W32 + W41)
  kDcPA1Re, kDcPA1Im,        // sum_p p a1              (kernel p)
  kDcHA1Re, kDcHA1Im,        // sum_p p_host a1         (host p, DiracEnergy)
  kDcBetaRe, kDcBetaIm,      // sum_p (W11 + W22 - W33 - W44)
  kDcNeg,                    // sum_p min(Re tr W, 0)
  kDcPad
};

// Per-step record, complex quantities as (re, im).  Row k = state before iteration
// first_step + k (row n_steps = the final state); the antiparticle population of
// the final row is filled by the closing row-sum pass.
enum DiracObs {
  kDoX = 0, kDoP = 2, kDoXP = 4, kDoXX = 6, kDoPP = 8, kDoA1 = 10, kDoDV = 12, kDoXDV = 14,
  kDoPDV = 16, kDoPA1 = 18, kDoXA1 = 20, kDoEnergy = 22,
  kDoNeg = 24, kDoTrans = 25, kDoAnti = 26, kDoNorm = 27, kDoRawNorm = 28
};

struct DiracPlan {
  void* W;                     // [16][P][X]
  double* colsums;             // [X][kDNCol]
  double* rowpart;             // [2][P] antiparticle contributions of plane groups 0 and 3
  double* scale;               // [1] pending real factor 1/norm, applied on the next load
  double* obs;                 // [rows][kDNObs]
  const int* step;             // device time index
  const double* pot;           // [pot_rows][kMaxDeg+1] V0(t, x) coefficients
  const double* pot_dv;        // [P][X] V0(x - theta/2) - V0(x + theta/2) when pot_deg < 0
  const double* theta_decay;   // [P] exp(-dt D_Theta theta^2) or null
  const double* xmask;         // [X] edge absorber or null
  const double* x_host;        // [X] host grid
  const double* p_host;        // [P] host grid
  const double* theta_host;    // [P] host grid (Caldeira-Leggett)
  const double* dv0;           // [X] dV0/dx at kernel x (time-independent V0)
  const double* vsum;          // [X] sum_k V_k at kernel x (DiracEnergy, time-independent)
  const double* mass_dm;       // [P][X] F(x - theta/2) - F(x + theta/2), F = Mass(x) c^2 dt, or null
  const double* mass_sm;       // [P][X] F(x - theta/2) + F(x + theta/2)
  double* rho;                 // [X] Gross-Pitaevskii density Re sum_p tr W (normalised state)
  double* G;                   // [P][X] Re(B- - B+), B-/+ = rho(x -/+ theta/2)
  const double* lambda_host;   // [X] host grid (shift theorem)
  const void* twX64;           // FP64 twiddles for X (density shift)
  const void* twX_hi;
  const void* twX_lo;
  const void* twP_hi;
  const void* twP_lo;
  int X, P, pot_rows, pot_deg;
  int obs_base, energy, antiparticle, pad0;
  double dx, dp, dtheta, dlambda;   // kernel steps ('%f')
  double mh_c;                      // mass_half * c          (kinetic shell)
  double c_dt;                      // c * dt
  double mass_phase;                // F = mass_half c^2 dt   (constant mass of the potential factors, A = 0)
  double dt;                        // kernel dt
  double proj_mc2, proj_c;          // mass c^2, c            (projector, kernel constants)
  double cl_a;                      // dt * gamma of Caldeira-Leggett order 2 (host values); 0 = off
  double dxdp, c_host, mc2_host;    // observables (host constants)
  double transmission_x;            // x threshold of the transmission record
  double gp_coeff;                  // aGPitaevskii (host value); 0 = off
  // vector potentials A1..A3 (Potential_1..3) as polynomials in x; general = any non-zero
  int general, adeg0, adeg1, adeg2;
  double acoef[3][kMaxDeg + 1];
  // ODM damping (DampingODM_Kernel): odm_k = mass c dt gammaDamping, 1/lambdaBar; odm = on
  int odm, kvn;                     // ODM damping; the KvN variant (GPU_WignerDiracKvN2D_4x4)
  double odm_k, odm_inv_lb;
  // antiparticle step filter (filter = 1): the edge absorber then follows the filter (fmask)
  // instead of riding on the potential factors (xmask = null)
  int filter, pad1;
  const double* fmask;
  // time-dependent vector potentials: a_rows rows of [3][kMaxDeg + 1] (row = time index,
  // clamped); a_rows = 0: acoef above
  const double* acoef_rows;
  int a_rows, pad2;
  // V0(t, x) and dV0/dx(t, x) as expression programs (the host's pot_deg WH_POT_EXPR; pot_deg
  // is the table form here): pot_dv before each potential sweep, dv0 and vsum before each reduce
  const double* pot_expr;
  const double* dpot_expr;
  int pot_expr_len, dpot_expr_len, pot_expr_depth, dpot_expr_depth;
  double pot_t_step;
  // Mass(t, x) as an expression program: mass_dm / mass_sm refilled before each potential sweep
  const double* mass_expr;
  int mass_expr_len, mass_expr_depth;
  double mass_c, mass_dt, mass_t_step;
  // A1..A3 as expression programs: a_tab [3][2][P][X] = A_k(x - theta/2), A_k(x + theta/2) at the
  // potential sweep's time index; asum [X] = sum_k A_k(x) at the record row's (null: polynomials)
  const double* a_expr[3];
  int a_expr_len[3], a_expr_depth;
  const double* a_tab;
  const double* asum;
  double a_t_step;
};

}  // namespace wh
