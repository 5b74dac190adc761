// Argument block and control layout of the imaginary-time (Bloch) Gross-Pitaevskii solver
// (GPU_Wigner2D_GPitaevskii_Bloch, GPU_Wigner2D_GPitaevskii.py:2064-2543).
#pragma once
#include <hip/hip_runtime.h>

namespace wh {

constexpr int kMaxDegB = 8;
constexpr double kPiB = 3.14159265358979323846;
// column sums: Re S, Im S, sum p Re W, sum p^2 Re W, sum K(p) Re W, Re sum W^2, Im sum W^2,
// Re sum W W_0, Im sum W W_0
constexpr int kBCol = 9;
// history row: dt used, <x>, <x^2>, <p>, <p^2>, Hamiltonian_average, total energy,
// non-linear energy, purity (re, im), kept (1) / refused (0) / untested (-1), Im total energy
constexpr int kBHist = 12;
// device control block (doubles)
enum BlochCtrl { kBDt = 0, kBIter, kBNr, kBNi, kBEr, kBEi, kBCr, kBCi, kBDecision, kBCtrlSize };
// decisions of the finish kernel
constexpr double kBKeepNoCopy = 2.0;  // normalise, do not store as the last kept state
constexpr double kBKeep = 1.0;        // normalise and store as the last kept state
constexpr double kBRefuse = 0.0;      // restore the last kept state, halve dt
enum BlochMode { kBInit = 0, kBStep = 1 };

struct BlochArgs {
  void* W;               // [P][X] complex128 state
  void* Wstep;           // [P][X] last kept state
  const void* Wground;   // [P][X] ground state to remove (Run_ExitedState1) or null
  const double* twX;     // FP64 twiddles of X and P
  const double* twP;
  int X, P;
  double kdx, kdp, kdtheta, kdlambda, khbar, ka;  // kernel constants
  int pot_deg, pot_rows;
  const double* pot;     // [pot_rows][kMaxDegB+1] V(t, x) coefficients
  const double* vsum;    // [P][X] V(x - hbar theta/2) + V(x + hbar theta/2), or null
  int kdeg, gp;
  double kcoef[kMaxDegB + 1];
  double* rho;           // [X] dP sum_p Re W
  double* G;             // [P][X] B- + B+
  const double* theta_host;   // [P]
  const double* lambda_host;  // [X]
  const double* x_host;       // [X]
  const double* p_host;       // [P]
  const double* k_host;       // [P] K(p) on the host grid (Hamiltonian)
  const double* v0_host;      // [X] V(0, x) on the host grid (Hamiltonian)
  double* colsums;       // [X][kBCol]
  const double* g_colsums;    // [X][kBCol] of W_0
  double* ctrl;          // [kBCtrlSize]
  double* hist;          // [iterations][kBHist]
  int hist_base, first_test;
  double dxdp, dx_host, dp_host, hbar_host, gp_host;
};

hipError_t launch_bloch_psweep(const BlochArgs& a, int colsums, hipStream_t s);
hipError_t launch_bloch_xsweep(const BlochArgs& a, hipStream_t s);
hipError_t launch_bloch_gshift(const BlochArgs& a, hipStream_t s);
hipError_t launch_bloch_colsum(const BlochArgs& a, const void* state, double* colsums, int density, hipStream_t s);
hipError_t launch_bloch_finish(const BlochArgs& a, int mode, hipStream_t s);
hipError_t launch_bloch_commit(const BlochArgs& a, hipStream_t s);
hipError_t launch_bloch_set(double* ctrl, double dt, int iter, hipStream_t s);

}  // namespace wh
