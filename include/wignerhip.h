/*
 * wignerhip.h -- C ABI of libwignerhip.so, the MI355X (gfx950) Wigner
 * split-operator library.
 *
 * Conventions (all entry points):
 *   - return an int status, 0 = success; codes 1-9 mirror the cuFFT status set
 *     that cufft_wrapper.py:15-73 maps to Python exceptions, 10+ are new;
 *     wh_last_error() gives a message for the calling thread's last failure;
 *   - device memory is borrowed (plain device pointers, e.g. torch
 *     tensor.data_ptr()); the library never frees caller memory;
 *   - every launching call takes an explicit stream (hipStream_t as void*,
 *     NULL = the null stream) and is asynchronous with respect to the host;
 *   - layouts: complex (re,im interleaved) FP64 (WH_C128) or FP32 (WH_C64),
 *     W[p][x] row-major, both axes FFT ordered (index 0 = origin), exactly the
 *     layout of the reference's W_GPU (GPU_Wigner2D_FFT.py:110-114, 571-581).
 */
#ifndef WIGNERHIP_H
#define WIGNERHIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
/* 1-9: cufftExceptions of cufft_wrapper.py:55-65 */
#define WH_SUCCESS 0
#define WH_INVALID_PLAN 1
#define WH_ALLOC_FAILED 2
#define WH_INVALID_TYPE 3
#define WH_INVALID_VALUE 4
#define WH_INTERNAL_ERROR 5
#define WH_EXEC_FAILED 6
#define WH_SETUP_FAILED 7
#define WH_INVALID_SIZE 8
#define WH_UNALIGNED_DATA 9
/* new */
#define WH_INVALID_DESCRIPTOR 10
#define WH_LAUNCH_FAILED 11
#define WH_NO_DEVICE 12

/* precisions: CUFFT_Z2Z / CUFFT_C2C of cufft_wrapper.py:8-9 */
#define WH_C128 0
#define WH_C64 1
#define WH_C64_WIDE 2 /* complex64 state, FP64 arithmetic inside the sweeps (Wigner2D plans and ensembles) */
/* directions: CUFFT_FORWARD / CUFFT_INVERSE of cufft_wrapper.py:11-12 */
#define WH_FORWARD (-1)
#define WH_INVERSE 1

/* split schemes */
#define WH_SCHEME_LIE 0    /* K(dt); V(dt)        GPU_Wigner2D_FFT.Run, GPU_Wigner2D_FFT.py:1884-1899 */
#define WH_SCHEME_STRANG 1 /* V(dt/2); K; V(dt/2) GPU_Wigner2D_GPitaevskii.Run, GP.py:1911-1956 */

#define WH_MAX_DEG 8 /* degree bound of the polynomial V(t,x) and K(p) */
#define WH_POT_TABLE (-1) /* pot_deg value: V given as a [P][X] Bopp-difference table */
#define WH_POT_EXPR (-2)  /* pot_deg value: V(t, x) given as an expression program (below) */

/* Expression programs (pot_deg = WH_POT_EXPR): V(t, x) as WH_EXPR_MAX or fewer (op, value)
   pairs of doubles for a stack machine, evaluated per cell on the device before each p sweep
   at time t = step * pot_t_step -- the counterpart of the reference splicing its Potential
   string into expPotential (GPU_Wigner2D_FFT.py:88-94, 869-873).  Ops: push value / x / t,
   negate, the binary + - * / pow (second operand on top), and the functions exp sin cos tan
   sqrt fabs tanh sinh cosh log atan Heaviside (0 below 0, else 1, as the reference's device
   Heaviside).  The stack holds at most WH_EXPR_STACK values. */
#define WH_EXPR_MAX 64
#define WH_EXPR_STACK 16
#define WH_EXPR_CONST 0
#define WH_EXPR_X 1
#define WH_EXPR_T 2
#define WH_EXPR_NEG 3
#define WH_EXPR_ADD 4
#define WH_EXPR_SUB 5
#define WH_EXPR_MUL 6
#define WH_EXPR_DIV 7
#define WH_EXPR_POW 8
#define WH_EXPR_EXP 9
#define WH_EXPR_SIN 10
#define WH_EXPR_COS 11
#define WH_EXPR_TAN 12
#define WH_EXPR_SQRT 13
#define WH_EXPR_FABS 14
#define WH_EXPR_TANH 15
#define WH_EXPR_SINH 16
#define WH_EXPR_COSH 17
#define WH_EXPR_LOG 18
#define WH_EXPR_ATAN 19
#define WH_EXPR_HEAVISIDE 20
#define WH_NOBS 16   /* doubles per observable record */

/* observable record columns (one row per recorded step) */
#define WH_OBS_X 0
#define WH_OBS_X2 1
#define WH_OBS_P 2
#define WH_OBS_P2 3
#define WH_OBS_XP 4
#define WH_OBS_DV 5
#define WH_OBS_PDV 6
#define WH_OBS_XDV 7
#define WH_OBS_H 8
#define WH_OBS_NEG 9
#define WH_OBS_NORM 10
#define WH_OBS_RAWNORM_RE 11
#define WH_OBS_RAWNORM_IM 12
#define WH_OBS_PROBSQ 13

/* ---- library ---------------------------------------------------------- */
/* Message of the calling thread's last failed call ("" if none). */
const char* wh_last_error(void);
/* Static text for a status code. */
const char* wh_status_string(int status);
/* Library version, (major << 16) | minor. */
int wh_version(void);

/* ---- plain axis transforms: replaces cufft_wrapper.py ----------------- */
typedef struct wh_fft_plan wh_fft_plan;

/* cuFFT type codes, as cufft_wrapper.py:8-9 passes them */
#define WH_CUFFT_Z2Z 0x69
#define WH_CUFFT_C2C 0x29

/*
 * Replaces cufftPlanMany (cufft_wrapper.py:118-145) with its own argument list, so
 * that the reference's ctypes binding passes its arguments unchanged: n, inembed
 * and onembed are addresses of int arrays (n.ctypes.data, :165-167, :199-201,
 * :234-236), type is CUFFT_Z2Z or CUFFT_C2C.  Accepted (every length a power of
 * two in 64..2048, in place, so onembed/ostride/odist equal the input's):
 *   rank 1, inembed NULL            batch contiguous transforms (Plan_Z2Z((n,)),
 *                                   :160-167, istride 1, idist 0 as passed there)
 *   rank 1, istride 1, idist n[0]   batch contiguous rows (Plan_Z2Z_2D_Axis1, :225-236)
 *   rank 1, istride batch, idist 1  batch interleaved columns of an [n][batch] array
 *                                   (Plan_Z2Z_2D_Axis0, :190-201)
 *   rank 2, inembed NULL            batch contiguous [n[0]][n[1]] 2-D transforms
 *                                   (Plan_Z2Z((P, X)), GPU_Wigner2D_FFT.py:860)
 * inembed[0] (the outermost extent) is not read, as in cuFFT.  Allocates the
 * twiddle tables on the current device.
 */
int wh_cufft_plan_many(wh_fft_plan** plan, int rank, const int* n, const int* inembed, int istride, int idist,
                       const int* onembed, int ostride, int odist, int type, int batch);
/*
 * The rank-1 plans of wh_cufft_plan_many with the length passed by value and the
 * library's precision codes: length n (a power of two, 64..2048), either
 *   istride == 1, idist == n      batch contiguous rows  (Plan_Z2Z_2D_Axis1, :218-250;
 *                                                         Plan_Z2Z 1-D, :147-180;
 *                                                         with batch 1 any idist)
 *   istride == batch, idist == 1  batch interleaved columns of an [n][batch] array
 *                                                         (Plan_Z2Z_2D_Axis0, :183-215)
 * type: WH_C128 (Z2Z) or WH_C64 (C2C).  Allocates the twiddle table on the
 * current device.
 */
int wh_fft_plan_many(wh_fft_plan** plan, int n, int istride, int idist, int type, int batch);
/*
 * Replaces cufftExecZ2Z / cufftExecC2C (cufft_wrapper.py:81-103): in place
 * (idata == odata), direction WH_FORWARD or WH_INVERSE, inverse unnormalised.
 */
int wh_fft_exec(wh_fft_plan* plan, void* idata, void* odata, int direction, void* stream);
/* As wh_fft_exec, with the 1/n of the callers' "*= 1./n" (GPU_Wigner2D_FFT.py:790-801) folded in. */
int wh_fft_exec_normalized(wh_fft_plan* plan, void* idata, void* odata, int direction, void* stream);
/* Replaces cufftDestroy (cufft_wrapper.py:109-115); any plan of the two creators. */
int wh_fft_destroy(wh_fft_plan* plan);

/* ---- fused Wigner2D propagator ---------------------------------------- */
/*
 * Replaces, for one P x X Wigner function, the per-instance set-up of the
 * reference: SetCUDA_Constants + SetFFT_Plans + SetCUDA_Functions and the
 * device grids of SetPhaseSpaceGrid2D (GPU_Wigner2D_FFT.py:569-596, 833-906;
 * GPU_Wigner2D_GPitaevskii.py:854-937).  All pointers in the descriptor are
 * HOST arrays, copied at creation.  Constants are the kernel-side values (the
 * caller applies any '%f' truncation the reference performs).
 */
typedef struct {
  int struct_size; /* sizeof(wh_wigner2d_desc) */
  int precision;   /* WH_C128 or WH_C64 */
  int x_dim, p_dim; /* powers of two, 128..2048 */
  int scheme;      /* WH_SCHEME_LIE or WH_SCHEME_STRANG */
  /* kernel constants (expPotential / expPLambda sources, GPU_Wigner2D_FFT.py:80-200) */
  double dt, dx, dp, dtheta, dlambda, hbar, mass;
  /* kinetic phase: 0 = dt*lambda*p/mass (expPLambda, FFT.py:153);
     1 = dt*(K(p+hbar*lambda/2)-K(p-hbar*lambda/2))/hbar, K = sum kin_coef[n] p^n */
  int kinetic_poly;
  int kinetic_deg;
  double kin_coef[WH_MAX_DEG + 1];
  /* potential V(t,x) = sum_n pot_coef[row*(WH_MAX_DEG+1)+n] x^n, row = time index
     (clamped to pot_rows-1); pot_rows = 1 for a time-independent V.
     pot_deg = WH_POT_TABLE: V is not a polynomial; the p sweep reads the Bopp
     difference from pot_dv instead (pot_coef is then unused).
     pot_deg = WH_POT_EXPR: V(t, x) is the program pot_expr (pot_coef unused); before each
     p sweep the library tabulates its Bopp difference at the sweep's time index (two
     [P][X] FP64 tables of device memory per such member) */
  int pot_deg;
  int pot_rows;
  const double* pot_coef;
  /* 1-D tables (NULL = all ones) */
  const double* theta_decay;  /* [P] exp(-frac*D_theta*theta^2*dt) (FFT.py:117; GP.py:80) */
  const double* lambda_decay; /* [X] exp(-D_lambda*lambda^2) (FFT.py:154) */
  const double* x_mask;       /* [X] absorbing mask (FFT.py:121-123, times :514-516 with GP) */
  /* dissipator applied after the V step */
  int damping;                /* 0 none, 1 Caldeira-Leggett order 3 (FFT.py:1145-1157), 2 ODM (below) */
  double gamma_dt;            /* dt*gammaDamping */
  const double* sigma_p;      /* [P] sign(p) f(p) at kernel p (theta_fp, FFT.py:399-431) */
  const double* theta_host;   /* [P] host Theta grid (FFT.py:578) */
  int renormalize;            /* divide by dXdP*sum(W) after each step (FFT.py:1912-1913) */
  /* Gross-Pitaevskii (FFT.py:1687-1705, 479-519; GP.py:466-496, 1743-1748) */
  double gp_coeff;            /* a_GP (kernel constant); 0 = off */
  double gp_rho_factor;       /* 1 (GPU_Wigner2D_FFT) or dP (GPU_Wigner2D_GPitaevskii) */
  const double* lambda_host;  /* [X] host Lambda grid, for exp(i lambda theta/2) */
  /* observables (FFT.py:1857-1878): host grids and dV/dx, V at t=0 */
  const double* x_host;       /* [X] */
  const double* p_host;       /* [P] */
  const double* dv_host;      /* [X] dPotential(0,x) */
  const double* v0_host;      /* [X] Potential(0,x) */
  double dxdp;                /* dX*dP */
  double inv_2m;              /* 1/(2 mass) (host mass) */
  double dx_host;             /* dX (ProbabilitySquare record) */
  /* [P][X] V(x - hbar*theta/2) - V(x + hbar*theta/2) at kernel coordinates, FFT ordered,
     for pot_deg = WH_POT_TABLE (time-independent V given as an arbitrary expression,
     Potential(t,x) spliced into expPotential, FFT.py:88-94, 869-873) */
  const double* pot_dv;
  /* ODM Lindblad dissipator (damping = 2; Lindbladian_ODM_Order1, FFT.py:1516-1634, 1907-1910;
     with WH_SCHEME_STRANG after the second half step, GP.py:1958-1968):
     odm_g [4][P][X] = g(p + hbar lambda/2), g(p - hbar lambda/2) and both times sign(p -/+ ...),
     g(q) = sqrt(L_material f(|q| + hbar/(2 L_material))) on host grids (SetA_ODM, :1196-1244);
     odm_x [2][X] = cos 2x, sin 2x; odm_th [2][P] = cos(hbar theta), sin(hbar theta) */
  const double* odm_g;
  const double* odm_x;
  const double* odm_th;
  double odm_sign;            /* 1 damping, -1 pumping */
  /* ODM channels: 0 or 1 = one (Lindbladian_ODM_Order1, weight odm_weight1 -- 1 in
     GPU_Wigner2D_FFT.Run); 2 = GPU_Wigner2D_ODM2 (FFT.py:1973-2230): W += L(W) with
     (odm_weight1, sign 1), then with (odm_weight2, sign -1) on the result (:2173-2181) */
  int odm_channels;
  double odm_weight1, odm_weight2;
  /* pot_deg = WH_POT_EXPR: the program of V(t, x) (pot_expr_len pairs, host memory) and the
     time per step index (the host dt, as the reference's t = tIndex*dt).  pot_expr_kvn = 1:
     the program is dV/dx(t, x) and the p sweep takes the classical KvN phase
     -dt dV/dx(t, x) theta instead of the Bopp difference (exp_PotentialKvN,
     GPU_Wigner2D_FFT.py:27-75) */
  const double* pot_expr;
  int pot_expr_len;
  int pot_expr_kvn;
  double pot_t_step;
} wh_wigner2d_desc;

typedef struct wh_wigner2d wh_wigner2d;

int wh_wigner2d_create(const wh_wigner2d_desc* desc, wh_wigner2d** out);
int wh_wigner2d_destroy(wh_wigner2d* plan);

/*
 * Replaces the time loop of GPU_Wigner2D_FFT.Run (GPU_Wigner2D_FFT.py:1843-1913)
 * and GPU_Wigner2D_GPitaevskii.Run (GP.py:1898-1999): advances W (device,
 * in place) by n_steps iterations with time indices first_step .. first_step+n_steps-1
 * and writes (n_steps+1) observable rows of WH_NOBS doubles to obs (device):
 *   Lie:    row k = state before iteration first_step+k (row n_steps = final state);
 *   Strang: row k = state after iteration first_step+k   (row n_steps unused).
 * normalize_initial: divide the incoming W by dXdP*sum(W) first (Run's set-up,
 * FFT.py:1778-1780).  On return W holds the (renormalised) final state.
 * use_graph: replay the per-step launch sequence from a captured hipGraph.
 */
int wh_wigner2d_run(wh_wigner2d* plan, void* W, int first_step, int n_steps, double* obs,
                    int normalize_initial, int use_graph, void* stream);

/* Single sweeps, for callers composing their own loop (no renormalisation). */
/* K step: FFT_x, kinetic multiplier, inverse (Fourier_X_To_Lambda_GPU; expPLambda; Fourier_Lambda_To_X_GPU, FFT.py:1885-1888) */
int wh_wigner2d_kinetic(wh_wigner2d* plan, void* W, void* stream);
/* V step at time index `step` (Fourier_P_To_Theta_GPU; expPotential; Fourier_Theta_To_P_GPU, FFT.py:1891-1899).
   fused = 1: the p sweep exactly as configured inside wh_wigner2d_run (damping and the
   observable column sums fused in; Lie scheme); fused = 0: the plain V step. */
int wh_wigner2d_potential(wh_wigner2d* plan, void* W, int step, int fused, void* stream);
/*
 * Adjoint of one Lie step at time index `step` (no counterpart in the reference,
 * which has no gradients): Lambda <- U_step^dagger Lambda, i.e. the p sweep with
 * the conjugated potential multiplier followed by the x sweep with the conjugated
 * kinetic multiplier, and grad[k] (device, WH_MAX_DEG+1 doubles) =
 * dL/da_k(step) = Re <Lambda_in, dU_step/da_k W_step>, where W_next = U_step W_step
 * is the forward state after the step and a_k(step) the coefficients of
 * V(t_step, x).  Lie scheme without damping or Gross-Pitaevskii only.
 */
int wh_wigner2d_adjoint_step(wh_wigner2d* plan, void* Lambda, const void* W_next, int step, double* grad,
                             void* stream);
/* Observables of W (FFT.py:1857-1878) into one row of WH_NOBS doubles (device). */
int wh_wigner2d_observe(wh_wigner2d* plan, const void* W, double* obs_row, void* stream);

/*
 * No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under hipGraph stream capture, serving loops that forbid hipMalloc): the
 * observables of wh_wigner2d_observe without a plan and without any allocation.  The
 * caller passes the state and the host grids (all device memory); the column sums go
 * to a fixed device array of the library (one per GPU, about 82 KiB, held until the
 * process ends), so x_dim is at most WH_NOALLOC_MAX_X (WH_INVALID_SIZE beyond).  Calls
 * that overlap on one GPU share that array and race.  The first call on a device looks
 * the array up (hipGetSymbolAddress, not a stream operation): make it outside capture.
 */
#define WH_NOALLOC_MAX_X 2048
typedef struct wh_observe_desc {
  int struct_size;            /* sizeof(wh_observe_desc) */
  int precision;              /* storage of W: WH_C128, or WH_C64 / WH_C64_WIDE (complex64) */
  int x_dim, p_dim;           /* W is [p_dim][x_dim], FFT ordered */
  const double* x_host;       /* [X] host x grid, as wh_member_desc */
  const double* p_host;       /* [P] */
  const double* dv_host;      /* [X] dPotential(0,x) */
  const double* v0_host;      /* [X] Potential(0,x) */
  double dxdp;                /* dX*dP */
  double inv_2m;              /* 1/(2 mass) */
} wh_observe_desc;
int wh_wigner2d_observe_noalloc(const wh_observe_desc* desc, const void* W, double* obs_row, void* stream);

/* ---- ensembles: many Wigner functions, uneven sizes, one launch per sweep --- */
/*
 * No counterpart in the reference (one grid per class instance there): the
 * reference's per-instance set-up (the wh_wigner2d_desc fields) split into the
 * part shared by all members and the per-member part.
 */
typedef struct {
  int struct_size; /* sizeof(wh_ensemble_desc) */
  int precision;   /* WH_C128 or WH_C64, all members */
  int scheme;      /* WH_SCHEME_LIE or WH_SCHEME_STRANG */
  double dt, hbar, mass;
  int kinetic_poly, kinetic_deg;
  double kin_coef[WH_MAX_DEG + 1];
  int damping;     /* 0 none, 1 Caldeira-Leggett order 3, 2 ODM */
  double gamma_dt;
  int renormalize;
  double gp_coeff;
  int odm_channels;           /* as in wh_wigner2d_desc */
  double odm_weight1, odm_weight2;
} wh_ensemble_desc;

typedef struct {
  int struct_size; /* sizeof(wh_member_desc) */
  int x_dim, p_dim;
  double dx, dp, dtheta, dlambda;
  int pot_deg, pot_rows;
  const double* pot_coef;
  const double *theta_decay, *lambda_decay, *x_mask, *sigma_p, *theta_host, *lambda_host;
  const double *x_host, *p_host, *dv_host, *v0_host;
  double dxdp, inv_2m, dx_host, gp_rho_factor;
  const double* pot_dv; /* [P][X] Bopp difference of V for pot_deg = WH_POT_TABLE */
  const double *odm_g, *odm_x, *odm_th; /* ODM tables, as in wh_wigner2d_desc */
  double odm_sign;
  const double* pot_expr; /* pot_deg = WH_POT_EXPR: as in wh_wigner2d_desc */
  int pot_expr_len;
  int pot_expr_kvn;
  double pot_t_step;
} wh_member_desc;

typedef struct wh_ensemble wh_ensemble;

int wh_ensemble_create(const wh_ensemble_desc* desc, const wh_member_desc* members, int n_members,
                       wh_ensemble** out);
int wh_ensemble_destroy(wh_ensemble* e);
/*
 * As wh_wigner2d_run for every member at once: states[m] is member m's device
 * state, obs is a device array [n_members][n_steps+1][WH_NOBS].
 */
int wh_ensemble_run(wh_ensemble* e, void* const* states, int first_step, int n_steps, double* obs,
                    int normalize_initial, int use_graph, void* stream);
/*
 * Launch variants (index into the kernel variant tables; -1 keeps the current
 * choice): x sweep, p sweep, p sweep with Caldeira-Leggett.  Rejected with
 * WH_INVALID_VALUE if a member's grid does not fit the variant.
 */
int wh_ensemble_set_variants(wh_ensemble* e, int xvar, int pvar, int pclvar);
/* Number of variants of each table and the (radix, lanes) of variant i.  The last
   variant of each table is the throughput form (same radix and lanes as variant 0;
   twiddles from the table, the exchanges between transform passes held in the
   state's storage precision). */
int wh_variant_info(int table, int i, int* radix, int* lanes);
/* Replaces member `member`'s potential coefficient rows ([rows][WH_MAX_DEG+1]
   doubles, host or device memory, copied on `stream`) -- e.g. between the runs of
   an optimisation loop.  No counterpart in the reference (which recompiles). */
int wh_ensemble_set_potential(wh_ensemble* e, int member, const double* coef, int rows, int deg, void* stream);
/* The single-grid plan's underlying ensemble (one member). */
wh_ensemble* wh_wigner2d_ensemble(wh_wigner2d* plan);

/* ---- imaginary-time (Bloch) Gross-Pitaevskii solver ------------------------ */
/*
 * Replaces the per-instance set-up of GPU_Wigner2D_GPitaevskii_Bloch
 * (GPU_Wigner2D_GPitaevskii.py:2064-2103: SetCUDA_Constants, the Kernel_Bloch
 * functions of :88-110, :160-175, :498-517, the Hamiltonian grid of :2094).
 * Host arrays, copied at creation; constants are the kernel-side values.
 */
#define WH_BNHIST 12 /* doubles per Bloch history row */
/* history columns: dt used by the step, <x>, <x^2>, <p>, <p^2>, Hamiltonian_average,
   total energy (with the non-linear term), non-linear energy, purity (re, im),
   1 kept / 0 refused / -1 untested (first iteration), Im of the total energy */
#define WH_BHIST_DT 0
#define WH_BHIST_X 1
#define WH_BHIST_X2 2
#define WH_BHIST_P 3
#define WH_BHIST_P2 4
#define WH_BHIST_H 5
#define WH_BHIST_E 6
#define WH_BHIST_NLE 7
#define WH_BHIST_PURITY 8
#define WH_BHIST_PURITY_IM 9
#define WH_BHIST_KEPT 10
typedef struct {
  int struct_size; /* sizeof(wh_bloch_desc) */
  int x_dim, p_dim; /* powers of two, 128..2048 */
  double dx, dp, dtheta, dlambda, hbar; /* kernel constants */
  /* V(t, x) polynomial rows as in wh_wigner2d_desc, or pot_deg = WH_POT_TABLE with
     pot_sum[P][X] = V(x - hbar theta/2) + V(x + hbar theta/2) at kernel coordinates */
  int pot_deg, pot_rows;
  const double* pot_coef;
  const double* pot_sum;
  int kinetic_deg;            /* K(p) = sum kin_coef[n] p^n */
  double kin_coef[WH_MAX_DEG + 1];
  double gp_coeff;            /* kernel a_GP; 0 = off */
  const double* theta_host;   /* [P] host grids */
  const double* lambda_host;  /* [X] */
  const double* x_host;       /* [X] */
  const double* p_host;       /* [P] */
  const double* k_host;       /* [P] KineticEnergy(p) on the host grid */
  const double* v0_host;      /* [X] Potential(0, x) on the host grid */
  double dxdp, dx_host, dp_host, hbar_host, gp_coeff_host;
} wh_bloch_desc;

typedef struct wh_bloch wh_bloch;

int wh_bloch_create(const wh_bloch_desc* desc, wh_bloch** out);
int wh_bloch_destroy(wh_bloch* plan);
/*
 * Replaces GPU_Wigner2D_GPitaevskii_Bloch.Run (W_ground NULL, :2105-2340) and
 * Run_ExitedState1 (W_ground = the ground state, :2343-2535): n_iter iterations
 * (the reference's timeSteps + 1) from dt.  W (device, [P][X] complex128) holds W_init
 * on entry and the final state on return; W_step (device, same size, caller-owned) is
 * the last kept state (it starts as W_init as given).  hist (device) receives n_iter
 * rows of WH_BNHIST doubles.  Every step is kept or refused on the device (a refused
 * step halves dt and restores W_step), so the call launches the whole loop without
 * waiting on the host.
 */
int wh_bloch_run(wh_bloch* plan, void* W, void* W_step, const void* W_ground, int n_iter, double dt, double* hist,
                 void* stream);

/* ---- Wigner-Dirac 4x4 (GPU_WignerDirac2D_4x4, GPU_WignerDirac2D_4x4.py:1880-3720) ---- */
#define WH_DNOBS 32 /* doubles per Dirac record row */
/* record columns (complex quantities as re, im pairs): the per-step lists of Run
   (:3412-3500): X_Average, P_Average, XP_Average, XX_Average, PP_Average,
   Alpha_1_Average, D_1_Potential_0_Average, X1_D_1_Potential_0_Average,
   P1_D_1_Potential_0_Average, P1_Alpha_1_Average, X1_Alpha_1_Average, Dirac_energy;
   then negativity, transmission, antiParticle_population, norm, raw norm */
#define WH_DOBS_X 0
#define WH_DOBS_P 2
#define WH_DOBS_XP 4
#define WH_DOBS_XX 6
#define WH_DOBS_PP 8
#define WH_DOBS_ALPHA1 10
#define WH_DOBS_DV0 12
#define WH_DOBS_XDV0 14
#define WH_DOBS_PDV0 16
#define WH_DOBS_PALPHA1 18
#define WH_DOBS_XALPHA1 20
#define WH_DOBS_ENERGY 22
#define WH_DOBS_NEG 24
#define WH_DOBS_TRANS 25
#define WH_DOBS_ANTI 26
#define WH_DOBS_NORM 27
#define WH_DOBS_RAWNORM 28

/*
 * Replaces the per-instance set-up of GPU_WignerDirac2D_4x4 (constants of
 * :1929-1942, Compiling_CUDA_functions :2019-2138, Allocate_GPUVariables
 * :1992-2017) for vector potentials A1 = A2 = A3 = 0 (the potential factors are
 * then diagonal in the spinor indices).  Host arrays, copied at creation.
 */
typedef struct {
  int struct_size; /* sizeof(wh_dirac_desc) */
  int precision;   /* WH_C128 or WH_C64 (state); factors are FP64 in both */
  int x_dim, p_dim; /* powers of two, 128..2048 */
  /* kernel constants (the caller applies the reference's '%f' splicing) */
  double dt, dx, dp, dtheta, dlambda, mass, mass_half, c;
  /* Potential_0(t,x): polynomial rows as in wh_wigner2d_desc, or pot_deg =
     WH_POT_TABLE with pot_dv[P][X] = V0(x - theta/2) - V0(x + theta/2) */
  int pot_deg, pot_rows;
  const double* pot_coef;
  const double* pot_dv;
  const double* theta_decay;  /* [P] exp(-dt*D_Theta*theta^2) (both potential halves), NULL = 1 */
  const double* x_mask;       /* [X] edge absorber 1-exp(-((j-X/2)/24)^2) (:111-155), NULL = off */
  const double* x_host;       /* [X] host grids */
  const double* p_host;       /* [P] */
  const double* theta_host;   /* [P] */
  const double* dv0;          /* [X] dV0/dx at kernel x (table mode) */
  const double* vsum;         /* [X] sum_k Potential_k at kernel x (table mode, DiracEnergy) */
  double gamma_dt;            /* CaldeiraDissipatorOrder2_4x4 (:3019-3058): dt*gammaDamping; 0 = off */
  double dxdp;                /* host dX*dP */
  double c_host, mc2_host;    /* host c and mass*c^2 (DiracEnergy, :2777-2805) */
  double transmission_x;      /* x threshold of the transmission record (:72-105: 10) */
  int antiparticle;           /* record AntiParticlePopulation (:2436-2466) */
  int energy;                 /* record DiracEnergy */
  /* position-dependent mass of the potential factors (GPU_WignerDirac2D_MassPotential_4x4,
     Mass(t,x) in place of mass_half, GPU_WignerDirac2D_MassPotential_4x4.py:387-391, 422-427):
     [P][X] F(x - theta/2) - F(x + theta/2) and F(x - theta/2) + F(x + theta/2) with
     F = Mass(x) c^2 dt; NULL = the constant mass_half */
  const double* mass_dm;
  const double* mass_sm;
  /* Gross-Pitaevskii term (MakeGrossPitaevskiiTerms :2695-2715, gpu_sum_axis0 :1791-1817):
     phase dt*a*Re(B- - B+), B-/+ = rho(x -/+ theta/2), rho = Re sum_p trace of the state */
  double gp_coeff;            /* grossPitaevskiiCoefficient (host value); 0 = off */
  const double* lambda_host;  /* [X] host Lambda grid (exp(i lambda theta/2)) */
  /* vector potentials Potential_1..3 as polynomials in x (a_deg[k] = -1: identically zero;
     a_deg[k] bounds the degree over all rows).  Any non-zero one switches the potential step
     to two sweeps (the factors then couple all 16 planes; with mass_dm/mass_sm their mass
     term is Mass(x -/+ theta/2) c^2 dt from those tables).
     Time-independent: a_coef, a_rows = 0.  Time-dependent: a_rows rows of [3][WH_MAX_DEG + 1]
     coefficients in a_coef_rows (row = time index, clamped to a_rows - 1; host memory). */
  int a_deg[3];
  double a_coef[3][WH_MAX_DEG + 1];
  /* ODM damping (dampingModel 'ODM': DampingODM_Kernel, GPU_WignerDirac2D_4x4.py:952-1195,
     applied between the potential factors and Theta_To_P, :3608-3615); needs gamma_dt = 0
     (the two damping models exclude each other).  With A the factor sweeps leave the
     planes in the theta picture and the ODM sweep transforms them back */
  int odm;                    /* 1 = on */
  double odm_gamma;           /* gammaDamping (kernel argument, full precision) */
  double lambda_bar;          /* lambdaBar (kernel constant) */
  /* the KvN variant (GPU_WignerDiracKvN2D_4x4.py): kinetic factors at q = +/- lambda/2 without
     the mass term on the diagonal (:183-189, 283-289; mass_half = 1e-5 there, :1961), potential
     step e^{-dt D_Theta theta^2 + i dt V0'(x) theta} on every plane (KvN_XTheta_Kernel, :356-420;
     V0' from the polynomial rows, dv0 or the program dpot_expr), the step filter always on with projectors at p
     (:775, 3599-3640); V0 only (no A, mass function, GP or damping) */
  int kvn;
  /* antiparticle step filter (antiParticleStepFiltering, :3636-3678): after the potential half,
     W <- Lambda_+(p + lambda/2) W Lambda_+(p - lambda/2) in (p, lambda), then the absorber */
  int step_filter;
  const double* a_coef_rows;  /* [a_rows][3][WH_MAX_DEG + 1], see a_deg */
  int a_rows;
  /* pot_deg = WH_POT_EXPR: V0(t, x) that is not a polynomial in x and depends on t, as the
     expression programs (see WH_EXPR_*) of V0 and of dV0/dx, evaluated on the device at
     t = time index * pot_t_step: the potential step's Bopp difference before each potential
     sweep (with kvn: dV0/dx of the KvN phase), dV0/dx and V0 of the record rows before
     each reduce. */
  const double* pot_expr;
  int pot_expr_len;
  const double* dpot_expr;
  int dpot_expr_len;
  double pot_t_step;
  /* Mass(t, x) that depends on t, as an expression program (instead of mass_dm/mass_sm):
     the library refills both mass tables on the device before each potential sweep, with
     F = Mass(x, t) c dt's kernel constants c and dt, t = time index * mass_t_step */
  const double* mass_expr;
  int mass_expr_len;
  double mass_t_step;
  /* vector potentials that are not polynomials in x: the expression programs of Potential_1..3
     (a_expr_len[k] = 0: identically zero; a_deg[k] = 0 for a non-zero one so that the two
     factor sweeps run), evaluated on the device at x -/+ theta/2 before each potential sweep
     and at x for the energy of each record row, t = time index * a_t_step */
  const double* a_expr[3];
  int a_expr_len[3];
  double a_t_step;
} wh_dirac_desc;

typedef struct wh_dirac wh_dirac;

int wh_dirac_create(const wh_dirac_desc* desc, wh_dirac** out);
int wh_dirac_destroy(wh_dirac* plan);
/*
 * Replaces the time loop of GPU_WignerDirac2D_4x4.Run (:3409-3690): advances the
 * 16 planes W[16][P][X] (device, plane 4i+j = W_(i+1)(j+1)) by n_steps iterations
 * with time indices first_step .. first_step+n_steps-1 (t = dt*index) and writes
 * (n_steps+1) rows of WH_DNOBS doubles to obs (device): row k = the state before
 * iteration first_step+k (row n_steps = the final state).  Each iteration ends with
 * the trace normalisation; W holds the normalised final state on return.
 */
int wh_dirac_run(wh_dirac* plan, void* W, int first_step, int n_steps, double* obs, int use_graph, void* stream);
/* Single sweeps for callers composing their own loop (no normalisation, no record):
   the kinetic factors (DiracPropagator_P_plus/minus_Lambda between the x transforms,
   :3514-3558) and the potential factors at time index `step` with damping and absorber
   (DiracPropagator_X_minus/plus_Theta between the p transforms, :3560-3683). */
int wh_dirac_kinetic(wh_dirac* plan, void* W, void* stream);
int wh_dirac_potential(wh_dirac* plan, void* W, int step, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* WIGNERHIP_H */
