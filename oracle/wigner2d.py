"""Oracle: NumPy complex128 restatement of the Wigner2D split-operator step.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Never shipped, never the
thing measured.

Follows, line by line in behaviour (not in code):

* grids and FFT ordering ........ GPU_Wigner2D_FFT.py:549-596 (SetPhaseSpaceBox2D/Grid2D)
* kernel coordinate maps ........ GPU_Wigner2D_FFT.py:104-114 (i=(row+P/2)%P, x=dx*(j-X/2))
* axis transforms ............... GPU_Wigner2D_FFT.py:788-801, cufft_wrapper.py:183-250,396-403
                                  (forward e^{-i}, inverse unnormalised then *=1/N)
* kinetic multiplier (Lie) ...... GPU_Wigner2D_FFT.py:131-158 (expPLambda: e^{-D_L l^2} e^{-i dt l p/m})
* kinetic multiplier (Strang) ... GPU_Wigner2D_GPitaevskii.py:117-158 (expPLambdaKinetic)
* potential multiplier .......... GPU_Wigner2D_FFT.py:80-126 (full dt) and
                                  GPU_Wigner2D_GPitaevskii.py:60-86 (dt/2, D_Theta/2)
* absorbing mask ................ GPU_Wigner2D_FFT.py:121-123 (10 dx), :514-516 (20 dx, GP kernel)
* Gross-Pitaevskii .............. GPU_Wigner2D_FFT.py:435-519, 1687-1705, 1845-1848;
                                  GPU_Wigner2D_GPitaevskii.py:466-496, 1743-1748, 1904-1956
* Caldeira-Leggett order 3 ...... GPU_Wigner2D_FFT.py:399-431, 1135-1157, 1903-1913
* ODM Lindblad (order 1) ........ GPU_Wigner2D_FFT.py:1196-1244, 1516-1634, 1907-1910; two
                                  channels of GPU_Wigner2D_ODM2, :1973-2230
* Lie driver / observables ...... GPU_Wigner2D_FFT.py:1754-1964
* Strang driver / observables ... GPU_Wigner2D_GPitaevskii.py:1809-2055
* '%f' constant splicing ........ GPU_Wigner2D_FFT.py:833-842 (optional emulation)

Pins (tests/test_oracle.py): the notebook known answers of CPU_GPU_Wigner2D.ipynb
(through oracle.cpu_notebook, which this module must agree with), the harmonic
oscillator / free-particle closed forms of the split step, norm conservation,
Hermitian symmetry, the shift theorem for the GP densities, the Ehrenfest
identity for the CL friction, and the algebraic expansion of the CL composite.
"""
from __future__ import annotations

import numpy as np

# FFT backend of the oracle: numpy.fft (pocketfft, one thread: the reference's CPU path uses
# single-threaded pyFFTW plans, CPU_GPU_Wigner2D.ipynb cells 6-11) or, for the all-cores CPU
# baseline, scipy.fft with `workers` threads.  Same transform definition either way.
_FFT_WORKERS = None


def use_fft_workers(workers=None):
    """None: numpy.fft on one thread; n: scipy.fft with n worker threads."""
    global _FFT_WORKERS
    _FFT_WORKERS = None if workers is None else int(workers)


def fft(a, axis=-1):
    if _FFT_WORKERS is None:
        return np.fft.fft(a, axis=axis)
    import scipy.fft

    return scipy.fft.fft(a, axis=axis, workers=_FFT_WORKERS)


def ifft(a, axis=-1):
    if _FFT_WORKERS is None:
        return np.fft.ifft(a, axis=axis)
    import scipy.fft

    return scipy.fft.ifft(a, axis=axis, workers=_FFT_WORKERS)


def percent_f(v: float) -> float:
    """Value of a constant after '%f' splicing into CUDA source (6 decimals).

    GPU_Wigner2D_FFT.py:833-842 builds '__constant__ double dt=%f;' etc.
    """
    return float("%f" % float(v))


def fft_order_axis(n: int, step: float) -> np.ndarray:
    """Kernel-side coordinate of FFT-ordered index k: step*(((k+n/2)%n) - n/2).

    GPU_Wigner2D_FFT.py:110-114 (i=(blockIdx.y+P/2)%P; theta=dtheta*(i-0.5*P)).
    """
    k = np.arange(n)
    return step * (((k + n // 2) % n) - 0.5 * n)


def host_axis(amplitude: float, n: int, step: float) -> np.ndarray:
    """fftshift(linspace(-A, A-d, n)) as built on the host, GPU_Wigner2D_FFT.py:571-581."""
    return np.fft.fftshift(np.linspace(-amplitude, amplitude - step, n))


class Grid:
    """Phase-space box, GPU_Wigner2D_FFT.py:549-596.  Layout W[p_index, x_index]."""

    def __init__(self, X: int, P: int, X_amp: float, P_amp: float):
        self.X, self.P = int(X), int(P)
        self.X_amp, self.P_amp = float(X_amp), float(P_amp)
        self.dX = 2.0 * X_amp / float(X)
        self.dP = 2.0 * P_amp / float(P)
        self.dTheta = 2.0 * np.pi / (2.0 * P_amp)
        self.dLambda = 2.0 * np.pi / (2.0 * X_amp)
        self.Theta_amp = self.dTheta * P / 2.0
        self.Lambda_amp = self.dLambda * X / 2.0
        # host grids (full precision), FFT ordered, broadcast shapes [1,X] / [P,1]
        self.x_host = host_axis(self.X_amp, self.X, self.dX)[None, :]
        self.p_host = host_axis(self.P_amp, self.P, self.dP)[:, None]
        self.theta_host = host_axis(self.Theta_amp, self.P, self.dTheta)[:, None]
        self.lambda_host = host_axis(self.Lambda_amp, self.X, self.dLambda)[None, :]


class Wigner2DOracle:
    """Restatement of GPU_Wigner2D_FFT (scheme='lie') and
    GPU_Wigner2D_GPitaevskii (scheme='strang').

    Parameters mirror the reference constructor (GPU_Wigner2D_FFT.py:1718-1720)
    plus the attributes the reference reads off ``self``:

    potential(t, x), dpotential(t, x): vectorised callables (the reference evals
        potentialString / dPotentialString with numpy, :674-692, and splices the
        same string into the CUDA kernel, :871-873).
    kinetic(p): only used by the Strang scheme (expPLambdaKinetic); the Lie scheme
        hard-codes dt*lambda*p/mass (FFT.py:153).
    fp_damping(p): f(p) of the CL friction sigma(p)=sign(p) f(p) (FFT.py:399-431).
    damping: None or 'CaldeiraLeggett'.
    gp_coeff: a_GP (grossPitaevskiiCoefficient / GPitaevskiiCoeff).
    mask: apply the absorbing edge mask of the potential kernel (always on in the
        reference; exposed so closed-form tests can switch it off).
    truncate_constants: emulate the '%f' splicing of the kernel constants.
    time_dependent: V depends on t (otherwise the multipliers are cached).
    kvn: the classical KvN potential step of GPU_FokkerPlank2D_FFT (exp_PotentialKvN,
        FFT.py:27-75, called at :2401): phase dt * (-dPotential(t, x) * theta), the same
        D_Theta decay and edge mask; with renormalize (its 'Wigner' normalization, :2408-2410).
    renormalize: None = after every step when gamma != 0 (FFT.py:1912-1913).
    """

    def __init__(self, X, P, X_amp, P_amp, hbar, mass, dt, D_theta=0.0, D_lambda=0.0,
                 gamma=0.0, potential=None, dpotential=None, kinetic=None,
                 fp_damping=None, damping=None, gp_coeff=0.0, scheme="lie",
                 mask=True, truncate_constants=False, time_dependent=False, L_material=None,
                 odm_weights=None, kvn=False, renormalize=None):
        self.g = Grid(X, P, X_amp, P_amp)
        self.hbar, self.mass, self.dt = float(hbar), float(mass), float(dt)
        self.D_theta, self.D_lambda, self.gamma = float(D_theta), float(D_lambda), float(gamma)
        self.potential = potential
        self.dpotential = dpotential if dpotential is not None else (lambda t, x: 0.0 * x)
        self.kinetic = kinetic if kinetic is not None else (lambda p: 0.5 * p * p / self.mass)
        self.fp_damping = fp_damping
        self.damping = damping
        self.gp_coeff = float(gp_coeff)
        self.L_material = L_material  # ODM (FFT.py:1196-1197)
        # ODM channels: (1,) for GPU_Wigner2D_FFT.Run (weight 1, sign 1, FFT.py:1907-1910);
        # GPU_Wigner2D_ODM2 applies (w1, sign 1) then (w2, sign -1) (FFT.py:2173-2181)
        self.odm_weights = (1.0,) if odm_weights is None else tuple(odm_weights)
        self.scheme = scheme
        self.mask = mask
        self.time_dependent = bool(time_dependent)
        self.kvn = bool(kvn)
        self.renormalize = renormalize
        self._cache = {}
        tr = percent_f if truncate_constants else float
        g = self.g
        # kernel-side constants (spliced into the CUDA source in the reference)
        self.k_dt, self.k_dx, self.k_dp = tr(dt), tr(g.dX), tr(g.dP)
        self.k_dtheta, self.k_dlambda = tr(g.dTheta), tr(g.dLambda)
        self.k_mass, self.k_hbar = tr(mass), tr(hbar)
        self.k_Dtheta, self.k_Dlambda = tr(D_theta), tr(D_lambda)
        self.k_agp = tr(gp_coeff)
        # kernel-side coordinates
        self.kx = fft_order_axis(g.X, self.k_dx)[None, :]
        self.kp = fft_order_axis(g.P, self.k_dp)[:, None]
        self.ktheta = fft_order_axis(g.P, self.k_dtheta)[:, None]
        self.klambda = fft_order_axis(g.X, self.k_dlambda)[None, :]

    # ------------------------------------------------------------------ multipliers
    def edge_mask(self, width_cells: float) -> np.ndarray:
        """(1-e^{-(x-xm)^2/(w dx)^2})(1-e^{-(x+xm)^2/(w dx)^2}), FFT.py:121-123."""
        x = self.kx
        xm = self.k_dx * (self.g.X - 1.0) / 2.0
        w2 = (width_cells * self.k_dx) ** 2
        return (1.0 - np.exp(-((x - xm) ** 2) / w2)) * (1.0 - np.exp(-((x + xm) ** 2) / w2))

    def kinetic_multiplier(self) -> np.ndarray:
        """Multiplier on the (p, lambda) grid."""
        lam, p = self.klambda, self.kp
        r = np.exp(-self.k_Dlambda * lam * lam)  # FFT.py:154 (no dt)
        if self.scheme == "lie":
            phase = self.k_dt * lam * p / self.k_mass  # FFT.py:153
        else:  # GP.py:130-151
            h = self.k_hbar
            phase = self.k_dt * (self.kinetic(p + h * lam / 2.0) - self.kinetic(p - h * lam / 2.0)) / h
        return r * np.cos(phase) - 1j * (r * np.sin(phase))

    def potential_multiplier(self, t: float) -> np.ndarray:
        """Multiplier on the (theta, x) grid (FFT.py:80-126; GP.py:60-86)."""
        th, x, h = self.ktheta, self.kx, self.k_hbar
        frac = 1.0 if self.scheme == "lie" else 0.5
        if self.kvn:  # Potential_XTheta = -dPotential(t, x) * theta (FFT.py:42-47, 60)
            dV = -self.dpotential(t, x) * th
        else:
            dV = (self.potential(t, x - h * th / 2.0) - self.potential(t, x + h * th / 2.0)) / h
        phase = frac * self.k_dt * dV
        r = np.exp(-frac * self.k_Dtheta * th * th * self.k_dt)
        m = r * np.cos(phase) - 1j * (r * np.sin(phase))
        if self.mask:
            m = m * self.edge_mask(10.0)
        return m

    # ------------------------------------------------------------------ Gross-Pitaevskii
    def gp_density(self, W: np.ndarray) -> np.ndarray:
        """rho(x) = Re sum_p W  (FFT.py:435-455; GP.py multiplies by dP, :1747)."""
        rho = np.real(W.sum(axis=0))
        if self.scheme == "strang":
            rho = rho * self.g.dP
        return rho

    def gp_shifted(self, rho: np.ndarray):
        """B-/+ (theta, x) = rho(x -/+ theta/2) by the shift theorem, FFT.py:1687-1705."""
        g = self.g
        rt = fft(rho.astype(np.complex128))[None, :]
        ph = np.exp(0.5j * g.lambda_host * g.theta_host)  # FFT.py:594-595 (no hbar)
        Bm = ifft(rt / ph, axis=1)
        Bp = ifft(rt * ph, axis=1)
        return Bm, Bp

    def gp_multiplier(self, Bm, Bp) -> np.ndarray:
        frac = 1.0 if self.scheme == "lie" else 0.5
        phase = -frac * self.k_dt * self.k_agp * np.real(Bm - Bp) / self.k_hbar
        m = np.cos(phase) + 1j * np.sin(phase)
        if self.scheme == "lie":  # FFT.py:514-516 (20 dx); GP.py kernel has no mask
            if self.mask:
                m = m * self.edge_mask(20.0)
        return m

    # ------------------------------------------------------------------ damping
    def sigma(self) -> np.ndarray:
        """sigma(p) = f(p) for p>=0, -f(p) for p<0, kernel coordinates (FFT.py:424-429)."""
        p = self.kp
        f = self.fp_damping(p) * np.ones_like(p)
        return np.where(p >= 0.0, f, -f)

    def cl_operator(self, W: np.ndarray) -> np.ndarray:
        """D(W) = F_p^{-1}[ theta * F_p[ sigma(p) W ] ] (Theta_fp_Damping, FFT.py:1135-1143)."""
        LW = W * self._cached("sigma", self.sigma)
        LW = fft(LW, axis=0) * self.g.theta_host  # host Theta grid
        return ifft(LW, axis=0)

    def caldeira_order3(self, W: np.ndarray) -> np.ndarray:
        """FFT.py:1145-1157: W + 2aD W + (10/3)a^2 D^2 W + (4/3)a^3 D^3 W, a = i dt gamma."""
        a = 1j * self.dt * self.gamma
        LW = W.copy()
        LW = LW + (2.0 / 3.0) * a * self.cl_operator(W)
        LW = LW + (2.0 / 2.0) * a * self.cl_operator(LW)
        return W + 2.0 * a * self.cl_operator(LW)

    # ------------------------------------------------------------------ ODM Lindblad dissipator
    def odm_tables(self):
        """SetA_ODM (FFT.py:1206-1244) on host grids: g(p) = sqrt(L f(|p| + hbar/(2L))) (g_ODM,
        :1196-1200) at p -/+ hbar lambda/2, its signed form, and the cos/sin of x +/- hbar theta/2."""
        g = self.g
        h, L = self.hbar, self.L_material

        def gfun(q):
            return np.sqrt(L * self.fp_damping(np.abs(q) + h / (2.0 * L)))

        qp = g.p_host + 0.5 * h * g.lambda_host
        qm = g.p_host - 0.5 * h * g.lambda_host
        gp_, gm_ = gfun(qp), gfun(qm)
        xp = g.x_host + 0.5 * h * g.theta_host
        xm = g.x_host - 0.5 * h * g.theta_host
        return dict(gp=gp_, gm=gm_, gsp=gp_ * np.sign(qp), gsm=gm_ * np.sign(qm),
                    cp=np.cos(xp), cm=np.cos(xm), sp=np.sin(xp), sm=np.sin(xm))

    def odm_operator(self, W, weight=1.0, sign=1.0, n=1.0):
        """_Lindbladian_ODM (FFT.py:1516-1634): sum over six terms of
        h_k(theta, x) F_p F_x^-1 [ g_k(p, lambda) F_x W ], back to p, times dt 2 gamma weight / n."""
        T = self._cached("odm", self.odm_tables)

        def term(gk):
            Y = ifft(fft(W, axis=1) * gk, axis=1)
            return fft(Y, axis=0)

        LW = term(T["gp"] * T["gm"]) * (T["cp"] * T["cm"])
        LW = LW + term(T["gp"] * T["gsm"]) * T["cm"] * T["sp"] * (1j * sign)
        LW = LW + term(T["gsp"] * T["gm"]) * T["sm"] * T["cp"] * (-1j * sign)
        LW = LW + term(T["gsp"] * T["gsm"]) * (T["sp"] * T["sm"])
        LW = LW + term(-0.5 * T["gp"] ** 2)
        LW = LW + term(-0.5 * T["gm"] ** 2)
        return ifft(LW, axis=0) * (self.dt * 2.0 * self.gamma * weight / n)

    # ------------------------------------------------------------------ sweeps
    def _cached(self, key, fn):
        if key not in self._cache:
            self._cache[key] = fn()
        return self._cache[key]

    def kinetic_step(self, W):
        K = self._cached("K", self.kinetic_multiplier)
        return ifft(fft(W, axis=1) * K, axis=1)

    def potential_step(self, W, t, gp_terms=None):
        if self.time_dependent:
            M = self.potential_multiplier(t)
        else:
            M = self._cached("V", lambda: self.potential_multiplier(0.0))
        B = fft(W, axis=0) * M
        if gp_terms is not None:
            B = B * self.gp_multiplier(*gp_terms)
        return ifft(B, axis=0)

    # ------------------------------------------------------------------ observables
    def observables(self, W: np.ndarray) -> dict:
        """dXdP * sum(W*A) for the reference's Ehrenfest set (FFT.py:1857-1878)."""
        g = self.g
        x, p = g.x_host, g.p_host
        dV = self.dpotential(0.0, x) * np.ones_like(x)
        H = p ** 2 / (2.0 * self.mass) + self.potential(0.0, x)  # FFT.py:1747
        dxdp = g.dX * g.dP
        Wr = np.real(W)
        s = lambda A: dxdp * float(np.sum(Wr * A))
        return {
            "X": s(x), "X2": s(x * x), "P": s(p), "P2": s(p * p), "XP": s(x * p),
            "dV": s(dV), "PdV": s(p * dV), "XdV": s(x * dV), "H": s(H),
            "neg": dxdp * float(np.sum(np.where(Wr < 0.0, Wr, 0.0))),
            "norm": dxdp * float(np.sum(Wr)),
        }

    # ------------------------------------------------------------------ drivers
    def normalise(self, W):
        W = np.array(W, dtype=np.complex128)
        return W / (W.sum() * self.g.dX * self.g.dP)

    def run(self, W_init, time_steps: int, record=True):
        """Runs time_steps+1 iterations as the reference loops do (FFT.py:1776, GP.py:1830).

        Returns (W_end complex128, observables dict of arrays, extras dict)."""
        W = self.normalise(W_init)
        keys = ["X", "X2", "P", "P2", "XP", "dV", "PdV", "XdV", "H", "neg", "norm"]
        obs = {k: [] for k in keys}
        extras = {"ProbabilitySquare": [], "time": []}
        dxdp = self.g.dX * self.g.dP
        for ti in range(time_steps + 1):
            t = ti * self.dt
            extras["time"].append(t)
            if self.scheme == "lie":
                gp = None
                if self.gp_coeff != 0.0:
                    gp = self.gp_shifted(self.gp_density(W))
                if record:
                    for k, v in self.observables(W).items():
                        obs[k].append(v)
                W = self.kinetic_step(W)
                W = self.potential_step(W, t, gp)
            else:
                gp = None
                if self.gp_coeff != 0.0:
                    gp = self.gp_shifted(self.gp_density(W))
                W = self.potential_step(W, t, gp)
                W = self.kinetic_step(W)
                if self.gp_coeff != 0.0:
                    rho = self.gp_density(W)
                    extras["ProbabilitySquare"].append(float(np.sum(rho ** 2) * self.g.dX))
                    gp = self.gp_shifted(rho)
                W = self.potential_step(W, t, gp)
            if self.gamma != 0.0:
                if self.damping == "CaldeiraLeggett":
                    W = self.caldeira_order3(W)
                elif self.damping == "ODM":  # Lindbladian_ODM_Order1 per channel (FFT.py:1907-1910, 2173-2181)
                    for k, w in enumerate(self.odm_weights):
                        W = W + self.odm_operator(W, weight=w, sign=1.0 if k == 0 else -1.0)
            if self.gamma != 0.0 if self.renormalize is None else self.renormalize:
                W = W / (W.sum() * dxdp)
            if self.scheme == "strang" and record:
                for k, v in self.observables(W).items():
                    obs[k].append(v)
        return W, {k: np.array(v) for k, v in obs.items()}, extras


# ---------------------------------------------------------------------- initial states
def gaussian_wigner(grid: Grid, x0=0.0, p0=0.0, s=1.0, hbar=1.0) -> np.ndarray:
    """(1/(pi hbar)) exp(-(x-x0)^2/s^2 - s^2 (p-p0)^2/hbar^2) on the host grid (FFT ordered)."""
    x, p = grid.x_host, grid.p_host
    return (1.0 / (np.pi * hbar)) * np.exp(-((x - x0) ** 2) / s ** 2 - (s ** 2) * (p - p0) ** 2 / hbar ** 2)


def harmonic_oscillator_wigner(grid: Grid, n, omega, x0, p0, mass=1.0):
    """Propagator_Base.Wigner_HarmonicOscillator, GPU_Wigner2D_FFT.py:715-728."""
    from scipy.special import eval_laguerre

    x, p = grid.x_host, grid.p_host
    r2 = mass * omega ** 2 * (x - x0) ** 2 + (p - p0) ** 2 / mass
    W = (-1) ** n * eval_laguerre(n, 2 * r2) * np.exp(-r2)
    return W / (np.sum(W) * grid.dX * grid.dP)


def cat_state_wigner(grid: Grid, a=2.0, s=1.0, hbar=1.0) -> np.ndarray:
    """g(x-a,p)+g(x+a,p)+2 g(x,p) cos(2 a p/hbar), g the s-Gaussian; normalised numerically."""
    x, p = grid.x_host, grid.p_host
    g = lambda xx: np.exp(-(xx ** 2) / s ** 2 - (s ** 2) * p ** 2 / hbar ** 2)
    W = g(x - a) + g(x + a) + 2.0 * g(x) * np.cos(2.0 * a * p / hbar)
    return W / (np.sum(W) * grid.dX * grid.dP)
