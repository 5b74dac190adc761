"""Benchmark of the Wigner split-operator hot path on one MI355X.

Workload (default): the damped Wigner2D propagator at 1024 x 1024,
x,p in [-12,12), Caldeira-Leggett friction gamma = 0.1 with D_theta = 2 m gamma kT
= 0.1 (kT = 0.5) and sigma(p) = p, double well V = 0.1 x^4 - x^2, dt = 0.005,
cat-state initial W, complex128, Lie step of GPU_Wigner2D_FFT (absorbing mask on,
renormalisation and the Ehrenfest observables every step).  A step is one full
dt of the whole grid; value = cell-steps/s (cells x steps / wall time).

Also reported (extra keys): the same workload in the FP32-bulk form and its
relative L2 error against the FP64 run; the undamped quartic-oscillator step at
1024^2 (FP64), alone and as HBM-resident batches (64 x 1024^2 FP64 against
128 x 1024^2 FP32); the seeded 256-member uneven ensemble (sides 128..1024, FP32,
1000 steps) with its error against the FP64 run of the same ensemble; and the
Wigner-Dirac 4x4 step at 512^2 x 16 planes in FP64 and FP32.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def cfg2(precision="fp64", X=1024):
    """1024^2 CL-damped double well (the Lindblad/Caldeira-Leggett damped workload)."""
    return dict(X=X, P=X, X_amp=12.0, P_amp=12.0, hbar=1.0, mass=1.0, dt=0.005,
                potential="0.1*pow(x,4) - pow(x,2)", dpotential="0.4*pow(x,3) - 2.0*x",
                scheme="lie", D_theta=2.0 * 1.0 * 0.1 * 0.5, gamma=0.1, damping="CaldeiraLeggett",
                fp_damping="fabs(p)", mask=True, precision=precision)


def plain(precision="fp64", X=1024):
    """Undamped quartic oscillator at 1024^2, x,p in [-10,10), no mask."""
    return dict(X=X, P=X, X_amp=10.0, P_amp=10.0, hbar=1.0, mass=1.0, dt=0.01,
                potential="0.5*pow(x,2) + 0.1*pow(x,4)", dpotential="x + 0.4*pow(x,3)",
                scheme="lie", mask=False, precision=precision)


def make(cfg, device):
    from pywignercuda_amd.wigner2d import Wigner2D

    c = dict(cfg)
    args = [c.pop(k) for k in ("X", "P", "X_amp", "P_amp", "hbar", "mass", "dt")]
    return Wigner2D(*args, device=device, **c)


def initial_state(prop, kind):
    import numpy as np

    x = prop.x_host[None, :]
    p = prop.p_host[:, None]
    if kind == "cat":
        g = lambda xx: np.exp(-(xx ** 2) - p ** 2)  # noqa: E731
        W = g(x - 2.0) + g(x + 2.0) + 2.0 * g(x) * np.cos(4.0 * p)
    else:
        W = np.exp(-((x - 2.0) ** 2) - p ** 2) / np.pi
    return W / (W.sum() * prop.dX * prop.dP)


def timed_run(prop, W, steps, warmup):
    """warmup untimed iterations (at least two; the first graph-enabled run captures and
    uploads the step graph, which runs of 33 or more iterations replay), then `steps` timed
    iterations into the same record buffer, allocated before the timing."""
    import torch

    from pywignercuda_amd._lib import WH_NOBS

    warm = max(int(warmup), 2)
    # one record buffer and one time origin for the warm-up and the timed run, as a caller's
    # steady-state loop has them: the timed call re-binds nothing
    obs = torch.zeros((max(steps, warm) + 1, WH_NOBS), dtype=torch.float64, device=W.device)
    prop.run(W, warm, first_step=warm, use_graph=True, obs=obs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prop.run(W, steps, first_step=warm, normalize_initial=False, use_graph=True, obs=obs)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, obs


def kernel_time(fn, reps):
    """Average duration of fn()'s launches on torch's current stream (HIP events)."""
    import torch

    fn()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps * 1e-3


def _gauss(prop_x, prop_p, dX, dP, x0=2.0, p0=0.0, s=1.0):
    import numpy as np

    x, p = prop_x[None, :], prop_p[:, None]
    W = np.exp(-((x - x0) ** 2) / s ** 2 - s ** 2 * (p - p0) ** 2) / np.pi
    return W / (W.sum() * dX * dP)


def timed_ensemble(ens, states, steps, warmup):
    import torch

    from pywignercuda_amd._lib import WH_NOBS

    warm = max(int(warmup), 2)
    ens.run(states, warm, use_graph=True)
    obs = torch.zeros((ens.n, steps + 1, WH_NOBS), dtype=torch.float64, device=ens.device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(states, steps, first_step=warm, normalize_initial=False, use_graph=True, obs=obs)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def hbm_batch(precision, n, steps, warmup, dev):
    """n copies of the plain 1024^2 quartic-oscillator step in one ensemble: the state
    (n x 16 MiB FP64 / 8 MiB FP32) is far beyond the 256 MB last-level cache."""
    from pywignercuda_amd.wigner2d import WignerEnsemble

    c = plain("fp64")
    m = dict(X=c["X"], P=c["P"], X_amp=c["X_amp"], P_amp=c["P_amp"], potential=c["potential"],
             dpotential=c["dpotential"], mask=False)
    ens = WignerEnsemble([m] * n, 1.0, 1.0, c["dt"], scheme="lie", precision=precision, device=dev)
    t = ens.tables[0]
    W0 = _gauss(t.x_host, t.p_host, t.dX, t.dP)
    states = ens.states_to_device([W0] * n)
    wall = timed_ensemble(ens, states, steps, warmup)
    cells = n * c["X"] * c["P"]
    return {"members": n, "cell_steps_per_s": cells * steps / wall, "steps": steps,
            "hbm_frac_algorithmic": cells * steps * (64 if precision == "fp64" else 32) / wall / 1e9 / HBM_PEAK_GBS,
            "state_bytes": cells * (16 if precision == "fp64" else 8)}


def uneven_ensemble_members(seed=20261202, n=256):
    """256 Wigner functions, sides U{128,256,512,1024}, V = a x^4 - b x^2 with a~U[0.05,0.2],
    b~U[0.5,2], Gaussian W0 with x0~U[-3,3], p0~U[-1,1], s~U[0.7,1.3]; box [-10,10)^2."""
    import numpy as np

    rng = np.random.default_rng(seed)
    sides = rng.choice([128, 256, 512, 1024], size=n)
    a, b = rng.uniform(0.05, 0.2, n), rng.uniform(0.5, 2.0, n)
    x0, p0, s = rng.uniform(-3, 3, n), rng.uniform(-1, 1, n), rng.uniform(0.7, 1.3, n)
    members, init = [], []
    for i in range(n):
        tab = np.zeros((1, 9))
        tab[0, 4], tab[0, 2] = a[i], -b[i]
        ai, bi = float(a[i]), float(b[i])
        members.append(dict(X=int(sides[i]), P=int(sides[i]), X_amp=10.0, P_amp=10.0,
                            potential=f"{ai!r}*pow(x,4) - {bi!r}*pow(x,2)",
                            dpotential=f"{4 * ai!r}*pow(x,3) - {2 * bi!r}*x", pot_table=tab, mask=False))
        init.append((x0[i], p0[i], s[i]))
    return members, init


def uneven_ensemble(steps, warmup, dev, precisions=("fp32", "fp32fast", "fp64")):
    """FP32 run of the seeded uneven ensemble (one launch per sweep for all 256 members) and its
    relative L2 error against the FP64 run of the same ensemble."""
    import numpy as np
    import torch

    from pywignercuda_amd.wigner2d import WignerEnsemble

    members, init = uneven_ensemble_members()
    out = {}
    finals = {}
    for prec in precisions:
        ens = WignerEnsemble(members, 1.0, 1.0, 0.01, scheme="lie", precision=prec, device=dev)
        W0 = [_gauss(t.x_host, t.p_host, t.dX, t.dP, *ini) for t, ini in zip(ens.tables, init)]
        warm = ens.states_to_device(W0)  # graph capture and first launches outside the timing
        ens.run(warm, 20, use_graph=True)
        del warm
        states = ens.states_to_device(W0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.run(states, steps, use_graph=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        cells = sum(t.X * t.P for t in ens.tables)
        out[prec] = {"cell_steps_per_s": cells * steps / wall, "wall_s": wall}
        finals[prec] = torch.cat([W.reshape(-1).to(torch.complex128) for W in states]).cpu().numpy()
        del ens, states
    for prec in ("fp32", "fp32fast"):
        if prec not in finals or "fp64" not in finals:
            continue
        d = finals[prec] - finals["fp64"]
        out[prec]["rel_l2_vs_fp64"] = float(np.linalg.norm(d) / np.linalg.norm(finals["fp64"]))
    out["cells"] = int(sum(m["X"] * m["P"] for m in members))
    out["members"] = len(members)
    out["steps"] = steps
    return out


def dirac_workload(precision, steps, dev):
    """Wigner-Dirac 4x4, 512^2 x 16 planes, c = m = 1, V0 = -2/sqrt(x^2+1), dt = 0.002, positive-energy
    Gaussian spinor at x0 = -2, px = 6 (X_amp = 18, P_amp = 16)."""
    import numpy as np
    import torch

    from pywignercuda_amd.dirac import WignerDirac4x4

    plan = WignerDirac4x4(512, 512, 18.0, 16.0, 1.0, 1.0, 0.002,
                          potentials=("-2/sqrt(x*x+1)", "0.", "0.", "0."), precision=precision, device=dev)
    x = plan.x_host[None, :]
    th = plan.theta_host[:, None]

    def spinor(X):
        rho = np.exp(1j * X * 6.0) * np.exp(-0.5 * (X + 2.0) ** 2)
        p0 = np.sqrt(36.0 + 1.0)
        return np.array([rho * (p0 + 1.0), 0 * X, 0 * X, rho * 6.0])

    L, R = spinor(x - 0.5 * th), spinor(x + 0.5 * th)
    W = np.fft.ifft(L[:, None] * np.conj(R[None, :]), axis=2)
    W = W / (np.real(W[0, 0] + W[1, 1] + W[2, 2] + W[3, 3]).sum() * plan.dX * plan.dP)
    Wd = plan.to_device(W)
    plan.run(Wd, 10)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.run(Wd, steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = 512 * 512
    return {"cell_steps_per_s": cells * steps / wall, "steps": steps,
            "hbm_frac_algorithmic": cells * steps * (1024 if precision == "fp64" else 512) / wall / 1e9
            / HBM_PEAK_GBS}


def calibrate_copy(dev):
    """8 plain device copies of 256 MiB: the scale of the HBM byte counters (FETCH_SIZE,
    WRITE_SIZE) against known traffic."""
    import torch

    src = torch.empty(256 << 20, dtype=torch.uint8, device=dev).fill_(1)
    dst = torch.empty_like(src)
    for _ in range(8):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"bytes_read_per_copy": 256 << 20, "bytes_written_per_copy": 256 << 20, "copies": 8}


def _uneven_one(prec):
    def run(steps, warmup, dev):
        r = uneven_ensemble(steps, warmup, dev, precisions=(prec,))
        return dict(r[prec], cells=r["cells"], members=r["members"])
    return run


def damped_expression(steps, warmup, dev):
    """The damped 1024^2 step with a time-dependent non-polynomial potential
    V = 0.1x^4 - x^2 + 0.05 sin(t) e^{-x^2}: an expression program tabulated on the device
    before each p sweep (one expr_table_kernel launch per step)."""
    c = cfg2("fp64")
    c["potential"] = "0.1*pow(x,4) - pow(x,2) + 0.05*sin(t)*exp(-x*x)"
    c["dpotential"] = "0.4*pow(x,3) - 2.0*x - 0.1*x*sin(t)*exp(-x*x)"
    prop = make(c, dev)
    W = prop.to_device(initial_state(prop, "cat"))
    wall, _ = timed_run(prop, W, steps, warmup)
    cells = prop.X * prop.P
    return {"cell_steps_per_s": cells * steps / wall, "steps": steps, "us_per_step": wall / steps * 1e6}


# extra workloads runnable alone (--workload NAME): (runner(steps, warmup, dev), description)
SIDE_WORKLOADS = {
    "batch-fp64": (lambda st, w, d: hbm_batch("fp64", 64, st, w, d), "64 x 1024^2 plain quartic Lie step, FP64"),
    "batch-fp32": (lambda st, w, d: hbm_batch("fp32", 128, st, w, d),
                   "128 x 1024^2 plain quartic Lie step, complex64 state with FP64 arithmetic"),
    "batch-fp32fast": (lambda st, w, d: hbm_batch("fp32fast", 128, st, w, d),
                       "128 x 1024^2 plain quartic Lie step, FP32 arithmetic"),
    "uneven-fp32": (_uneven_one("fp32"), "256-member uneven ensemble (sides 128..1024, seeded), 'fp32'"),
    "uneven-fp32fast": (_uneven_one("fp32fast"), "256-member uneven ensemble, 'fp32fast'"),
    "uneven-fp64": (_uneven_one("fp64"), "256-member uneven ensemble, FP64"),
    "damped-expr": (damped_expression, "damped 1024^2 Lie step, FP64, V(t, x) = 0.1x^4 - x^2 + 0.05 sin(t) e^{-x^2} "
                    "as a device-evaluated expression program"),
    "dirac-fp64": (lambda st, w, d: dirac_workload("fp64", st, d), "Wigner-Dirac 4x4, 512^2 x 16 planes, FP64"),
    "dirac-fp32": (lambda st, w, d: dirac_workload("fp32", st, d), "Wigner-Dirac 4x4, 512^2 x 16 planes, FP32"),
    "adjoint": (lambda st, w, d: adjoint_workload(st, w, d),
                "512^2 differentiable control: forward + adjoint backward (64 Fourier parameters, checkpoints every 50)"),
}


def adjoint_workload(steps, warmup, dev, n_params=64, K=50):
    """The differentiable-control configuration: 512^2, V = 0.25x^4 - x^2 - E(t)x with E(t) a
    64-term Fourier series (N(0, 0.1) coefficients, T = 20), dt = 0.01, loss 1 - overlap with a
    Gaussian at x = +2, checkpoints every K steps.  Times one forward + backward (torch.autograd
    over the adjoint step) after an untimed one."""
    import numpy as np
    import torch

    from pywignercuda_amd.adjoint import fourier_control_table, propagate
    from pywignercuda_amd.wigner2d import Wigner2D

    dt, T, X = 0.01, 20.0, 512
    base = np.zeros(9)
    base[4], base[2] = 0.25, -1.0
    prop = Wigner2D(X, X, 10.0, 10.0, 1.0, 1.0, dt, potential="0.25*pow(x,4) - pow(x,2)",
                    dpotential="pow(x,3) - 2*x", mask=False, device=dev, pot_table=np.tile(base, (steps + 2, 1)))
    rng = np.random.default_rng(20261203)
    params = torch.tensor(rng.normal(0.0, 0.1, n_params), dtype=torch.float64, device=dev, requires_grad=True)
    x, p = prop.x_host[None, :], prop.p_host[:, None]
    W0 = np.exp(-((x + 1.0) ** 2) - p ** 2) / np.pi
    W0 = prop.to_device(W0 / (W0.sum() * prop.dX * prop.dP))
    Wt = torch.tensor(np.exp(-((x - 2.0) ** 2) - p ** 2) / np.pi, dtype=torch.float64, device=dev)
    c = 2 * np.pi * prop.dX * prop.dP

    def grad_once():
        L = 1.0 - c * (propagate(prop, fourier_control_table(params, steps, dt, T, base), W0, steps,
                                 checkpoint_every=K) * Wt).sum()
        (g,) = torch.autograd.grad(L, params)
        return g

    grad_once()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grad_once()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = X * X
    return {"cell_steps_per_s": cells * steps / wall, "forward_backward_s": wall, "steps": steps,
            "params": n_params, "checkpoint_every": K}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(target_seconds=12.0):
    """The oracle (NumPy complex128) on the same workload, bounded samples: one thread with
    numpy.fft (the reference's CPU path runs single-threaded FFTW plans), then all usable host
    cores with scipy.fft workers (capped at 16, the host share of one GPU)."""
    import numpy as np
    import scipy

    from oracle import wigner2d as ow

    c = cfg2()
    orc = ow.Wigner2DOracle(c["X"], c["P"], c["X_amp"], c["P_amp"], c["hbar"], c["mass"], c["dt"],
                            D_theta=c["D_theta"], gamma=c["gamma"],
                            potential=lambda t, x: 0.1 * x ** 4 - x ** 2,
                            dpotential=lambda t, x: 0.4 * x ** 3 - 2.0 * x,
                            fp_damping=lambda p: np.abs(p), damping="CaldeiraLeggett", scheme="lie",
                            mask=True)
    W = ow.cat_state_wigner(orc.g, a=2.0)

    def sample(workers):
        ow.use_fft_workers(workers)
        try:
            orc.run(W, 0)  # multipliers cached, one iteration
            n = 0
            t0 = time.perf_counter()
            while True:
                orc.run(W, 0)  # one Lie iteration (observables + K + V + CL + renorm)
                n += 1
                el = time.perf_counter() - t0
                if el > target_seconds:
                    return n, el
        finally:
            ow.use_fft_workers(None)

    n1, el1 = sample(None)
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    workers = max(1, min(ncores, 16))
    nn, eln = sample(workers)
    cells = c["X"] * c["P"]
    model = _cpu_model()
    return {"value": cells * n1 / el1, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n1} Lie+CL iterations of the 1024^2 damped workload with the NumPy "
                      f"complex128 oracle (numpy.fft, 1 thread) in {el1:.1f} s on {model}",
            "all_cores": {"value": cells * nn / eln, "cores": workers,
                          "sample": f"{nn} iterations, scipy.fft workers={workers} (elementwise "
                                    f"work single-threaded numpy) in {eln:.1f} s"},
            "cpu_model": model, "os_cpu_count": os.cpu_count(), "usable_cores": ncores,
            "numpy": np.__version__, "scipy": scipy.__version__}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--workload", choices=("damped", "plain") + tuple(SIDE_WORKLOADS), default="damped",
                    help="timed workload: the damped 1024^2 step (default) or the undamped quartic step; "
                         "the others run one of the extra workloads alone for --steps steps (profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run 8 plain device copies of 256 MiB (scale for the HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the hot path would shard by ensemble member, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    dev = torch.device("cuda:0")
    if a.workload in SIDE_WORKLOADS:
        fn, desc = SIDE_WORKLOADS[a.workload]
        r = fn(a.steps, a.warmup, dev)
        side = {"metric": "cell-steps/s (one extra workload alone)", "value": r["cell_steps_per_s"],
                "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                "higher_is_better": True, "data": "synthetic", "config": {"workload": desc}, "detail": r}
        if a.calibrate_copy:
            side["copy_calibration"] = calibrate_copy(dev)
        print(json.dumps(side))
        return
    if a.workload == "plain":
        prop = make(plain("fp64"), dev)
        W = prop.to_device(initial_state(prop, "gauss"))
    else:
        prop = make(cfg2("fp64"), dev)
        W = prop.to_device(initial_state(prop, "cat"))
    cells = prop.X * prop.P
    # dominant kernel: the fused p sweep (V step + CL friction + column sums), in the variant
    # the plan runs (tuned choice; the last variant of a table is its throughput form)
    from pywignercuda_amd import _lib
    from pywignercuda_amd.wigner2d import tuned_variants

    cl = a.workload == "damped"
    table = 2 if cl else 1
    vi = tuned_variants("fp64", [prop.X], [prop.P], cl=cl)[table]
    vi = max(vi, 0)
    L = _lib.lib()
    rr, bb = ctypes.c_int(), ctypes.c_int()
    L.wh_variant_info(table, vi, ctypes.byref(rr), ctypes.byref(bb))
    tp = vi == L.wh_variant_info(table, -1, None, None) - 1
    kname = (f"psweep_kernel<double, double, {rr.value}, {bb.value}, {'true' if cl else 'false'}, "
             f"{'true' if tp else 'false'}>")
    # the dominant kernels alone first (on a copy of the state): their launches also bring the
    # GPU out of idle before the headline run, as in a running application
    Wk = W.clone()
    t_p = kernel_time(lambda: prop.potential(Wk, 0, fused=True), a.kernel_reps)
    t_x = kernel_time(lambda: prop.kinetic(Wk), a.kernel_reps)
    del Wk
    wall, obs = timed_run(prop, W, a.steps, a.warmup)
    value = cells * a.steps / wall
    bytes_per_launch = cells * 16 * 2  # read + write the complex128 state once
    achieved = bytes_per_launch / t_p / 1e9
    # HBM bytes per launch of the same kernel and workload from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated against a plain copy)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_wigner2d_cl_1024_fp64.json")
    if a.workload == "damped" and os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["bytes_per_launch"]
    out = {
        "metric": "cell-steps/s at 1024^2: plain FP64 vs batched FP32 @ stated error; HBM GB/s",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": "Wigner2D Lie step + Caldeira-Leggett order-3 friction, 1024x1024, "
                               "x,p in [-12,12), gamma=0.1, kT=0.5 (D_theta=0.1), V=0.1x^4-x^2, "
                               "dt=0.005, cat state, observables + renormalisation every step",
                   "grid": [prop.P, prop.X], "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kname, "kernel_us": t_p * 1e6,
                     "xsweep_us": t_x * 1e6,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
    }
    if not a.no_extra:
        import numpy as np

        ref = W.cpu().numpy()
        f32 = {}
        for prec in ("fp32", "fp32fast"):
            p32 = make(cfg2(prec), dev)
            W32 = p32.to_device(initial_state(p32, "cat"))
            wall32, _ = timed_run(p32, W32, a.steps, a.warmup)
            f32[prec] = (cells * a.steps / wall32,
                         float(np.linalg.norm(W32.cpu().numpy() - ref) / np.linalg.norm(ref)))
        pp = make(plain("fp64"), dev)
        Wp = pp.to_device(initial_state(pp, "gauss"))
        wallp, _ = timed_run(pp, Wp, a.steps, a.warmup)
        out["extra"] = {
            "fp32_cell_steps_per_s": f32["fp32"][0],
            "fp32_rel_l2_vs_fp64": f32["fp32"][1],
            "fp32fast_cell_steps_per_s": f32["fp32fast"][0],
            "fp32fast_rel_l2_vs_fp64": f32["fp32fast"][1],
            "plain_fp64_cell_steps_per_s": cells * a.steps / wallp,
            "plain_fp64_hbm_frac": cells * 64 * a.steps / wallp / 1e9 / HBM_PEAK_GBS,
        }
        del p32, W32, pp, Wp
        torch.cuda.empty_cache()
        bsteps = max(a.steps // 10, 50)
        out["extra"]["hbm_batch_fp64_64x1024^2"] = hbm_batch("fp64", 64, bsteps, 5, dev)
        torch.cuda.empty_cache()
        out["extra"]["hbm_batch_fp32_128x1024^2"] = hbm_batch("fp32", 128, bsteps, 5, dev)
        torch.cuda.empty_cache()
        out["extra"]["hbm_batch_fp32fast_128x1024^2"] = hbm_batch("fp32fast", 128, bsteps, 5, dev)
        torch.cuda.empty_cache()
        b64 = out["extra"]["hbm_batch_fp64_64x1024^2"]["cell_steps_per_s"]
        out["extra"]["batched_fp32_over_fp64"] = out["extra"]["hbm_batch_fp32_128x1024^2"]["cell_steps_per_s"] / b64
        out["extra"]["batched_fp32fast_over_fp64"] = (
            out["extra"]["hbm_batch_fp32fast_128x1024^2"]["cell_steps_per_s"] / b64)
        out["extra"]["uneven_ensemble_256"] = uneven_ensemble(1000, 0, dev)
        torch.cuda.empty_cache()
        # the batched FP32 forms against the plain single-grid FP64 step
        plain64 = out["extra"]["plain_fp64_cell_steps_per_s"]
        out["extra"]["batched_fp32_over_plain_fp64"] = (
            out["extra"]["hbm_batch_fp32_128x1024^2"]["cell_steps_per_s"] / plain64)
        out["extra"]["uneven_fp32_over_plain_fp64"] = (
            out["extra"]["uneven_ensemble_256"]["fp32"]["cell_steps_per_s"] / plain64)
        out["extra"]["dirac_512^2_fp64"] = dirac_workload("fp64", 1000, dev)
        out["extra"]["dirac_512^2_fp32"] = dirac_workload("fp32", 1000, dev)
        out["extra"]["adjoint_512^2_2000_steps"] = adjoint_workload(2000, 0, dev)
    if a.calibrate_copy:
        out["copy_calibration"] = calibrate_copy(dev)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
