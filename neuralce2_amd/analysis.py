"""Mean energy and heat capacity from parallel-tempering energy samples (reference
scripts/mc/ptni_ladder_analyze.py:17-37).

  e      per-site energies (n_samples, R) in eV, one column per temperature
  C_v    = N var(e) / (k_B T)^2 in k_B per site, population variance over the samples (second half of the run)
  peak   of C_v(T): the parabola through the maximum and its two neighbours in log T (the maximum itself at an
         edge of the grid or when the three points are not concave)
"""
from __future__ import annotations

import json
import os

import numpy as np

KB = 8.617333e-5


def heat_capacity(e, T, N, kb=KB):
    """e (n_samples, R) eV/site, T (R,) K -> C_v (R,) in k_B per site."""
    e = np.asarray(e, np.float64)
    T = np.asarray(T, np.float64)
    return N * e.var(axis=0) / (kb * T) ** 2


def peak_interp(T, cv):
    T = np.asarray(T, np.float64)
    cv = np.asarray(cv, np.float64)
    k = int(np.argmax(cv))
    if k == 0 or k == len(T) - 1:
        return float(T[k])
    x = np.log(T[k - 1:k + 2])
    y = cv[k - 1:k + 2]
    if not np.allclose(np.diff(x), x[1] - x[0], rtol=1e-6, atol=0):
        # unequal spacing: vertex of the interpolating parabola through the three points
        A = np.stack([x * x, x, np.ones(3)], 1)
        a, b, _ = np.linalg.solve(A, y)
        return float(np.exp(-b / (2 * a))) if a < 0 else float(T[k])
    curv = y[0] - 2 * y[1] + y[2]
    if curv >= 0:
        return float(T[k])
    h = x[1] - x[0]
    return float(np.exp(x[1] - 0.5 * h * (y[2] - y[0]) / curv))


def analyze_run(out_dir, n_sites, framework="icet"):
    """Read a finished run's outputs (icet: mcmc_result.npz; nce: energies.npz + mcmc_summary.json) and return
    T, <E>(T), C_v(T) from the second half of the samples, the C_v peak and the acceptance rates."""
    if framework == "icet":
        z = np.load(os.path.join(out_dir, "mcmc_result.npz"), allow_pickle=False)
        E = np.asarray(z["energies"], np.float64)
        T = np.asarray(z["temperatures"], np.float64)
        e = E[E.shape[0] // 2:] / n_sites
        acc = float(z["swap_accept_rate"])
        exch = float(z["exchange_accept_rate"])
    else:
        z = np.load(os.path.join(out_dir, "energies.npz"), allow_pickle=False)
        tr = np.asarray(z["traj"], np.float64)                # (R, samples) eV/site
        T = np.asarray(z["temps"], np.float64)
        e = tr[:, tr.shape[1] // 2:].T
        s = json.load(open(os.path.join(out_dir, "mcmc_summary.json")))
        acc = float(s["acc_mean"])
        exch = float(s.get("exchange_acc") or 0.0)
    cv = heat_capacity(e, T, n_sites)
    return dict(T=T, E_mean=e.mean(0), C_v=cv, T_peak=peak_interp(T, cv), cv_max=float(cv.max()),
                acceptance=acc, exchange_acceptance=exch, n_samples=int(e.shape[0]))
