"""Scoring of recovered memberships against ground truth (the evaluation of CPU/Pvalue.cpp:18-235).

For every estimated community i and true community j, the Pearson correlation r of the two
membership columns over the nodes is tested against zero with the statistic
t = r sqrt((n - 2) / (1 - r^2)) on n - 2 degrees of freedom (Calculate_Tstat, Pvalue.cpp:80-85). The
two-sided p-value is I_x((n - 2)/2, 1/2) with x = (n - 2) / ((n - 2) + t^2), the regularised
incomplete beta function (Calculate_Pvalue, Pvalue.cpp:199-235). Here it is evaluated by
scipy.special.betainc instead of the continued fraction of Pvalue.cpp:87-197 (relative tolerance
3e-7 there). Pvalue.cpp:199-235 also has special cases: the Cauchy tail for n - 2 = 1 and the normal
tail beyond 10^7 degrees of freedom. As written, both are overwritten by p = 0.5 (the else branch of
the last test), so they return 1. They are kept as intended here. A column with zero variance has
no correlation and gets p = 1.

A pair (i, j) matches when p < threshold. The recovery ratio is the fraction of true communities
with at least one matching estimate. The error of a recovered community is the smallest mean
absolute difference (1/n) ||pi_hat_i - pi_j||_1 over its matches; the reported error is its mean
over the recovered communities. Memberships are compared as given (normalised Pi-hat against 0/1
or Dirichlet truth).
"""
from __future__ import annotations

import math

import numpy as np


def correlations(est: np.ndarray, truth: np.ndarray) -> np.ndarray:
    """[k_est, k_true] Pearson correlations of the columns of est [n, k_est] and truth [n, k_true]
    (NaN where a column has zero variance)."""
    a = np.asarray(est, np.float64)
    b = np.asarray(truth, np.float64)
    if a.shape[0] != b.shape[0]:
        raise ValueError("est and truth must have the same number of nodes (rows)")
    a = a - a.mean(0)
    b = b - b.mean(0)
    na = np.sqrt((a * a).sum(0))
    nb = np.sqrt((b * b).sum(0))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = (a.T @ b) / np.outer(na, nb)
    r[(na == 0)[:, None] | (nb == 0)[None, :]] = np.nan
    return np.clip(r, -1.0, 1.0)


def pvalues(r: np.ndarray, n: int) -> np.ndarray:
    """Two-sided p-values of correlations r over n nodes (t test on n - 2 degrees of freedom)."""
    from scipy.special import betainc
    r = np.asarray(r, np.float64)
    dof = n - 2
    p = np.ones_like(r)
    ok = np.isfinite(r)
    if dof <= 0:
        return p
    rr = r[ok]
    with np.errstate(divide="ignore", invalid="ignore"):
        t = -np.abs(rr * np.sqrt(dof / (1.0 - rr * rr)))
    if dof == 1:
        half = 0.5 + np.arctan(t) / math.pi
    elif dof > 1e7:
        from scipy.special import erfc
        half = 0.5 * erfc(-t / math.sqrt(2.0))
    else:
        x = np.where(np.isfinite(t), dof / (dof + t * t), 0.0)    # |r| = 1: t = -inf, x = 0, p = 0
        half = 0.5 * betainc(0.5 * dof, 0.5, x)
    p[ok] = 2.0 * half
    return p


def evaluate(est: np.ndarray, truth: np.ndarray, threshold: float = 0.01) -> dict:
    """Matches, recovery ratio and error of estimated memberships est [n, k_est] against truth
    [n, k_true] (rows are nodes)."""
    est = np.asarray(est, np.float64)
    truth = np.asarray(truth, np.float64)
    n = est.shape[0]
    r = correlations(est, truth)
    p = pvalues(r, n)
    match = p < threshold
    l1 = np.full(match.shape, np.inf)   # (1/n) ||est_i - truth_j||_1, for the matching pairs only
    for i, j in zip(*np.nonzero(match)):
        l1[i, j] = np.abs(est[:, i] - truth[:, j]).mean()
    recovered = match.any(0)
    err = l1.min(0)
    return {
        "n": n, "k_est": est.shape[1], "k_true": truth.shape[1], "threshold": threshold,
        "recovery_ratio": float(recovered.mean()),
        "error": float(err[recovered].mean()) if recovered.any() else float("nan"),
        "matches": int(match.sum()),
        "pvalues": p, "correlations": r, "l1": l1,
    }


def read_membership_pairs(path: str, n: int, k: int) -> np.ndarray:
    """Dense [n, k] 0/1 memberships from a file of 1-based `node community` lines (the Pi_true_*
    files shipped with the reference data); pairs outside [1, n] x [1, k] are rejected."""
    pairs = np.loadtxt(path, dtype=np.int64, ndmin=2)
    if pairs.shape[1] < 2:
        raise ValueError(f"{path}: expected two columns")
    node, comm = pairs[:, 0] - 1, pairs[:, 1] - 1
    if (node < 0).any() or (node >= n).any() or (comm < 0).any() or (comm >= k).any():
        raise ValueError(f"{path}: node or community index out of range for n={n}, k={k}")
    out = np.zeros((n, k), np.float64)
    out[node, comm] = 1.0
    return out
