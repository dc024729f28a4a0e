"""The evaluation harness (mmsb_amd.evaluate): p-values of the membership correlation t test
(CPU/Pvalue.cpp:80-85,199-235) against scipy.stats, matching and recovery on the generator's ground
truth, and the reader of the Pi_true membership files (host code, CPU)."""
import importlib
import os

import numpy as np
import pytest

REF_DATA = "/root/reference/CommunityDetectionSoftware_CPU/data"


@pytest.fixture(scope="module")
def ev(mmsb):
    return importlib.import_module("mmsb_amd.evaluate")


@pytest.mark.parametrize("n", [3, 4, 10, 57, 1000])
def test_pvalues_match_pearsonr(ev, n):
    from scipy import stats
    rng = np.random.default_rng(n)
    for trial in range(20):
        x = rng.standard_normal(n)
        y = rng.standard_normal(n) + (trial % 5 - 2) * 0.3 * x    # correlations of both signs
        r = ev.correlations(x[:, None], y[:, None])[0, 0]
        ref = stats.pearsonr(x, y)
        assert r == pytest.approx(ref.statistic, rel=1e-12, abs=1e-14)
        assert ev.pvalues(np.array([[r]]), n)[0, 0] == pytest.approx(ref.pvalue, rel=1e-9, abs=1e-300)


def test_pvalue_edge_cases(ev):
    # zero-variance columns have no correlation: p = 1; |r| = 1: p = 0; two nodes: no test
    x = np.arange(10.0)
    p = ev.pvalues(ev.correlations(np.stack([x, np.ones(10), 2 * x + 1], 1), x[:, None]), 10)
    assert p[1, 0] == 1.0 and p[0, 0] == 0.0 and p[2, 0] == 0.0
    assert ev.pvalues(np.array([[0.5]]), 2)[0, 0] == 1.0
    # the normal tail beyond 1e7 degrees of freedom agrees with the t test just below it
    r = np.array([[1e-3]])
    assert ev.pvalues(r, 10_000_010)[0, 0] == pytest.approx(ev.pvalues(r, 9_999_990)[0, 0], rel=1e-3)


def test_recovery_on_generator_truth(mmsb, ev):
    g = mmsb.generate(300, 300, 300, 300, 6, 0.8, 0.5, 0.02, seed=3, keep_truth=True)
    pi = g["pi"]
    rng = np.random.default_rng(0)
    perfect = ev.evaluate(pi, pi)
    assert perfect["recovery_ratio"] == 1.0 and perfect["error"] == 0.0
    perm = rng.permutation(6)
    noisy = np.clip(pi[:, perm] + 0.05 * rng.standard_normal(pi.shape), 0, None)
    res = ev.evaluate(noisy, pi)
    assert res["recovery_ratio"] == 1.0 and 0.0 < res["error"] < 0.06
    assert all(res["pvalues"][np.argsort(perm)[j], j] < 1e-10 for j in range(6))   # each column's own match
    junk = ev.evaluate(rng.random(pi.shape), pi)
    assert junk["recovery_ratio"] < 0.5


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference data not present")
def test_read_pi_true(ev):
    path = os.path.join(REF_DATA, "Pi_true_x.txt")
    pairs = np.loadtxt(path, dtype=np.int64)
    n, k = int(pairs[:, 0].max()), int(pairs[:, 1].max())
    M = ev.read_membership_pairs(path, n, k)
    assert M.shape == (n, k) and M.sum() == len(np.unique(pairs, axis=0))
    with pytest.raises(ValueError):
        ev.read_membership_pairs(path, n - 1, k)
