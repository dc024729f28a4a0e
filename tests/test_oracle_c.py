"""The C / OpenMP restatement of the STGD loop (oracle/stgd_cpu.c, the compiled CPU baseline)
against the numpy oracle (core.stgd): two restatements written apart agree to rounding on Phi and
on the ||dPhi||^2 log, over odd k (edge tiles of its register blocking), partial and single-node
minibatches, a minibatch larger than the partition, both Newton-Schulz degrees, both coefficient
scalings and theta != 1; and its result does not depend on the thread count."""
import os
import subprocess

import numpy as np
import pytest

from oracle import core

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.fixture(scope="module")
def cstgd():
    from oracle import cstgd as mod
    if not mod.available():
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    return mod


def _inputs(seed, n, k):
    rng = np.random.default_rng(seed)
    Y = [rng.standard_normal((n, k)) * 0.3 for _ in range(3)]
    mu = [rng.standard_normal(k) * 0.1 for _ in range(3)]
    return Y, mu, core.init_phi(seed + 1, k)


CASES = [  # n, k, batch, degree, scaling, theta
    (300, 9, 64, 3, "paper", 1.0),       # partial last minibatch
    (257, 17, 50, 5, "gpu", 0.7),        # quintic, the GPU coefficients, theta
    (100, 33, 1000, 3, "paper", 1.0),    # minibatch larger than the partition
    (64, 8, 1, 3, "paper", 1.0),         # one node per step
    (700, 130, 300, 3, "paper", 1.0),    # full 6 x 8 register tiles plus edges
]


@pytest.mark.parametrize("n,k,batch,degree,scaling,theta", CASES)
def test_c_restatement_matches_numpy_oracle(cstgd, n, k, batch, degree, scaling, theta):
    Y, mu, phi0 = _inputs(n + k, n, k)
    kw = dict(alpha0=1.5, eta=0.05, batch=batch, epochs=2, theta=theta, ns_iters=2, ns_degree=degree,
              scaling=scaling)
    ref, ref_log = core.stgd(*Y, *mu, phi0, **kw)
    phi, log, phases = cstgd.stgd(*Y, *mu, phi0, **kw)
    assert np.abs(ref - phi0).max() > 1e-3                 # the steps moved Phi
    assert np.abs(phi - ref).max() < 1e-12 * np.abs(ref).max()
    assert log.shape == ref_log.shape
    assert np.abs(log - ref_log).max() < 1e-10 * np.abs(ref_log).max()
    assert set(phases) == set(cstgd.PHASES) and min(phases.values()) >= 0.0


def test_c_restatement_thread_count_independent(cstgd):
    Y, mu, phi0 = _inputs(5, 900, 70)
    kw = dict(alpha0=2.0, eta=0.02, batch=256, epochs=1)
    one = cstgd.stgd(*Y, *mu, phi0, threads=1, **kw)
    four = cstgd.stgd(*Y, *mu, phi0, threads=4, **kw)
    assert np.array_equal(one[0], four[0]) and np.array_equal(one[1], four[1])


def test_c_restatement_fixed_steps_and_rejects_bad_arguments(cstgd):
    Y, mu, phi0 = _inputs(7, 50, 6)
    ref, ref_log = core.stgd(*Y, *mu, phi0, alpha0=1.0, eta=0.1, batch=16, epochs=1, steps=9)
    phi, log, _ = cstgd.stgd(*Y, *mu, phi0, alpha0=1.0, eta=0.1, batch=16, steps=9)   # wraps past the end of X
    assert log.shape == (9,) and np.abs(phi - ref).max() < 1e-12
    with pytest.raises(ValueError):
        cstgd.stgd(*Y, *mu, phi0, alpha0=1.0, eta=0.1, batch=16, ns_degree=4)
    with pytest.raises(ValueError):
        cstgd.stgd(*Y, *mu, phi0, alpha0=1.0, eta=0.1, batch=0)
    with pytest.raises(ValueError):
        cstgd.stgd(Y[0][:, :5], Y[1], Y[2], *mu, phi0, alpha0=1.0, eta=0.1, batch=16)


def test_pipeline_timing_script_c1(cstgd):
    """The per-phase CPU timing of the restatements (oracle/time_pipeline.py) at C1: every phase is
    timed and the two STGD restatements agree on Phi after C1's 200 epochs."""
    from oracle import time_pipeline
    out = time_pipeline.run("c1", 1)
    assert {"whiten", "yvectors", "stgd_numpy", "stgd_c_openmp", "eigen_post"} <= set(out["phases_s"])
    assert out["stgd_c_vs_numpy_max_abs"] < 1e-10
    assert out["end_to_end_s"] > 0 and out["nnz_xa"] > 0
