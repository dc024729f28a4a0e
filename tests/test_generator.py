"""The library's seeded MMSB generator reproduces the oracle bit-for-bit (host code, CPU)."""
import numpy as np
import pytest

from oracle import gen

CASES = [
    dict(n_x=120, n_a=100, n_b=90, n_c=110, k=5, alpha0=1.0, p_in=0.4, p_out=0.03, seed=0),
    dict(n_x=200, n_a=200, n_b=200, n_c=200, k=10, alpha0=0.5, p_in=0.2, p_out=0.01, seed=7, degree_corrected=True),
    dict(n_x=300, n_a=250, n_b=280, n_c=260, k=8, alpha0=2.0, p_in=0.05, p_out=0.002, seed=3, mode="balls"),
    dict(n_x=400, n_a=400, n_b=400, n_c=400, k=50, alpha0=2.0, p_in=0.08, p_out=0.0016, seed=11,
         degree_corrected=True, mode="balls"),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c.get('mode','auto')}-k{c['k']}-s{c['seed']}")
def test_generator_bit_exact_vs_oracle(mmsb, case):
    got = mmsb.generate(**case, keep_truth=True, threads=3)
    g = gen.generate(gen.GenParams(**case))
    assert np.array_equal(got["theta"], g.theta)
    assert np.array_equal(got["pi"], g.pi)
    assert np.array_equal(got["part"], g.part) and np.array_equal(got["local"], g.local)
    assert got["n_edges"] == len(g.edges)
    for name in ("XA", "XB", "XC", "BA", "CA"):
        rp, ci, vals, nr, nc = got[name]
        orp, oci, onr, onc = g.blocks[name]
        assert (nr, nc) == (onr, onc) and vals is None
        assert np.array_equal(rp, orp), name
        assert np.array_equal(ci, oci), name


def test_generator_independent_of_thread_count(mmsb):
    case = dict(n_x=500, n_a=500, n_b=500, n_c=500, k=20, alpha0=1.0, p_in=0.05, p_out=0.001, seed=5,
                degree_corrected=True, mode="balls")
    a = mmsb.generate(**case, threads=1)
    b = mmsb.generate(**case, threads=7)
    for name in ("XA", "XB", "XC", "BA", "CA"):
        assert np.array_equal(a[name][0], b[name][0]) and np.array_equal(a[name][1], b[name][1])


def test_generator_large_graph_matches_oracle_at_any_thread_count(mmsb):
    """1.65M edges: above the 2^20 keys where the edge list is sorted and deduplicated in parallel
    buckets. 7 threads split the keys into chunks whose bounds n c / 7 are not integers, which once
    made two chunks share one histogram (edges were lost, and the count varied run to run)."""
    case = dict(n_x=12000, n_a=12000, n_b=12000, n_c=12000, k=30, alpha0=1.5, p_in=0.02, p_out=0.0008, seed=4,
                degree_corrected=True, mode="balls")
    a = mmsb.generate(**case, threads=1)
    b = mmsb.generate(**case, threads=7)
    g = gen.generate(gen.GenParams(**case))
    assert a["n_edges"] == b["n_edges"] == len(g.edges) > (1 << 20)
    for name in ("XA", "XB", "XC", "BA", "CA"):
        orp, oci, _, _ = g.blocks[name]
        for got in (a, b):
            assert np.array_equal(got[name][0], orp) and np.array_equal(got[name][1], oci), name


def test_generator_rejects_bad_params(mmsb):
    with pytest.raises(mmsb.MMSBError):
        mmsb.generate(0, 10, 10, 10, 3, 1.0, 0.1, 0.01)
