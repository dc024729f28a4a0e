"""The library's host-side I/O (csrc/io.cpp through the C-ABI) against the oracle's restatement of
the reference formats, bit for bit: the reader on the shipped excerpt of DATA/Gx_a.txt in both
orientations, under every duplicate policy, with and without the while(!feof) repeat, binary and
weighted; the writers byte for byte. No GPU is used."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import core, io

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EXCERPT = os.path.join(GOLD, "Gx_a_head3000.txt")
REF_DATA = "/root/reference/CommunityDetectionSoftware_CPU/data"


def _same(got, want, weighted_vals=False):
    rp, ci, v, nr, nc = got
    rp2, ci2, v2, nr2, nc2 = want
    assert (nr, nc) == (nr2, nc2)
    assert np.array_equal(rp, np.asarray(rp2, np.int32))
    assert np.array_equal(ci, np.asarray(ci2, np.int32))
    if v2 is None:
        assert v is None
    else:
        # the library stores values in FP32: the oracle's FP64 sums, rounded once
        assert v is not None and np.array_equal(v, np.asarray(v2, np.float64).astype(np.float32))


@pytest.fixture(scope="module")
def excerpt_dims():
    rows = np.loadtxt(EXCERPT, dtype=np.int64)
    return int(rows[:, 0].max()), int(rows[:, 1].max())


@pytest.mark.parametrize("first_is_row", [True, False])
@pytest.mark.parametrize("dup", ["one", "last", "sum"])
@pytest.mark.parametrize("feof", [False, True])
def test_reader_matches_oracle_on_shipped_excerpt(mmsb, excerpt_dims, first_is_row, dup, feof):
    a, b = excerpt_dims
    nr, nc = (a, b) if first_is_row else (b, a)
    got = mmsb.read_coo(EXCERPT, nr, nc, first_is_row=first_is_row, duplicates=dup, feof_repeat=feof)
    want = io.read_coo(EXCERPT, nr, nc, first_is_row=first_is_row, duplicates=dup, feof_repeat=feof)
    _same(got, want)
    assert got[1].size == 3000        # the excerpt holds no repeated pair
    if dup == "sum" and feof:
        # read_G_sparse adds 1.0 per line: the repeated last line weighs 2 (IOfn.cpp:36-51)
        last = np.loadtxt(EXCERPT, dtype=np.int64)[-1] - 1
        r, c = (last[0], last[1]) if first_is_row else (last[1], last[0])
        rp, ci, v = got[0], got[1], got[2]
        e = rp[r] + np.searchsorted(ci[rp[r]:rp[r + 1]], c)
        assert ci[e] == c and v[e] == 2.0
        assert np.count_nonzero(v != 1.0) == 1


@pytest.fixture(scope="module")
def weighted_file(tmp_path_factory, excerpt_dims):
    """The excerpt with a third column of weights, some pairs repeated with other weights, one pair
    whose weights cancel (pruned by the summing reader) and a last line that repeats an earlier pair."""
    rows = np.loadtxt(EXCERPT, dtype=np.int64)[:600]
    rng = np.random.default_rng(0)
    w = np.round(rng.uniform(-2, 3, rows.shape[0]), 3)
    extra_idx = rng.choice(rows.shape[0], 40, replace=False)
    extra = rows[extra_idx]
    extra_w = np.round(rng.uniform(-1, 1, 40), 4)
    extra_w[0] = -w[extra_idx[0]]               # this pair sums to zero
    lines = [f"{r}\t{c}\t{v:.6g}" for (r, c), v in zip(rows, w)]
    lines += [f"{r} {c} {v:.6g}" for (r, c), v in zip(extra, extra_w)]
    lines.append(f"{rows[5, 0]}\t{rows[5, 1]}\t0.125")
    p = tmp_path_factory.mktemp("w") / "Gw.txt"
    p.write_text("\n".join(lines) + "\n")
    return str(p)


@pytest.mark.parametrize("first_is_row", [True, False])
@pytest.mark.parametrize("dup", ["one", "last", "sum"])
@pytest.mark.parametrize("feof", [False, True])
def test_weighted_reader_matches_oracle(mmsb, excerpt_dims, weighted_file, first_is_row, dup, feof):
    a, b = excerpt_dims
    nr, nc = (a, b) if first_is_row else (b, a)
    got = mmsb.read_coo(weighted_file, nr, nc, first_is_row=first_is_row, weighted=True, duplicates=dup,
                        feof_repeat=feof)
    want = io.read_coo(weighted_file, nr, nc, first_is_row=first_is_row, weighted=True, duplicates=dup,
                       feof_repeat=feof)
    _same(got, want)
    if dup == "sum":
        assert got[1].size == 599          # 600 distinct pairs, one cancelled and pruned
    else:
        assert got[1].size == 600


def test_reader_errors_are_status_codes(mmsb, tmp_path):
    with pytest.raises(mmsb.MMSBError) as e:
        mmsb.read_coo(str(tmp_path / "missing.txt"), 3, 3)
    assert e.value.code == mmsb.MMSB_ERR_IO
    f = tmp_path / "oob.txt"
    f.write_text("1 2\n4 1\n")
    with pytest.raises(mmsb.MMSBError) as e:
        mmsb.read_coo(str(f), 3, 3)
    assert e.value.code == mmsb.MMSB_ERR_IO and "outside" in str(e.value)
    with pytest.raises(ValueError):
        io.read_coo(str(f), 3, 3)
    g = tmp_path / "empty.txt"
    g.write_text("")
    rp, ci, v, nr, nc = mmsb.read_coo(str(g), 4, 5, feof_repeat=True)
    assert rp.tolist() == [0] * 5 and ci.size == 0 and (nr, nc) == (4, 5)


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="reference data not present")
def test_reader_full_data_matches_digests(mmsb):
    """The library reader on the five shipped partition files: nnz and SHA-256 of the CSR arrays
    equal the committed digests (which the oracle reader also reproduces)."""
    dig = json.load(open(os.path.join(GOLD, "data_digests.json")))
    for name, d in dig.items():
        rp, ci, v, _, _ = mmsb.read_coo(os.path.join(REF_DATA, name + ".txt"), d["rows"], d["cols"])
        assert v is None and ci.size == d["nnz"] == d["lines"]
        assert hashlib.sha256(rp.astype(np.int32).tobytes()).hexdigest() == d["rowptr_sha256"], name
        assert hashlib.sha256(ci.astype(np.int32).tobytes()).hexdigest() == d["colidx_sha256"], name


def test_value_writer_is_byte_identical(mmsb, tmp_path):
    rng = np.random.default_rng(1)
    v = np.r_[rng.standard_normal(50) * 10.0 ** rng.integers(-30, 30, 50), 0.0, -0.0, 1.0, 1 / 3, 1e-300, 2.5e300]
    path = tmp_path / "vals.txt"
    mmsb.check(mmsb.lib.mmsb_write_values(str(path).encode(), np.ascontiguousarray(v).ctypes.data, v.size))
    assert path.read_text() == io.format_mat(v)


def test_pi_triplet_writer_is_byte_identical(mmsb, tmp_path):
    rng = np.random.default_rng(2)
    k, n = 7, 40
    pi = rng.uniform(0, 1, (n, k)) * (rng.random((n, k)) < 0.4)
    pi[3] = 0.0
    pi[5, 2] = 1e-7
    pi[6, 1] = 123456.789
    path = tmp_path / "pi.txt"
    mmsb.check(mmsb.lib.mmsb_write_pi_triplets(str(path).encode(), np.ascontiguousarray(pi).ctypes.data, n, k))
    assert path.read_text() == io.format_pi_triplets(pi.T)
