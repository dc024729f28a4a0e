"""On-disk formats of the reference, restated for the oracle (TEST INFRASTRUCTURE).

Readers:
  read_G (GPU, pre_proc.c:49-80): whitespace-separated numbers parsed with %lf; per entry the
    FIRST number is the X-side (column of the stored NA x NX buffer) and the SECOND the partner
    index, 1-based; BINARY sets the entry to 1 (pre_proc.c:67-68), EXPECTED reads a third number
    into it (pre_proc.c:64-65), so a repeated entry keeps the last value.
  read_G_sparse (CPU, IOfn.cpp:21-54): first number = row, second = column, 1-based; triplets
    are summed by setFromTriplets (IOfn.cpp:51) and |v| <= 1e-18 pruned (IOfn.cpp:52).
  Both loop on `while(!feof)`: after the last newline one more iteration re-uses the previous
  indices, so the last entry is applied twice (pre_proc.c:60-78, IOfn.cpp:36-49). With BINARY /
  last-value semantics that is harmless; with summed triplets the last edge weighs 2.
Writers:
  write_mat (post_proc.c:45-59) and the evals/evecs/frob_norm writers (stpm.c:926-951): one
    "%5.25e" value per line; Pi_* files hold the k x N column-major buffer, i.e. node-major.
  write_pi (IOfn.cpp:57-67): "row+1 \\t col+1 \\t value" per nonzero in column-major order with
    the ostream default (6 significant digits).
  sim_info (stpm.c:957-975): fixed key = value lines.
"""
from __future__ import annotations

import numpy as np

from .gen import coo_to_csr


def parse_numbers(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        data = f.read()
    return np.array(data.split(), dtype=np.float64)


def read_coo(path: str, n_rows: int, n_cols: int, first_is_row: bool = True, weighted: bool = False,
             duplicates: str = "one", feof_repeat: bool = False):
    """Read a 1-based COO text file into CSR (rowptr, colidx, vals or None, n_rows, n_cols).

    duplicates: "one" (BINARY: entry = 1), "last" (EXPECTED on the GPU path: last value wins),
    "sum" (CPU path: setFromTriplets sums, also for two-column files, whose lines count 1 each).
    feof_repeat applies the last entry once more, as the reference's while(!feof) loop does when
    the file ends with a newline.
    """
    nums = parse_numbers(path)
    w = 3 if weighted else 2
    nums = nums[: (nums.size // w) * w].reshape(-1, w)
    if feof_repeat and nums.shape[0]:
        with open(path, "rb") as f:
            f.seek(-1, 2)
            ends_nl = f.read(1) in (b"\n", b" ", b"\t", b"\r")
        if ends_nl:
            nums = np.vstack([nums, nums[-1:]])
    a = nums[:, 0].astype(np.int64) - 1
    b = nums[:, 1].astype(np.int64) - 1
    rows, cols = (a, b) if first_is_row else (b, a)
    if rows.size and (rows.min() < 0 or rows.max() >= n_rows or cols.min() < 0 or cols.max() >= n_cols):
        raise ValueError(f"{path}: index out of range for {n_rows} x {n_cols}")
    if duplicates == "one" or (not weighted and duplicates == "last"):
        rp, ci, nr, nc = coo_to_csr(rows, cols, n_rows, n_cols)
        return rp, ci, None, nr, nc
    # weighted, or summed binary triplets (read_G_sparse pushes 1.0 per line, IOfn.cpp:45-46, so the
    # feof repeat of the last line gives that entry the value 2)
    vals = nums[:, 2] if weighted else np.ones(nums.shape[0])
    key = rows * n_cols + cols
    if duplicates == "last":
        order = np.argsort(key, kind="stable")
        key_s, val_s = key[order], vals[order]
        last = np.r_[key_s[1:] != key_s[:-1], True]
        ukey, uval = key_s[last], val_s[last]
    elif duplicates == "sum":
        ukey, inv = np.unique(key, return_inverse=True)
        uval = np.zeros(ukey.size)
        np.add.at(uval, inv, vals)
    else:
        raise ValueError(duplicates)
    keep = np.abs(uval) > 1e-18
    ukey, uval = ukey[keep], uval[keep]
    r, c = ukey // n_cols, ukey % n_cols
    rowptr = np.zeros(n_rows + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    return np.cumsum(rowptr).astype(np.int32), c.astype(np.int32), uval, n_rows, n_cols


def write_coo(path: str, rowptr, colidx, vals=None):
    """1-based COO text, one 'row<TAB>col' (or 'row<TAB>col<TAB>val') line per nonzero."""
    with open(path, "w") as f:
        for i in range(len(rowptr) - 1):
            for e in range(rowptr[i], rowptr[i + 1]):
                if vals is None:
                    f.write(f"{i + 1}\t{colidx[e] + 1}\n")
                else:
                    f.write(f"{i + 1}\t{colidx[e] + 1}\t{vals[e]:.17g}\n")


def format_mat(values) -> str:
    return "".join("%5.25e\n" % v for v in np.asarray(values, np.float64).ravel())


def format_pi_triplets(pi_kn: np.ndarray) -> str:
    """write_pi of a k x N matrix: nonzeros in column-major order, '%g'-style 6 digits."""
    out = []
    k, n = pi_kn.shape
    for col in range(n):
        for row in range(k):
            v = pi_kn[row, col]
            if v != 0.0:
                out.append("%d\t%d\t%g\n" % (row + 1, col + 1, v))
    return "".join(out)


SIM_INFO_LINES = (
    ("Total number of nodes          = %d", "n_total"),
    ("Number of nodes in partition X = %d", "nx"),
    ("Number of nodes in partition A = %d", "na"),
    ("Number of nodes in partition B = %d", "nb"),
    ("Number of nodes in partition C = %d", "nc"),
    ("Number of nodes communities    = %d", "ncom"),
    ("Alpha_0 (Dirichlet parameter)  = %5.25e", "alpha_0"),
    ("Learning rate = %5.25e", "learn_rate"),
    ("Threshold for STPM convergence    = %5.25e", "thresh"),
    ("Pseudoinverse numerical tolerance = %5.25e", "pinv_tol"),
    ("Total number of Frob Norm values written to file = %d", "n_frob"),
    ("Number of iterations between convergence test    = %d", "iters_between_tests"),
    ("Maximum number of stochastic iterations allowed  = %d", "sgd_iter_max"),
    ("Exec Time svd1  = %5.25e (Seconds)", "t_svd1"),
    ("Exec Time svd2  = %5.25e (Seconds)", "t_svd2"),
    ("Exec Time Pre Proc  = %5.25e (Seconds)", "t_pre"),
    ("Exec Time STPM      = %5.25e (Seconds)", "t_stpm"),
    ("Exec Time Post Proc = %5.25e (Seconds)", "t_post"),
    ("Total Exec Time     = %5.25e (Seconds)", "t_total"),
)


def format_sim_info(v: dict) -> str:
    """The sim_info file, line for line as stpm.c:957-975 prints it (the total time is the sum of
    the three phase times, stpm.c:975)."""
    return "".join((fmt % v[key]) + "\n" for fmt, key in SIM_INFO_LINES)


def parse_sim_info(text: str) -> dict:
    """Inverse of format_sim_info: the value of every line, keyed as SIM_INFO_LINES."""
    lines = text.split("\n")
    out = {}
    for (fmt, key), line in zip(SIM_INFO_LINES, lines):
        val = line.split("=", 1)[1].split("(")[0].strip()
        out[key] = int(val) if fmt.endswith("%d") else float(val)
    return out
