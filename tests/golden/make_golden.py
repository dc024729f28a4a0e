"""Writes the committed fixtures under tests/golden/ from the reference's data files.

Run in the build container only (the reference tree is not present where GPU tests run):
    python tests/golden/make_golden.py /root/reference/CommunityDetectionSoftware_CPU/data
Outputs (data only, no reference source):
  Gx_a_head3000.txt      first 3000 lines of data/Gx_a.txt (reader fixture)
  INVLAMPHI_WRITE.txt    the 10 x 10 Lambda^-1 Phi^T triplets the CPU path wrote (k = 10)
  Partition.txt          partition sizes
  data_digests.json      per input file: line count, nnz and SHA-256 of the CSR arrays built by
                         the oracle reader in the reference's CPU orientation (row = first column)
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.io import read_coo  # noqa: E402


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(data):
    with open(os.path.join(data, "Gx_a.txt")) as f, open(os.path.join(HERE, "Gx_a_head3000.txt"), "w") as g:
        for i, line in enumerate(f):
            if i == 3000:
                break
            g.write(line)
    for name in ("INVLAMPHI_WRITE.txt", "Partition.txt"):
        shutil.copy(os.path.join(data, name), os.path.join(HERE, name))
    nx, na, nb, nc = 65879, 65879, 65879, 65880          # data/Partition.txt
    shapes = {"Gx_a": (nx, na), "Gx_b": (nx, nb), "Gx_c": (nx, nc), "Gb_a": (nb, na), "Gc_a": (nc, na)}
    out = {}
    for name, (r, c) in shapes.items():
        path = os.path.join(data, name + ".txt")
        lines = sum(1 for _ in open(path))
        rp, ci, _, _, _ = read_coo(path, r, c, first_is_row=True)
        out[name] = dict(lines=lines, rows=r, cols=c, nnz=int(ci.size), rowptr_sha256=digest(rp.astype(np.int32)),
                         colidx_sha256=digest(ci.astype(np.int32)), file_sha256=hashlib.sha256(open(path, "rb").read()).hexdigest())
    with open(os.path.join(HERE, "data_digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps({k: (v["lines"], v["nnz"]) for k, v in out.items()}))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/CommunityDetectionSoftware_CPU/data")
