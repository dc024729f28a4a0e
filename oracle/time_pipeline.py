"""Times the FP64 CPU restatements end to end, phase by phase, on a configuration's seeded graph --
TEST INFRASTRUCTURE ONLY (the CPU side of the end-to-end comparison; nothing in the library runs it).

    python -m oracle.time_pipeline --config c2 --threads 1

Phases: generation (reported apart, as the reference reads text files instead), whitening, the
whitened neighbour vectors, STGD (numpy; and the C / OpenMP restatement of the same epochs when it is
built), eigenvalues + Pi of the four partitions + normProbVector. BLAS and OpenMP threads are capped
at --threads. Prints one JSON line.
"""
import argparse
import json
import os
import time

import numpy as np

from . import core, cstgd, gen

CONFIGS = {  # n per partition, k, alpha0, p_in, p_out, minibatch, learning rate, epochs
    "c1": dict(n=1000, k=10, alpha0=1.0, p_in=0.6, p_out=0.02, b=100, eta=1e-3, epochs=200),
    "c2": dict(n=100000, k=100, alpha0=1.0, p_in=0.01, p_out=2e-4, b=1024, eta=1e-3, epochs=1),
}


def run(name: str, threads: int) -> dict:
    c = CONFIGS[name]
    n, k = c["n"], c["k"]
    t = {}
    t0 = time.time()
    g = gen.generate(gen.GenParams(n, n, n, n, k, c["alpha0"], c["p_in"], c["p_out"], seed=0))
    t["generate"] = time.time() - t0
    bl = {nm: core.Csr(rp, ci, nr, nc) for nm, (rp, ci, nr, nc) in g.blocks.items()}
    nnz = int(bl["XA"].colidx.size)
    del g
    wcfg = core.WhitenConfig(k=k, alpha0=c["alpha0"], oversample=10, power_iters=1, seed=1)
    t0 = time.time()
    wh = core.whiten(bl["XA"], bl["XB"], bl["XC"], wcfg)
    t["whiten"] = time.time() - t0
    t0 = time.time()
    ys = core.yvectors(bl["XA"], bl["XB"], bl["XC"], wh)
    t["yvectors"] = time.time() - t0
    args = (ys["Y_A"], ys["Y_B"], ys["Y_C"], ys["mu_A"], ys["mu_B"], ys["mu_C"], core.init_phi(1, k))
    kw = dict(alpha0=c["alpha0"], eta=c["eta"], batch=c["b"], epochs=c["epochs"])
    t0 = time.time()
    phi, _ = core.stgd(*args, **kw)
    t["stgd_numpy"] = time.time() - t0
    out = {"config": name, "n_per_partition": n, "k": k, "nnz_xa": nnz, "threads": threads,
           "stgd_epochs": c["epochs"], "minibatch": c["b"]}
    if cstgd.available():
        t0 = time.time()
        phi_c, _, phases = cstgd.stgd(*args, threads=threads, **kw)
        t["stgd_c_openmp"] = time.time() - t0
        out["stgd_c_phases_s"] = {key: round(v, 4) for key, v in phases.items()}
        out["stgd_c_vs_numpy_max_abs"] = float(np.abs(phi_c - phi).max())
    t0 = time.time()
    _, lam_p, _ = core.eigenvalues(phi)
    pis = core.post_proc(wh, phi, lam_p, bl["XA"], bl["BA"], bl["CA"])
    _ = [core.norm_prob_rows(P) for P in pis]
    t["eigen_post"] = time.time() - t0
    stgd_s = min(v for key, v in t.items() if key.startswith("stgd_"))
    out["phases_s"] = {key: round(v, 4) for key, v in t.items()}
    out["end_to_end_s"] = round(t["whiten"] + t["yvectors"] + stgd_s + t["eigen_post"], 4)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", choices=sorted(CONFIGS), default="c1")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=a.threads):
        print(json.dumps(run(a.config, a.threads)), flush=True)


if __name__ == "__main__":
    main()
