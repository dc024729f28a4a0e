"""Benchmark of the MMSB tensor-method pipeline on one MI355X.

Metric: STGD nodes/s (one step = one STGD epoch over the n_X nodes of partition X, minibatch b),
plus the end-to-end seconds of whitening + STGD + post-processing, on the synthetic workload
"1M nodes per partition, k = 500, alpha0 = 2, degree-corrected (Pareto 2.5), mean degree ~80".

Prints ONE JSON line. Extra objects: "roofline" for the dominant STGD kernel (MFMA-bound),
"spmm" (HBM roofline of the CSR SpMM on the same graph), "cpu_baseline" (the FP64 oracle's STGD
step timed on this host on a bounded sample of the same minibatches).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c1] [--precision bf16|fp32]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"bf16": 2.5e15, "fp32": 157.3e12, "hbm": 8.0e12}

# workloads: n per partition, k, alpha0, degree-corrected, target mean degree (or p_in/p_out), b
WORKLOADS = {
    "c1": dict(n=1000, k=10, alpha0=1.0, dc=False, p_in=0.6, p_out=0.02, b=100, eta=1e-3),
    "c2": dict(n=100000, k=100, alpha0=1.0, dc=False, p_in=0.01, p_out=2e-4, b=1024, eta=1e-3),
    "c3": dict(n=1000000, k=500, alpha0=2.0, dc=True, mean_degree=80.0, b=1024, eta=1e-3),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def c3_probs(n, k, mean_degree):
    # mean degree ~ N (p_out + (p_in - p_out)/k) with p_in = (k+1) p_out: half of the edges
    # fall inside communities.
    N = 4 * n
    p_out = mean_degree / (2.0 * N)
    return (k + 1) * p_out, p_out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="bf16", choices=["fp32", "bf16"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, "
              "'How the path would shard'); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch  # noqa: F401  (one HIP runtime: torch's)
    import __graft_entry__ as ge
    m = ge.load_package()

    w = WORKLOADS[args.workload]
    n, k, b = w["n"], w["k"], w["b"]
    if "mean_degree" in w:
        p_in, p_out = c3_probs(n, k, w["mean_degree"])
    else:
        p_in, p_out = w["p_in"], w["p_out"]
    t0 = time.time()
    g = m.generate(n, n, n, n, k, w["alpha0"], p_in, p_out, seed=0, degree_corrected=w["dc"], threads=16)
    t_gen = time.time() - t0
    nnz = {nm: int(g[nm][1].size) for nm in ("XA", "XB", "XC", "BA", "CA")}
    mean_deg = 2.0 * g["n_edges"] / (4 * n)
    log(f"generated {args.workload}: {g['n_edges']} edges, mean degree {mean_deg:.1f}, {t_gen:.1f}s")

    cfg = m.config(nx=n, na=n, nb=n, nc=n, ncom=k, alpha_0=w["alpha0"], learn_rate=w["eta"], minibatch=b,
                   epochs=1, polish_epochs=0, oversample=10, power_iters=1, seed=1, precision=args.precision)
    # ---- end-to-end: CSR on host -> whitening -> 1 STGD epoch -> eigen -> Pi on host ----------
    t0 = time.time()
    p = m.Pipeline(cfg)
    p.set_graph(g)
    t_up = time.time() - t0
    p.whiten()
    t_wh = time.time() - t0 - t_up
    phi0 = None
    p.stgd(0, "fp32")                      # seeded orthonormal Phi0 ...
    phi0 = p.get("phi")                    # ... kept for the precision comparison below
    t1 = time.time()
    p.stgd(-1, args.precision)
    t_st = time.time() - t1
    t2 = time.time()
    p.eigen()
    p.post()
    t_post = time.time() - t2
    pin = [p.get_f32("pin_" + s) for s in "xabc"]
    t_e2e = time.time() - t0
    t_dl = time.time() - t2 - t_post
    log(f"end-to-end {t_e2e:.2f}s (upload {t_up:.2f}, whiten {t_wh:.2f}, stgd {t_st:.2f}, post {t_post:.2f}, "
        f"download {t_dl:.2f})")
    diag = {"nonfinite_pin": int(sum((~np.isfinite(a)).sum() for a in pin)),
            "nonfinite_phi": int((~np.isfinite(p.get("phi"))).sum()),
            "nonfinite_W_rows0_1000": int((~np.isfinite(p.get_rows("W", 0, 1000))).sum()),
            "lam_m2_top": [float(x) for x in p.get("lam_m2")[:3]],
            "lam_m2_last": [float(x) for x in p.get("lam_m2")[-3:]],
            "frob_last": float(p.get("frob")[-1])}
    log(f"diagnostics {diag}")
    del pin

    # ---- timed STGD epochs in both precisions --------------------------------------------------
    spe = math.ceil(n / b)
    rates = {}
    precs = ("fp32", "bf16")
    for prec in precs:
        for _ in range(args.warmup):
            p.stgd(spe, prec)
        m.check(m.lib.mmsb_sync(p.ctx))
        ts = time.time()
        for _ in range(args.steps):
            p.stgd(spe, prec)
        m.check(m.lib.mmsb_sync(p.ctx))
        dt = time.time() - ts
        rates[prec] = (n * args.steps / dt, 1e3 * dt / args.steps)
        log(f"STGD {prec}: {rates[prec][0]:.3e} nodes/s, {rates[prec][1]:.2f} ms/epoch")
    prec_used = args.precision if args.precision in rates else "fp32"
    nodes_per_s, ms_per_epoch = rates[prec_used]
    flops_epoch = 6.0 * k * k * n + 4.0 * k ** 3 * spe           # algorithmic STGD flops / epoch
    ns_flops_epoch = 2 * 2 * 2.0 * k ** 3 * spe                  # Newton-Schulz renormalisation (2 iters)
    epoch_achieved = flops_epoch / (ms_per_epoch * 1e-3)
    peak = PEAK[prec_used]

    # ---- per-kernel device time: one more epoch launched directly, HIP events around each phase ----
    # on the context stream (the stream every STGD kernel runs on). P1 -- the Y_A Phi, Y_B Phi
    # contractions fused with the C_g epilogue, plus (Phi^T Phi)^o2 -- is the largest kernel.
    ph_ms = p.stgd_timed(spe, prec_used)
    phases_us = {nm: 1e3 * t / spe for nm, t in zip(("p1", "p2", "renorm"), ph_ms)}
    p1_flops = (4.0 * k * k * n + 2.0 * k ** 3 * spe) / spe       # per launch: 2 b k^2 (P_A) + 2 b k^2 (P_B) + 2 k^3
    p1_achieved = p1_flops / (phases_us["p1"] * 1e-6)
    log(f"phases (us/step): {phases_us}; P1 {p1_achieved / 1e12:.1f} TFLOP/s")

    # ---- Pi agreement: 2 epochs FP32 vs 1 epoch BF16 + 1 FP32 polish epoch from the same Phi0 ------
    pis = {}
    for mode in precs:
        p.set_phi(phi0)
        p.stgd(spe, mode)
        p.stgd(spe, "fp32")
        p.eigen()
        p.post()
        pis[mode] = p.get_f32("pin_x")
    pi_err = (float(np.abs(pis["bf16"] - pis["fp32"]).max() / max(np.abs(pis["fp32"]).max(), 1e-30))
              if "bf16" in pis else None)
    del pis
    log(f"Pi_X(bf16+polish) vs Pi_X(fp32): max rel diff {pi_err}")

    # ---- SpMM roofline on G_XA (n x n, nnz) times a tall n x r FP32 matrix ----------------------
    spmm = spmm_roofline(m, g["XA"], r=k + 10)

    out = {
        "metric": f"STGD nodes/s (n={n}/partition, k={k})",
        "value": nodes_per_s,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_epoch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec_used,
        "data": "synthetic (seeded MMSB generator)",
        "config": {"workload": f"mmsb-{args.workload}: {n} nodes/partition, k={k}, alpha0={w['alpha0']}, "
                               f"{'degree-corrected Pareto(2.5), ' if w['dc'] else ''}mean degree {mean_deg:.1f}, "
                               f"minibatch {b}, one step = one STGD epoch over X",
                   "n_per_partition": n, "k": k, "minibatch": b, "nnz": nnz, "parallelism": "none"},
        "end_to_end_s": t_e2e,
        "stages_s": {"upload_and_transpose": t_up, "whiten": t_wh, "stgd_1_epoch": t_st, "eigen_post": t_post,
                     "download_pi": t_dl, "generate": t_gen},
        "diagnostics": diag,
        "stgd_nodes_per_s": {pr: v[0] for pr, v in rates.items()},
        "bf16_speedup_vs_fp32": rates["bf16"][0] / rates["fp32"][0] if "bf16" in rates else None,
        "pi_rel_diff_bf16_polish_vs_fp32": pi_err,
        "roofline": {"bound": "mfma", "achieved": p1_achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": p1_achieved / peak, "traffic": None,
                     "kernel": f"p1_kernel ({prec_used}): Y_A Phi, Y_B Phi + C_g epilogue, (Phi^T Phi)^o2",
                     "flops_per_launch": p1_flops, "avg_us": phases_us["p1"], "phases_us_per_step": phases_us,
                     "epoch_tflops": epoch_achieved / 1e12, "epoch_frac": epoch_achieved / peak,
                     "flops_per_epoch": flops_epoch, "renorm_flops_per_epoch": ns_flops_epoch},
        "spmm": spmm,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, n, k, b, w, args.cpu_seconds)
    p.close()
    print(json.dumps(out), flush=True)


def spmm_roofline(m, blk, r):
    import torch
    rowptr, colidx, _, nr, nc = blk
    dev = torch.device("cuda:0")
    ld = (r + 3) // 4 * 4
    d_rp = torch.from_numpy(rowptr).to(dev)
    d_ci = torch.from_numpy(colidx).to(dev)
    B = torch.randn(nc, ld, device=dev, dtype=torch.float32)
    Y = torch.empty(nr, ld, device=dev, dtype=torch.float32)
    stream = torch.cuda.current_stream()
    for _ in range(2):
        m.spmm_f32(d_rp, d_ci, None, B, Y, r=r, n=nc)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    ev0.record(stream)
    for _ in range(reps):
        m.spmm_f32(d_rp, d_ci, None, B, Y, r=r, n=nc)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    nnz = colidx.size
    bytes_alg = 4 * (nr + 1) + 4 * nnz + 4 * r * nnz + 4 * r * nr   # binary CSR: no value array
    gbs = bytes_alg / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s", "frac": gbs * 1e9 / PEAK["hbm"],
            "ms": ms, "rows": nr, "nnz": int(nnz), "r": r, "bytes_per_launch": bytes_alg}


def cpu_baseline(p, n, k, b, w, seconds):
    """The oracle's FP64 STGD step (numpy) on the first minibatches of the same whitened data."""
    import numpy as np
    from oracle import core
    nsample = min(n, 64 * b)
    ya, yb, yc = (p.get_rows(nm, 0, nsample) for nm in ("YA", "YB", "YC"))
    mus = [p.get(nm) for nm in ("muw_a", "muw_b", "muw_c")]
    phi = p.get("phi")
    steps = 0
    nbatch = max(1, nsample // b)
    t0 = time.time()
    while time.time() - t0 < seconds:
        lo = (steps % nbatch) * b
        phi = core.stgd_step(phi, ya[lo:lo + b], yb[lo:lo + b], yc[lo:lo + b], *mus, w["alpha0"], w["eta"])
        steps += 1
    dt = time.time() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": steps * b / dt, "unit": "nodes/s", "cores": threads, "kind": "port",
            "sample": f"{steps} STGD minibatch steps (b={b}, k={k}, cycling over the first {nsample} whitened nodes) "
                      f"of the oracle's FP64 numpy step, {dt:.1f}s"}


if __name__ == "__main__":
    main()
