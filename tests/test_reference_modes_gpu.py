"""The reference-faithful modes through the C-ABI, against the FP64 oracle: the per-step QR
renormalisation (stpm.c:774-786), the per-step end test of the STGD loop (stpm.c:747-765), the
tensor coefficients that go with each M2 scaling, the program's output files byte for byte
(sim_info, stpm.c:957-975; the CPU program's triplet files, CommunityDetection.cpp:121-144), the
SpMM paths without column strips, and the release of the process-wide resources."""
import os

import numpy as np
import pytest

from oracle import core, csr_spmm, io as oio

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _inputs(n, k, seed, mu=0.2):
    rng = np.random.default_rng(seed)
    ys = [rng.standard_normal((n, k)).astype(np.float32).astype(np.float64) for _ in range(3)]
    mus = [mu * rng.standard_normal(k) for _ in range(3)]
    return ys, mus


def _pipe(mmsb, n, k, b, ys, mus, phi0, **kw):
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, oversample=0, minibatch=b, **kw)
    p = mmsb.Pipeline(cfg)
    p.set_yvectors(*ys, *mus)
    p.set_phi(phi0)
    return p


# b = 1: the reference's one-node step; graphs on and off; BF16 contractions with the FP32 QR
@pytest.mark.parametrize("k,b,prec,graphs", [(6, 1, "fp32", 1), (10, 64, "fp32", 1), (24, 100, "fp32", 0),
                                             (16, 64, "bf16", 1)])
def test_qr_renormalisation_matches_oracle(mmsb, k, b, prec, graphs):
    n = 640 if b > 1 else 80
    ys, mus = _inputs(n, k, seed=k + b)
    phi0 = core.init_phi(3, k)
    print(f"case k={k} b={b} prec={prec} graphs={graphs}", flush=True)
    p = _pipe(mmsb, n, k, b, ys, mus, phi0, alpha_0=1.2, learn_rate=3e-3, renorm_mode="qr", use_graphs=graphs,
              m2_scaling="gpu")
    spe = -(-n // b)
    p.stgd(spe + 3, prec)           # a whole epoch (captured when graphs are on) and three more steps
    phi_ref, log_ref = core.stgd(*ys, *mus, phi0, alpha0=1.2, eta=3e-3, batch=b, epochs=1, steps=spe + 3,
                                 scaling="gpu", renorm_mode="qr")
    phi = p.get("phi")
    tol = 1e-4 if prec == "fp32" else 1e-2
    assert rel_err(phi, phi_ref) < tol
    assert np.abs(phi.T @ phi - np.eye(k)).max() < 1e-5          # Q is orthonormal in every mode
    # R = Q^T Phi~ of the last step is upper triangular with a positive diagonal: Phi is that Q
    np.testing.assert_allclose(p.get("frob"), log_ref, rtol=(1e-3 if prec == "fp32" else 5e-2), atol=1e-10)
    p.close()


@pytest.mark.parametrize("graphs", [1, 0])
def test_thresh_ends_the_loop_mid_epoch(mmsb, graphs):
    """THRESH above every logged value with SGD_ITER_MIN = 5: the loop ends after step index 6
    (7 steps), inside the first of three epochs, also when the epoch runs as a captured graph; Phi is
    that of 7 uninterrupted steps."""
    n, k, b = 1024, 12, 64
    ys, mus = _inputs(n, k, seed=3)
    phi0 = core.init_phi(4, k)
    p = _pipe(mmsb, n, k, b, ys, mus, phi0, learn_rate=2e-3, epochs=3, thresh=1e9, sgd_iter_min=5, use_graphs=graphs)
    p.stgd(-1, "fp32")
    log = p.get("frob")
    assert log.size == 7, log.size
    phi_ref, log_ref = core.stgd(*ys, *mus, phi0, alpha0=1.0, eta=2e-3, batch=b, epochs=3, thresh=1e9, iter_min=5)
    assert log_ref.size == 7
    assert rel_err(p.get("phi"), phi_ref) < 1e-4
    q = _pipe(mmsb, n, k, b, ys, mus, phi0, learn_rate=2e-3, use_graphs=0)
    q.stgd(7, "fp32")
    assert np.array_equal(p.get("phi"), q.get("phi"))
    p.close()
    q.close()


def test_iter_max_caps_the_loop(mmsb):
    """SGD_ITER_MAX = 20 with a convergence test that never fires (THRESH = 0): the last step is
    index 21, so 22 steps run -- the captured first epoch (16 steps) and 6 steps of the second, whose
    remaining kernels return at once -- and Phi equals that of 22 uninterrupted BF16 steps (the FP32
    polishing epoch, the last of four, is never reached)."""
    n, k, b = 1024, 8, 64
    ys, mus = _inputs(n, k, seed=5)
    phi0 = core.init_phi(5, k)
    p = _pipe(mmsb, n, k, b, ys, mus, phi0, learn_rate=2e-3, epochs=4, polish_epochs=1, thresh=0.0, sgd_iter_min=1,
              sgd_iter_max=20)
    p.stgd(-1, "bf16")
    assert p.get("frob").size == 22
    q = _pipe(mmsb, n, k, b, ys, mus, phi0, learn_rate=2e-3, use_graphs=0)
    q.stgd(22, "bf16")
    assert np.array_equal(p.get("phi"), q.get("phi"))
    np.testing.assert_allclose(p.get("frob"), q.get("frob"), rtol=1e-12)
    p.close()
    q.close()


def test_reference_defaults_run_nx_plus_four_single_node_steps(mmsb):
    """The reference's own settings: one node per step, THRESH = 1, SGD_ITER_MIN = NX,
    SGD_ITER_MAX = NX + 2 (stpm.h:94-97), QR renormalisation. With a frob that stays above THRESH the
    loop runs NX + 4 steps -- one epoch and four nodes of the next -- exactly as the oracle's loop."""
    n, k = 60, 5
    ys, mus = _inputs(n, k, seed=9)
    phi0 = core.init_phi(6, k)
    p = _pipe(mmsb, n, k, 1, ys, mus, phi0, learn_rate=1e-2, epochs=5, polish_epochs=0, thresh=1e-30, sgd_iter_min=n,
              sgd_iter_max=n + 2, renorm_mode="qr", m2_scaling="gpu")
    p.stgd(-1, "fp32")
    log = p.get("frob")
    phi_ref, log_ref = core.stgd(*ys, *mus, phi0, alpha0=1.0, eta=1e-2, batch=1, epochs=5, scaling="gpu",
                                 renorm_mode="qr", thresh=1e-30, iter_min=n, iter_max=n + 2)
    assert log_ref.size == n + 4 and log.size == n + 4
    assert rel_err(p.get("phi"), phi_ref) < 1e-4
    np.testing.assert_allclose(log, log_ref, rtol=1e-3)
    p.close()


def test_gpu_m2_scaling_gives_the_reference_gpu_step(mmsb):
    """m2_scaling = GPU whitens with the GPU M2 (pre_proc.c:292-302) and steps with init_buff's
    coefficients (stpm.c:1026-1038): one-node steps at LEARN_RATE on the library's own vectors equal
    the oracle's GPU-scaled steps; the paper scaling takes (a0+1)(a0+2)/2 times the T step."""
    from tests.test_pipeline_gpu import make_case
    g, cfg, bl, wcfg = make_case(mmsb, n=400, k=5, alpha0=2.0, m2_scaling="gpu", minibatch=1, learn_rate=1e-3)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    p.whiten()
    ref = core.whiten(bl["XA"], bl["XB"], bl["XC"], wcfg)
    assert rel_err(p.get("W"), ref["W"]) < 1e-4
    ya, yb, yc = (p.get(nm) for nm in ("YA", "YB", "YC"))
    mus = [p.get(nm) for nm in ("muw_a", "muw_b", "muw_c")]
    phi0 = core.init_phi(2, 5)
    p.set_phi(phi0)
    p.stgd(30, "fp32")
    phi_gpu, _ = core.stgd(ya, yb, yc, *mus, phi0, alpha0=2.0, eta=1e-3, batch=1, epochs=1, steps=30, scaling="gpu")
    phi_pap, _ = core.stgd(ya, yb, yc, *mus, phi0, alpha0=2.0, eta=1e-3, batch=1, epochs=1, steps=30,
                           scaling="paper")
    d_got, d_gpu, d_pap = p.get("phi") - phi0, phi_gpu - phi0, phi_pap - phi0   # what the 30 steps moved
    assert rel_err(d_got, d_gpu) < 1e-3
    assert rel_err(d_got, d_pap) > 0.3
    p.close()


def test_output_files_match_the_reference_formats(mmsb, tmp_path):
    """mmsb_write_outputs: sim_info line for line as stpm.c:957-975 (values parsed back and
    re-printed with the reference's format strings give the same bytes); the CPU program's pi_ and
    white_ triplet files (CommunityDetection.cpp:121-144) equal write_pi of the normalised Pi of all
    partitions concatenated and of W."""
    n, k = 300, 4
    g = mmsb.generate(n, n, n, n, k, 1.0, 0.5, 0.03, seed=7)
    cfg = mmsb.config(nx=n, na=n, nb=n, nc=n, ncom=k, folder_name=str(tmp_path) + "/", minibatch=64, epochs=2,
                      alpha_0_str="1.0", learn_rate_str="1e-3", sgd_iter_max=1000)
    p = mmsb.Pipeline(cfg)
    p.set_graph(g)
    p.whiten()
    p.stgd(-1)
    p.eigen()
    p.post()
    p.write_outputs()
    text = (tmp_path / f"sim_info_{k}_1.0_1e-3.txt").read_text()
    vals = oio.parse_sim_info(text)
    assert oio.format_sim_info(vals) == text
    assert (vals["nx"], vals["ncom"], vals["n_frob"], vals["sgd_iter_max"]) == (n, k, p.get("frob").size, 1000)
    assert vals["iters_between_tests"] == 1 and vals["t_svd1"] > 0 and vals["t_svd2"] > 0
    assert vals["t_total"] == pytest.approx(vals["t_pre"] + vals["t_stpm"] + vals["t_post"], rel=1e-12)
    pin = np.concatenate([p.get(nm) for nm in ("pin_x", "pin_a", "pin_b", "pin_c")])
    pin32 = pin.astype(np.float32).astype(np.float64)
    assert (tmp_path / f"pi_{k}_1.0_1e-3.txt").read_text() == oio.format_pi_triplets(pin32.T)
    W = p.get("W")
    assert (tmp_path / f"white_{k}_1.0_1e-3.txt").read_text() == oio.format_pi_triplets(W)
    # the normalised rows are normProbVector of the raw Pi (the oracle's closed form)
    raw = np.concatenate([p.get(nm) for nm in ("pi_x", "pi_a", "pi_b", "pi_c")])
    np.testing.assert_allclose(pin, core.norm_prob_rows(raw), atol=1e-6)
    # explicit paths, one file skipped
    p.write_triplets(pi_path=tmp_path / "only_pi.txt")
    assert (tmp_path / "only_pi.txt").read_text() == oio.format_pi_triplets(pin32.T)
    p.close()


@pytest.mark.parametrize("r", [300, 510, 1010, 1100, 2100])
def test_spmm_without_strips_matches_oracle(mmsb, dev, r):
    """A strip limit of 0 bytes sends every row wider than 256 columns to the one-wave-per-row kernels
    (2, 4, 8 or 16 float4 chunks per lane) that only very large gathered matrices take by default."""
    import torch
    rng = np.random.default_rng(r)
    m, n = 300, 240
    dense = rng.random((m, n)) < 0.07
    rowptr = np.r_[0, np.cumsum(dense.sum(1))].astype(np.int32)
    colidx = np.nonzero(dense)[1].astype(np.int32)
    ld = (r + 3) // 4 * 4
    B = rng.standard_normal((n, ld)).astype(np.float32)
    t = lambda a: torch.from_numpy(a).to(dev)
    ref = csr_spmm(rowptr, colidx, None, B[:, :r].astype(np.float64), m)
    outs = []
    prev = mmsb.lib.mmsb_set_strip_limit(0)
    try:
        for limit in (0, -1):
            mmsb.lib.mmsb_set_strip_limit(limit)
            Y = torch.zeros((m, ld), dtype=torch.float32, device=dev)
            mmsb.spmm_f32(t(rowptr), t(colidx), None, t(B), Y, r=r, n=n)
            torch.cuda.synchronize()
            outs.append(Y[:, :r].cpu().numpy())
    finally:
        mmsb.lib.mmsb_set_strip_limit(prev)
    np.testing.assert_allclose(outs[0], ref, rtol=1e-4, atol=1e-4)
    assert np.array_equal(outs[0], outs[1])         # the same sums in the same order either way


def test_recomputed_rows_without_strips_match_oracle(mmsb):
    """The checkpointed minibatch fill (yrows_kernel) without strips, k = 130 (two float4 chunks per
    lane): the rows it rebuilds equal G_XA W from the oracle's row loop."""
    from tests.test_pipeline_gpu import make_case
    prev = mmsb.lib.mmsb_set_strip_limit(0)
    try:
        n, k = 1400, 130
        g, cfg, bl, wcfg = make_case(mmsb, n=n, k=k, p_in=0.6, p_out=0.02, power_iters=1, minibatch=128,
                                     y_mode="recompute")
        p = mmsb.Pipeline(cfg)
        p.set_graph(g)
        p.whiten()
        W = p.get("W")
        rows = p.get_rows("YA", 100, 400)
        np.testing.assert_allclose(rows, bl["XA"].mul(W)[100:500], rtol=1e-4, atol=1e-4 * np.abs(rows).max())
        p.close()
    finally:
        mmsb.lib.mmsb_set_strip_limit(prev)


def test_release_and_reuse(mmsb, dev):
    """mmsb_release frees the process-wide staging and temporaries; the next calls re-create them and
    give the same results."""
    import torch
    rng = np.random.default_rng(4)
    m, n, r = 2000, 1500, 64
    dense = rng.random((m, n)) < 0.01
    rowptr = np.r_[0, np.cumsum(dense.sum(1))].astype(np.int32)
    colidx = np.nonzero(dense)[1].astype(np.int32)
    B = torch.from_numpy(rng.standard_normal((n, r)).astype(np.float32)).to(dev)
    outs = []
    for _ in range(2):
        Y = torch.zeros((m, r), dtype=torch.float32, device=dev)
        mmsb.spmm_f32_host_csr(rowptr, colidx, None, B, Y, r=r, n=n, chunk_nnz=5000)
        torch.cuda.synchronize()
        outs.append(Y.cpu().numpy())
        mmsb.lib.mmsb_release()
    assert np.array_equal(outs[0], outs[1])
    g = mmsb.generate(400, 400, 400, 400, 4, 1.0, 0.5, 0.03, seed=2)
    ws = []
    for _ in range(2):
        p = mmsb.Pipeline(mmsb.config(nx=400, na=400, nb=400, nc=400, ncom=4, minibatch=64))
        p.set_graph(g)
        p.whiten()
        ws.append(p.get("W"))
        p.close()
        mmsb.lib.mmsb_release()
    assert np.array_equal(ws[0], ws[1])
