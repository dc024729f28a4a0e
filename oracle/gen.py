"""Seeded synthetic MMSB graph generator of the oracle (TEST INFRASTRUCTURE).

The reference ships no generator (its inputs are text files, stpm.h:63-69); the synthetic
workloads are defined here and reproduced bit-for-bit by csrc/gen.cpp. Model:

  * theta_i: 1, or Pareto(shape a, x_m = 1) normalised to mean 1 (degree-corrected MMSB);
  * pi_i ~ Dirichlet(alpha0/k * 1) via independent Gamma draws g_ic / sum_c g_ic, where
    Gamma(alpha) = Gamma(1 + alpha) * U^(1/alpha) (Marsaglia & Tsang, ACM TOMS 26(3), 2000,
    with the boost for shape < 1); a component whose factor U^(1/alpha) is below e^CUT (e^-20)
    is set to 0 (its mass is < 2e-9 of the node's Gamma scale);
  * edge {i,j}, i != j, with probability p_ij = theta_i theta_j (p_out + (p_in - p_out) pi_i.pi_j):
      - "exact": one Bernoulli draw per pair (N <= 5e4);
      - "balls": background balls (i, j ~ theta) and per-community balls (i, j ~ theta*pi_c),
        ball counts fixed at their means round(p_out W^2/2), round((p_in-p_out) S_c^2/2),
        duplicates collapsed and self-pairs dropped (first-order in p_ij; used above N = 5e4);
  * split: nodes ordered by a Philox key, contiguous quarters become X, A, B, C;
  * blocks G_XA, G_XB, G_XC (rows X), G_BA (rows B), G_CA (rows C), columns sorted, binary.

Every random draw is Philox4x32-10 at a counter built from its indices (oracle/philox.py), every
sum is taken sequentially in index order, and det_log/det_exp replace libm, so the result does
not depend on vectorisation or thread count.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .philox import (STREAM_BALL, STREAM_EDGE, STREAM_GAMMA, STREAM_POW, STREAM_SPLIT, STREAM_THETA,
                     det_exp, det_log, draw, u53, u53_open0)

CUT = -20.0
EXACT_MAX_NODES = 50000


@dataclass
class GenParams:
    n_x: int
    n_a: int
    n_b: int
    n_c: int
    k: int
    alpha0: float
    p_in: float
    p_out: float
    seed: int = 0
    degree_corrected: bool = False
    pareto_shape: float = 2.5
    mode: str = "auto"  # "exact" | "balls" | "auto"

    @property
    def n_total(self) -> int:
        return self.n_x + self.n_a + self.n_b + self.n_c

    def resolved_mode(self) -> str:
        if self.mode != "auto":
            return self.mode
        return "exact" if self.n_total <= EXACT_MAX_NODES else "balls"


@dataclass
class Graph:
    params: GenParams
    theta: np.ndarray          # [N]
    pi: np.ndarray             # [N, k] global node order
    part: np.ndarray           # [N] 0..3 = X, A, B, C
    local: np.ndarray          # [N] index inside the partition
    edges: np.ndarray          # [E, 2] global (i < j), sorted
    blocks: dict = field(default_factory=dict)  # name -> (rowptr, colidx, n_rows, n_cols)

    def pi_true(self, p: int) -> np.ndarray:
        """Ground-truth memberships of partition p in local order ([n_p, k])."""
        idx = np.nonzero(self.part == p)[0]
        out = np.empty((idx.size, self.pi.shape[1]))
        out[self.local[idx]] = self.pi[idx]
        return out


# ---- theta ---------------------------------------------------------------------------------
def gen_theta(p: GenParams) -> np.ndarray:
    n = p.n_total
    if not p.degree_corrected:
        return np.ones(n)
    w = draw(p.seed, STREAM_THETA, np.arange(n, dtype=np.uint32), 0)
    u = u53_open0(w[0], w[1])
    t = det_exp(-det_log(u) / p.pareto_shape)
    total = np.cumsum(t)[-1]           # sequential sum
    mean = total / n
    return t / mean


# ---- Gamma(1 + alpha) by Marsaglia-Tsang ---------------------------------------------------
def gamma1p(seed: int, node: np.ndarray, comp: np.ndarray, alpha: float) -> np.ndarray:
    """Gamma(1+alpha) draws for (node, comp) pairs; counter (node, comp, STREAM_GAMMA, j)."""
    a = 1.0 + alpha
    d = a - 1.0 / 3.0
    cc = 1.0 / np.sqrt(9.0 * d)
    m = node.size
    out = np.zeros(m)
    j = np.zeros(m, np.uint32)
    pend = np.arange(m)
    while pend.size:
        # Marsaglia polar normal: redraw until 0 < s < 1
        x = np.empty(pend.size)
        need = np.arange(pend.size)
        while need.size:
            idx = pend[need]
            w = draw(seed, STREAM_GAMMA, node[idx], comp[idx], j[idx])
            j[idx] += 1
            u1 = 2.0 * u53(w[0], w[1]) - 1.0
            u2 = 2.0 * u53(w[2], w[3]) - 1.0
            s = u1 * u1 + u2 * u2
            ok = (s > 0.0) & (s < 1.0)
            sv = np.where(ok, s, 0.5)
            num = -2.0 * det_log(sv)
            q = num / sv
            f = np.sqrt(q)
            x[need[ok]] = (u1 * f)[ok]
            need = need[~ok]
        t = cc * x
        v = 1.0 + t
        pos = v > 0.0
        v3 = (v * v) * v
        idx = pend
        w = draw(seed, STREAM_GAMMA, node[idx], comp[idx], j[idx])
        # the uniform draw is consumed only by attempts with v > 0
        j[idx[pos]] += 1
        u = u53_open0(w[0], w[1])
        x2 = x * x
        thr = 1.0 - (0.0331 * x2) * x2
        acc1 = pos & (u < thr)
        v3s = np.where(pos, v3, 1.0)
        inner = (1.0 - v3s) + det_log(v3s)
        rhs = 0.5 * x2 + d * inner
        acc2 = pos & ~acc1 & (det_log(u) < rhs)
        acc = acc1 | acc2
        out[idx[acc]] = d * v3[acc]
        pend = pend[~acc]
    return out


def gen_pi(p: GenParams) -> np.ndarray:
    n, k = p.n_total, p.k
    alpha = p.alpha0 / k
    node = np.repeat(np.arange(n, dtype=np.uint32), k)
    comp = np.tile(np.arange(k, dtype=np.uint32), n)
    w = draw(p.seed, STREAM_POW, node, comp, 0)
    u = u53_open0(w[0], w[1])
    lp = det_log(u) / alpha
    alive = lp >= CUT
    g = np.zeros(n * k)
    if alive.any():
        gg = gamma1p(p.seed, node[alive], comp[alive], alpha)
        g[alive] = gg * det_exp(lp[alive])
    g = g.reshape(n, k)
    total = np.zeros(n)
    for c in range(k):                      # sequential over components
        total = total + g[:, c]
    pi = np.zeros((n, k))
    nz = total > 0.0
    pi[nz] = g[nz] / total[nz, None]
    if (~nz).any():                          # every component cut: one-hot at the largest U
        uu = u.reshape(n, k)
        z = np.nonzero(~nz)[0]
        pi[z, np.argmax(uu[z], axis=1)] = 1.0
    return pi


# ---- edges ---------------------------------------------------------------------------------
def edges_exact(p: GenParams, theta: np.ndarray, pi: np.ndarray) -> np.ndarray:
    n, k = p.n_total, p.k
    dp = p.p_in - p.p_out
    out = []
    for i in range(n - 1):
        j = np.arange(i + 1, n)
        dot = np.zeros(j.size)
        for c in range(k):
            dot = dot + pi[i, c] * pi[j, c]
        pij = (theta[i] * theta[j]) * (p.p_out + dp * dot)
        w = draw(p.seed, STREAM_EDGE, np.uint32(i), j.astype(np.uint32), 0)
        u = u53(w[0], w[1])
        hit = j[u < pij]
        if hit.size:
            out.append(np.stack([np.full(hit.size, i), hit], 1))
    if not out:
        return np.zeros((0, 2), np.int64)
    return np.concatenate(out).astype(np.int64)


def _round_half_up(x: float) -> int:
    return int(np.floor(x + 0.5))


def ball_plan(p: GenParams, theta: np.ndarray, pi: np.ndarray):
    """Cumulative weights and ball counts of the background and of each community."""
    n, k = p.n_total, p.k
    cum_theta = np.cumsum(theta)
    W = cum_theta[-1]
    m_bg = _round_half_up(((p.p_out * W) * W) * 0.5)
    members, cums, m_c = [], [], []
    dp = p.p_in - p.p_out
    for c in range(k):
        mem = np.nonzero(pi[:, c] > 0.0)[0]
        cw = np.cumsum(theta[mem] * pi[mem, c])
        S = cw[-1] if cw.size else 0.0
        members.append(mem)
        cums.append(cw)
        m_c.append(_round_half_up(((dp * S) * S) * 0.5) if cw.size else 0)
    return cum_theta, m_bg, members, cums, np.array(m_c, np.int64)


def _pick(cum: np.ndarray, u: np.ndarray) -> np.ndarray:
    x = u * cum[-1]
    idx = np.searchsorted(cum, x, side="right")
    return np.minimum(idx, cum.size - 1)


def edges_balls(p: GenParams, theta: np.ndarray, pi: np.ndarray) -> np.ndarray:
    cum_theta, m_bg, members, cums, m_c = ball_plan(p, theta, pi)
    pairs = []
    e0 = 0
    def balls(count, e_start):
        e = np.arange(e_start, e_start + count, dtype=np.uint64)
        w = draw(p.seed, STREAM_BALL, (e & np.uint64(0xFFFFFFFF)).astype(np.uint32),
                 (e >> np.uint64(32)).astype(np.uint32), 0)
        return u53(w[0], w[1]), u53(w[2], w[3])
    if m_bg:
        u1, u2 = balls(m_bg, 0)
        pairs.append(np.stack([_pick(cum_theta, u1), _pick(cum_theta, u2)], 1))
    e0 = m_bg
    for c in range(p.k):
        if m_c[c]:
            u1, u2 = balls(int(m_c[c]), e0)
            mem, cw = members[c], cums[c]
            pairs.append(np.stack([mem[_pick(cw, u1)], mem[_pick(cw, u2)]], 1))
        e0 += int(m_c[c])
    if not pairs:
        return np.zeros((0, 2), np.int64)
    pr = np.concatenate(pairs).astype(np.int64)
    pr = pr[pr[:, 0] != pr[:, 1]]
    pr = np.sort(pr, axis=1)
    key = pr[:, 0] * p.n_total + pr[:, 1]
    key = np.unique(key)
    return np.stack([key // p.n_total, key % p.n_total], 1)


# ---- split and blocks ----------------------------------------------------------------------
def split_nodes(p: GenParams):
    n = p.n_total
    w = draw(p.seed, STREAM_SPLIT, np.arange(n, dtype=np.uint32), 0)
    key = (w[0].astype(np.uint64) << np.uint64(32)) | w[1].astype(np.uint64)
    order = np.lexsort((np.arange(n), key))   # by key, ties by node id
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    bounds = np.cumsum([0, p.n_x, p.n_a, p.n_b, p.n_c])
    part = np.searchsorted(bounds, rank, side="right") - 1
    local = rank - bounds[part]
    return part.astype(np.int64), local.astype(np.int64)


BLOCKS = {"XA": (0, 1), "XB": (0, 2), "XC": (0, 3), "BA": (2, 1), "CA": (3, 1)}


def build_blocks(p: GenParams, edges: np.ndarray, part: np.ndarray, local: np.ndarray) -> dict:
    sizes = [p.n_x, p.n_a, p.n_b, p.n_c]
    out = {}
    i, j = edges[:, 0], edges[:, 1]
    for name, (pr, pc) in BLOCKS.items():
        m1 = (part[i] == pr) & (part[j] == pc)
        m2 = (part[j] == pr) & (part[i] == pc)
        rows = np.concatenate([local[i[m1]], local[j[m2]]])
        cols = np.concatenate([local[j[m1]], local[i[m2]]])
        out[name] = coo_to_csr(rows, cols, sizes[pr], sizes[pc])
    return out


def coo_to_csr(rows, cols, n_rows, n_cols):
    """Binary CSR (rowptr, colidx, n_rows, n_cols): duplicates merged, columns ascending."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    key = np.unique(rows * n_cols + cols)
    r, c = key // n_cols, key % n_cols
    rowptr = np.zeros(n_rows + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    return rowptr, c.astype(np.int32), n_rows, n_cols


def generate(p: GenParams) -> Graph:
    theta = gen_theta(p)
    pi = gen_pi(p)
    edges = edges_exact(p, theta, pi) if p.resolved_mode() == "exact" else edges_balls(p, theta, pi)
    part, local = split_nodes(p)
    g = Graph(p, theta, pi, part, local, edges)
    g.blocks = build_blocks(p, edges, part, local)
    return g
