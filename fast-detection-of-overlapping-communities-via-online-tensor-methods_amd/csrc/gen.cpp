// Seeded synthetic MMSB graph generator (host, multi-threaded, deterministic).
//
// The reference consumes externally produced partition files (stpm.h:63-69) and ships no
// generator; this produces seeded synthetic workloads (1000 to 4M nodes per partition, plain or
// degree-corrected) in the same block layout. The model and every
// operation order are those of oracle/gen.py, which this file reproduces bit-for-bit:
//   theta_i = 1 or Pareto(a) / mean;  pi_i ~ Dirichlet(alpha0/k) by Gamma(1+alpha) U^(1/alpha)
//   (Marsaglia-Tsang, polar normals), components with U^(1/alpha) < e^-20 set to 0;
//   edges: one Bernoulli(theta_i theta_j (p_out + (p_in-p_out) pi_i.pi_j)) per pair ("exact"), or
//   background + per-community balls with counts fixed at their means ("balls");
//   split by sorting Philox keys; blocks G_XA, G_XB, G_XC, G_BA, G_CA in CSR.
// Work is split over threads by index ranges; every draw is a function of its indices and every
// sum runs sequentially in index order, so the output does not depend on the thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "host_util.h"
#include "mmsb_stpm.h"
#include "philox.h"

namespace mmsb {

static constexpr double kCut = -20.0;

static void glog(const mmsb_gen_params& p, const char* fmt, ...) {
  if (!p.verbose) return;
  va_list ap;
  va_start(ap, fmt);
  std::vprintf(fmt, ap);
  va_end(ap);
  std::fflush(stdout);
}

static double gamma1p(uint64_t seed, uint32_t node, uint32_t comp, double alpha) {
  const double a = 1.0 + alpha;
  const double d = a - 1.0 / 3.0;
  const double cc = 1.0 / std::sqrt(9.0 * d);
  uint32_t j = 0;
  for (;;) {
    double x;
    for (;;) {
      const u32x4 w = draw(seed, kStreamGamma, node, comp, j);
      ++j;
      const double u1 = 2.0 * u53(w.x, w.y) - 1.0;
      const double u2 = 2.0 * u53(w.z, w.w) - 1.0;
      const double s = u1 * u1 + u2 * u2;
      if (s > 0.0 && s < 1.0) {
        const double num = -2.0 * det_log(s);
        const double q = num / s;
        x = u1 * std::sqrt(q);
        break;
      }
    }
    const double t = cc * x;
    const double v = 1.0 + t;
    if (!(v > 0.0)) continue;
    const double v3 = (v * v) * v;
    const u32x4 w = draw(seed, kStreamGamma, node, comp, j);
    ++j;
    const double u = u53_open0(w.x, w.y);
    const double x2 = x * x;
    const double thr = 1.0 - (0.0331 * x2) * x2;
    if (u < thr) return d * v3;
    const double inner = (1.0 - v3) + det_log(v3);
    const double rhs = 0.5 * x2 + d * inner;
    if (det_log(u) < rhs) return d * v3;
  }
}

struct SparsePi {                // node-major sparse memberships (components with pi > 0)
  std::vector<int64_t> ptr;      // [N+1]
  std::vector<int32_t> comp;
  std::vector<double> val;
};

static void gen_theta(const mmsb_gen_params& p, int threads, std::vector<double>& theta) {
  const int64_t n = (int64_t)p.n_x + p.n_a + p.n_b + p.n_c;
  theta.assign(n, 1.0);
  if (!p.degree_corrected) return;
  parallel_for(n, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      const u32x4 w = draw(p.seed, kStreamTheta, (uint32_t)i, 0, 0);
      const double u = u53_open0(w.x, w.y);
      theta[i] = det_exp(-det_log(u) / p.pareto_shape);
    }
  });
  double total = 0.0;
  for (int64_t i = 0; i < n; ++i) total = total + theta[i];
  const double mean = total / (double)n;
  for (int64_t i = 0; i < n; ++i) theta[i] = theta[i] / mean;
}

static void gen_pi(const mmsb_gen_params& p, int threads, SparsePi& sp) {
  const int64_t n = (int64_t)p.n_x + p.n_a + p.n_b + p.n_c;
  const int k = p.k;
  const double alpha = p.alpha0 / (double)k;
  const int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(n, 64 * (int64_t)threads));
  std::vector<std::vector<int32_t>> ccomp(nchunk);
  std::vector<std::vector<double>> cval(nchunk);
  std::vector<std::vector<int32_t>> ccount(nchunk);
  parallel_for(nchunk, threads, [&](int64_t clo, int64_t chi) {
    std::vector<double> g(k), u(k);
    for (int64_t ch = clo; ch < chi; ++ch) {
      const int64_t lo = n * ch / nchunk, hi = n * (ch + 1) / nchunk;
      for (int64_t i = lo; i < hi; ++i) {
        double total = 0.0;
        for (int c = 0; c < k; ++c) {
          const u32x4 w = draw(p.seed, kStreamPow, (uint32_t)i, (uint32_t)c, 0);
          u[c] = u53_open0(w.x, w.y);
          const double lp = det_log(u[c]) / alpha;
          g[c] = (lp >= kCut) ? gamma1p(p.seed, (uint32_t)i, (uint32_t)c, alpha) * det_exp(lp) : 0.0;
          total = total + g[c];
        }
        int32_t cnt = 0;
        if (total > 0.0) {
          for (int c = 0; c < k; ++c) {
            const double v = g[c] / total;
            if (v > 0.0) {
              ccomp[ch].push_back(c);
              cval[ch].push_back(v);
              ++cnt;
            }
          }
        } else {  // every component cut: one-hot at the largest U (first on ties)
          int best = 0;
          for (int c = 1; c < k; ++c)
            if (u[c] > u[best]) best = c;
          ccomp[ch].push_back(best);
          cval[ch].push_back(1.0);
          cnt = 1;
        }
        ccount[ch].push_back(cnt);
      }
    }
  });
  sp.ptr.assign(n + 1, 0);
  int64_t node = 0, total = 0;
  for (int64_t ch = 0; ch < nchunk; ++ch)
    for (int32_t c : ccount[ch]) {
      total += c;
      sp.ptr[++node] = total;
    }
  sp.comp.resize(total);
  sp.val.resize(total);
  std::vector<int64_t> off(nchunk + 1, 0);
  for (int64_t ch = 0; ch < nchunk; ++ch) off[ch + 1] = off[ch] + (int64_t)ccomp[ch].size();
  parallel_for(nchunk, threads, [&](int64_t clo, int64_t chi) {
    for (int64_t ch = clo; ch < chi; ++ch) {
      std::copy(ccomp[ch].begin(), ccomp[ch].end(), sp.comp.begin() + off[ch]);
      std::copy(cval[ch].begin(), cval[ch].end(), sp.val.begin() + off[ch]);
    }
  });
}

// sequential dot of two sparse memberships in ascending component order (== dense sequential sum)
static double sparse_dot(const SparsePi& sp, int64_t i, int64_t j) {
  int64_t a = sp.ptr[i], ae = sp.ptr[i + 1], b = sp.ptr[j], be = sp.ptr[j + 1];
  double acc = 0.0;
  while (a < ae && b < be) {
    if (sp.comp[a] == sp.comp[b]) {
      acc = acc + sp.val[a] * sp.val[b];
      ++a;
      ++b;
    } else if (sp.comp[a] < sp.comp[b]) {
      ++a;
    } else {
      ++b;
    }
  }
  return acc;
}

static void edges_exact(const mmsb_gen_params& p, int threads, const std::vector<double>& theta,
                        const SparsePi& sp, std::vector<uint64_t>& keys) {
  const int64_t n = (int64_t)theta.size();
  const double dp = p.p_in - p.p_out;
  const int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(n, 256 * (int64_t)threads));
  std::vector<std::vector<uint64_t>> part(nchunk);
  // rows are assigned to chunks round-robin so the triangular work is balanced
  parallel_for(nchunk, threads, [&](int64_t clo, int64_t chi) {
    for (int64_t ch = clo; ch < chi; ++ch)
      for (int64_t i = ch; i < n - 1; i += nchunk)
        for (int64_t j = i + 1; j < n; ++j) {
          const double dot = sparse_dot(sp, i, j);
          const double pij = (theta[i] * theta[j]) * (p.p_out + dp * dot);
          const u32x4 w = draw(p.seed, kStreamEdge, (uint32_t)i, (uint32_t)j, 0);
          if (u53(w.x, w.y) < pij) part[ch].push_back((uint64_t)i * (uint64_t)n + (uint64_t)j);
        }
  });
  keys.clear();
  for (auto& v : part) keys.insert(keys.end(), v.begin(), v.end());
  parallel_sort_unique(keys, threads);
}

static int64_t pick(const double* cum, int64_t size, double u) {
  const double x = u * cum[size - 1];
  const int64_t idx = std::upper_bound(cum, cum + size, x) - cum;
  return idx < size - 1 ? idx : size - 1;
}

// pick() over one large cumulative array through a guide table: guide[j] is the first index whose
// cum exceeds t_j = j * W / G. For t_j <= x < t_{j+1} the first index with cum > x lies in
// [guide[j], guide[j+1]], so the search there returns exactly what the search over the whole array
// returns, touching one or two cache lines instead of ~log2(size) scattered ones.
struct GuidedCum {
  const double* cum = nullptr;
  int64_t size = 0;
  int64_t G = 0;
  std::vector<double> t;        // [G + 1] bucket starts
  std::vector<int64_t> guide;   // [G + 1]

  GuidedCum(const double* c, int64_t n, int threads) : cum(c), size(n) {
    G = std::max<int64_t>(1, n / 16);
    const double W = cum[size - 1];
    t.resize(G + 1);
    guide.resize(G + 1);
    parallel_for(G + 1, threads, [&](int64_t lo, int64_t hi) {
      for (int64_t j = lo; j < hi; ++j) {
        t[j] = (double)j * W / (double)G;
        guide[j] = std::upper_bound(cum, cum + size, t[j]) - cum;
      }
    });
  }
  int64_t operator()(double u) const {
    const double x = u * cum[size - 1];
    int64_t j = (int64_t)(x / cum[size - 1] * (double)G);
    j = std::min<int64_t>(std::max<int64_t>(j, 0), G - 1);
    while (j > 0 && x < t[j]) --j;
    while (j < G - 1 && x >= t[j + 1]) ++j;
    const int64_t lo = guide[j], hi = std::min<int64_t>(guide[j + 1] + 1, size);
    const int64_t idx = std::upper_bound(cum + lo, cum + hi, x) - cum;
    return idx < size - 1 ? idx : size - 1;
  }
};

static void edges_balls(const mmsb_gen_params& p, int threads, const std::vector<double>& theta,
                        const SparsePi& sp, std::vector<uint64_t>& keys) {
  const int64_t n = (int64_t)theta.size();
  const int k = p.k;
  const double dp = p.p_in - p.p_out;
  std::vector<double> cum_theta(n);
  double acc = 0.0;
  for (int64_t i = 0; i < n; ++i) cum_theta[i] = acc = acc + theta[i];
  const double W = cum_theta[n - 1];
  const int64_t m_bg = (int64_t)std::floor(((p.p_out * W) * W) * 0.5 + 0.5);
  // per-community members (ascending node id) and cumulative weights theta_i pi_ic
  std::vector<int64_t> cptr(k + 1, 0);
  for (size_t e = 0; e < sp.comp.size(); ++e) cptr[sp.comp[e] + 1]++;
  for (int c = 0; c < k; ++c) cptr[c + 1] += cptr[c];
  std::vector<int32_t> mem(sp.comp.size());
  std::vector<double> cw(sp.comp.size());
  {
    std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t e = sp.ptr[i]; e < sp.ptr[i + 1]; ++e) {
        const int c = sp.comp[e];
        mem[fill[c]] = (int32_t)i;
        cw[fill[c]] = theta[i] * sp.val[e];
        ++fill[c];
      }
  }
  std::vector<int64_t> m_c(k, 0), e0(k + 1, 0);
  parallel_for(k, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t c = lo; c < hi; ++c) {
      double a = 0.0;
      for (int64_t e = cptr[c]; e < cptr[c + 1]; ++e) cw[e] = a = a + cw[e];
      if (cptr[c + 1] > cptr[c]) m_c[c] = (int64_t)std::floor(((dp * a) * a) * 0.5 + 0.5);
    }
  });
  e0[0] = m_bg;
  for (int c = 0; c < k; ++c) e0[c + 1] = e0[c] + m_c[c];
  const int64_t total = e0[k];
  glog(p, "generate: dropping %lld balls\n", (long long)total);
  const GuidedCum bg(cum_theta.data(), n, threads);
  std::vector<uint64_t> raw(total);
  // every ball is a function of its index alone; the parts only pace the progress lines
  const int parts = total > ((int64_t)1 << 28) ? 8 : 1;
  for (int part = 0; part < parts; ++part) {
    const int64_t plo = total * part / parts, phi = total * (part + 1) / parts;
    parallel_for(phi - plo, threads, [&](int64_t lo0, int64_t hi0) {
      const int64_t lo = plo + lo0, hi = plo + hi0;
      int c = (int)(std::upper_bound(e0.begin(), e0.end(), lo) - e0.begin()) - 1;  // -1 = background
      for (int64_t e = lo; e < hi; ++e) {
        while (c < k - 1 && e >= e0[c + 1]) ++c;
        const u32x4 w = draw(p.seed, kStreamBall, (uint32_t)((uint64_t)e & 0xFFFFFFFFu), (uint32_t)((uint64_t)e >> 32), 0);
        const double u1 = u53(w.x, w.y), u2 = u53(w.z, w.w);
        int64_t i, j;
        if (c < 0) {
          i = bg(u1);
          j = bg(u2);
        } else {
          const int64_t b = cptr[c], sz = cptr[c + 1] - cptr[c];
          i = mem[b + pick(cw.data() + b, sz, u1)];
          j = mem[b + pick(cw.data() + b, sz, u2)];
        }
        raw[e] = (i == j) ? ~0ull : (uint64_t)std::min(i, j) * (uint64_t)n + (uint64_t)std::max(i, j);
      }
    });
    if (parts > 1) glog(p, "generate: balls %d/%d\n", part + 1, parts);
  }
  glog(p, "generate: sorting the edge keys\n");
  parallel_sort_unique(raw, threads);
  if (!raw.empty() && raw.back() == ~0ull) raw.pop_back();
  keys.swap(raw);
}

static void split_nodes(const mmsb_gen_params& p, int threads, std::vector<int8_t>& part, std::vector<int32_t>& local) {
  const int64_t n = (int64_t)p.n_x + p.n_a + p.n_b + p.n_c;
  std::vector<std::pair<uint64_t, int32_t>> key(n);
  parallel_for(n, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      const u32x4 w = draw(p.seed, kStreamSplit, (uint32_t)i, 0, 0);
      key[i] = {((uint64_t)w.x << 32) | w.y, (int32_t)i};
    }
  });
  std::sort(key.begin(), key.end());
  part.assign(n, 0);
  local.assign(n, 0);
  const int64_t b[5] = {0, p.n_x, (int64_t)p.n_x + p.n_a, (int64_t)p.n_x + p.n_a + p.n_b, n};
  int pp = 0;
  for (int64_t r = 0; r < n; ++r) {
    while (pp < 3 && r >= b[pp + 1]) ++pp;
    part[key[r].second] = (int8_t)pp;
    local[key[r].second] = (int32_t)(r - b[pp]);
  }
}

static int build_blocks(const mmsb_gen_params& p, int threads, const std::vector<uint64_t>& keys,
                        const std::vector<int8_t>& part, const std::vector<int32_t>& local, mmsb_graph* g) {
  const int64_t n = (int64_t)part.size();
  const int sizes[4] = {p.n_x, p.n_a, p.n_b, p.n_c};
  const int pr[5] = {0, 0, 0, 2, 3}, pc[5] = {1, 2, 3, 1, 1};
  mmsb_csr* out[5] = {&g->xa, &g->xb, &g->xc, &g->ba, &g->ca};
  // Each unordered pair of partitions feeds at most one block: route[part_i][part_j] = block, with
  // bit 3 set when node j indexes the block's rows.
  int8_t route[4][4];
  std::memset(route, -1, sizeof(route));
  for (int b = 0; b < 5; ++b) {
    route[pr[b]][pc[b]] = (int8_t)b;
    route[pc[b]][pr[b]] = (int8_t)(b | 8);
  }
  auto locate = [&](uint64_t key, int& b, int32_t& row, int32_t& col) -> bool {
    const int64_t i = (int64_t)(key / (uint64_t)n), j = (int64_t)(key % (uint64_t)n);
    const int8_t r = route[part[i]][part[j]];
    if (r < 0) return false;
    b = r & 7;
    if (r & 8) { row = local[j]; col = local[i]; } else { row = local[i]; col = local[j]; }
    return true;
  };
  std::vector<std::vector<int64_t>> cnt(5);
  for (int b = 0; b < 5; ++b) cnt[b].assign((size_t)sizes[pr[b]] + 1, 0);
  // Rows are filled in any order and sorted afterwards (keys are unique, so a row holds distinct
  // columns): the CSR does not depend on the thread count.
  const int64_t nkeys = (int64_t)keys.size();
  parallel_for(nkeys, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t t = lo; t < hi; ++t) {
      int b;
      int32_t r, c;
      if (locate(keys[t], b, r, c)) __atomic_fetch_add(&cnt[b][r + 1], (int64_t)1, __ATOMIC_RELAXED);
    }
  });
  static const char* const kNames[5] = {"XA", "XB", "XC", "BA", "CA"};
  std::vector<std::vector<int64_t>> fill(5);
  for (int b = 0; b < 5; ++b) {
    const int nr = sizes[pr[b]], nc = sizes[pc[b]];
    for (int i = 0; i < nr; ++i) cnt[b][i + 1] += cnt[b][i];
    const int64_t nnz = cnt[b][nr];
    if (nnz >= INT32_MAX) return set_error(MMSB_ERR_ARG, "generate: block %d has %lld nonzeros (int32 CSR)", b, (long long)nnz);
    glog(p, "generate: block %s (%lld nonzeros)\n", kNames[b], (long long)nnz);
    mmsb_csr* m = out[b];
    m->n_rows = nr;
    m->n_cols = nc;
    m->nnz = nnz;
    m->rowptr = (int32_t*)std::malloc(sizeof(int32_t) * (nr + 1));
    m->colidx = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(nnz, 1));
    m->vals = nullptr;
    if (!m->rowptr || !m->colidx) return set_error(MMSB_ERR_NOMEM, "generate: out of host memory");
    for (int i = 0; i <= nr; ++i) m->rowptr[i] = (int32_t)cnt[b][i];
    fill[b].assign(cnt[b].begin(), cnt[b].end() - 1);
  }
  parallel_for(nkeys, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t t = lo; t < hi; ++t) {
      int b;
      int32_t r, c;
      if (locate(keys[t], b, r, c)) out[b]->colidx[__atomic_fetch_add(&fill[b][r], (int64_t)1, __ATOMIC_RELAXED)] = c;
    }
  });
  for (int b = 0; b < 5; ++b) {
    mmsb_csr* m = out[b];
    parallel_for(m->n_rows, threads, [&](int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi; ++i) std::sort(m->colidx + m->rowptr[i], m->colidx + m->rowptr[i + 1]);
    });
  }
  glog(p, "generate: blocks built\n");
  return MMSB_OK;
}

}  // namespace mmsb

using namespace mmsb;

extern "C" void mmsb_gen_params_default(mmsb_gen_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->n_x = p->n_a = p->n_b = p->n_c = 1000;
  p->k = 10;
  p->alpha0 = 1.0;
  p->p_in = 0.6;
  p->p_out = 0.02;
  p->pareto_shape = 2.5;
}

extern "C" int mmsb_generate(const mmsb_gen_params* pp, int32_t keep_truth, mmsb_graph** out) {
  clear_error();
  MMSB_CHECK_ARG(pp && out, "generate: null argument");
  const mmsb_gen_params p = *pp;
  MMSB_CHECK_ARG(p.n_x > 0 && p.n_a > 0 && p.n_b > 0 && p.n_c > 0 && p.k > 0, "generate: sizes must be positive");
  MMSB_CHECK_ARG(p.alpha0 > 0 && p.p_in >= 0 && p.p_out >= 0 && p.p_in <= 1 && p.p_out <= 1, "generate: bad alpha0/p_in/p_out");
  const int64_t n = (int64_t)p.n_x + p.n_a + p.n_b + p.n_c;
  MMSB_CHECK_ARG(n < ((int64_t)1 << 32), "generate: at most 2^32 nodes");
  const int threads = p.threads > 0 ? p.threads : default_threads();
  int mode = p.mode;
  if (mode == MMSB_GEN_AUTO) mode = (n <= 50000) ? MMSB_GEN_EXACT : MMSB_GEN_BALLS;
  mmsb_graph* g = (mmsb_graph*)std::calloc(1, sizeof(mmsb_graph));
  if (!g) return set_error(MMSB_ERR_NOMEM, "generate: out of host memory");
  std::vector<double> theta;
  SparsePi sp;
  gen_theta(p, threads, theta);
  gen_pi(p, threads, sp);
  glog(p, "generate: memberships of %lld nodes (%lld nonzero components)\n", (long long)n, (long long)sp.comp.size());
  std::vector<uint64_t> keys;
  if (mode == MMSB_GEN_EXACT)
    edges_exact(p, threads, theta, sp, keys);
  else
    edges_balls(p, threads, theta, sp, keys);
  std::vector<int8_t> part;
  std::vector<int32_t> local;
  split_nodes(p, threads, part, local);
  int st = build_blocks(p, threads, keys, part, local, g);
  if (st != MMSB_OK) {
    mmsb_graph_free(g);
    return st;
  }
  g->n_nodes = n;
  g->n_edges = (int64_t)keys.size();
  g->k = p.k;
  if (keep_truth) {
    g->theta = (double*)std::malloc(sizeof(double) * n);
    g->part = (int8_t*)std::malloc(n);
    g->local = (int32_t*)std::malloc(sizeof(int32_t) * n);
    g->pi = (double*)std::calloc((size_t)n * p.k, sizeof(double));
    if (!g->theta || !g->part || !g->local || !g->pi) {
      mmsb_graph_free(g);
      return set_error(MMSB_ERR_NOMEM, "generate: out of host memory for ground truth");
    }
    std::memcpy(g->theta, theta.data(), sizeof(double) * n);
    std::memcpy(g->part, part.data(), n);
    std::memcpy(g->local, local.data(), sizeof(int32_t) * n);
    for (int64_t i = 0; i < n; ++i)
      for (int64_t e = sp.ptr[i]; e < sp.ptr[i + 1]; ++e) g->pi[i * p.k + sp.comp[e]] = sp.val[e];
  }
  *out = g;
  return MMSB_OK;
}

static void free_csr(mmsb_csr* m) {
  std::free(m->rowptr);
  std::free(m->colidx);
  std::free(m->vals);
  std::memset(m, 0, sizeof(*m));
}

extern "C" void mmsb_graph_free(mmsb_graph* g) {
  if (!g) return;
  free_csr(&g->xa);
  free_csr(&g->xb);
  free_csr(&g->xc);
  free_csr(&g->ba);
  free_csr(&g->ca);
  std::free(g->theta);
  std::free(g->part);
  std::free(g->local);
  std::free(g->pi);
  std::free(g);
}
