// CPU ORACLE — test infrastructure only.
//
// A plain C++/OpenMP restatement of the reference's maximal-clique path
// (yen-hsiang-chang/mce-gpu) used to check the GPU engine and as the timed CPU
// baseline. Nothing in the product links or calls this file; only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load liboracle.so.
//
// What it restates, and from where:
//   oracle_peel      batched k-core peeling: every round removes the current
//                    frontier (alive vertices of degree <= level) and assigns it
//                    the round index as priority; the next frontier is the set of
//                    vertices whose degree falls to `level` in that round; the
//                    level rises when a round finds nothing. Degeneracy = the
//                    last level that removed vertices.
//                    include/kcore.cuh:80-95 (process_degree), :107-143,
//                    :233-316 (findKcoreIncremental_async), :313-319 (Max Core).
//                    Pinned by: oracle_core_numbers (Batagelj-Zaversnik, an
//                    independent O(m) algorithm) and networkx core_number
//                    fixtures; the reference's own quirk of leaving the final
//                    batch at priority m (kcore.cuh:257, :275-277) orders the
//                    same way as giving it the last round index.
//   oracle_orient    row u -> [earlier neighbours asc | later neighbours asc],
//                    w later than u iff pri[u] < pri[w] or (== and u < w);
//                    split[u] = start of the later part.
//                    include/main_support.cuh:6-84, src/main.cu:102-146.
//                    Pinned by: out-degree <= degeneracy and the sorted-halves
//                    invariants checked in tests.
//   oracle_mce       L1: one subproblem per vertex v, R={v}, P=N+(v), X=N-(v);
//                    L2: one per oriented arc u->w, R={u,w}, P=N+(u)&N+(w),
//                    X=N(u)&N-(w); an L2 arc with P and X empty is a clique.
//                    mce/mce_kernel_nowl.cuh:45-108 (L1), :250-307 (L2).
//                    Bron-Kerbosch with pivoting: pivot = argmax |P & N(u)| over
//                    u in P and the X vertices held as bitmap rows; an X vertex
//                    covering all of P ends the branch; a leaf with P and X
//                    empty is a maximal clique. mce/mce_utils.cuh:338-524.
//                    IP keeps the X vertices outside the induced subgraph as an
//                    explicit list tested by membership; IPX gives each of them a
//                    bitmap row over P (mce/mce_utils.cuh:265-319).
//                    Pinned by: closed forms (complete, multipartite, Moon-Moser,
//                    cycles, Petersen, ...), brute force over all vertex subsets
//                    for n <= 16, and networkx find_cliques fixtures.
//   Tie-breaks (lowest index among equal pivot scores; candidates in ascending
//   bit order) are this build's own deterministic rules, shared with the GPU
//   engine so that tree-node counts agree exactly; counts, histograms and
//   hashes do not depend on them.
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t splitmix64(uint64_t x) {
  x += kGolden;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

constexpr int kHist = 4098;

struct Graph {
  uint32_t n;
  const uint64_t *rp;   // original CSR row pointers
  const uint32_t *ci;   // original CSR (sorted rows)
  std::vector<uint64_t> split;  // oriented: start of later neighbours
  std::vector<uint32_t> oc;     // oriented col
  // N-(v) = oc[rp[v], split[v]), N+(v) = oc[split[v], rp[v+1])
};

struct Acc {
  uint64_t count = 0, hash = 0, nodes = 0;
  uint64_t hist[kHist];
  double ops = 0, lds = 0;
  Acc() { memset(hist, 0, sizeof(hist)); }
};

// One Bron-Kerbosch subproblem over an induced vertex set S (sorted ids).
struct Solver {
  const Graph *g;
  bool ipx;
  bool work;  // accumulate OPS / LDS work figures
  uint32_t p = 0, W = 0, q = 0;
  std::vector<uint32_t> S;
  std::vector<uint64_t> A;    // p*W adjacency rows over S
  std::vector<uint64_t> XR;   // q*W bitmap rows of outside X vertices (IPX)
  std::vector<uint32_t> xid;  // outside X: row id (IPX) or vertex id (IP)
  std::vector<uint64_t> bits; // per-depth P and X words
  std::vector<std::vector<uint32_t>> xl;  // per-depth outside-X lists
  Acc *acc = nullptr;
  uint64_t nodes = 0, cliques = 0;

  int find(uint32_t y) const {
    auto it = std::lower_bound(S.begin(), S.end(), y);
    return (it != S.end() && *it == y) ? (int)(it - S.begin()) : -1;
  }
  bool adj_outside(uint32_t e, uint32_t ci) const {
    if (ipx) return (XR[(uint64_t)e * W + (ci >> 6)] >> (ci & 63)) & 1;
    const uint32_t c = S[ci];
    const uint32_t *b = g->oc.data() + g->rp[c], *en = g->oc.data() + g->split[c];
    return std::binary_search(b, en, e);
  }
  uint64_t *Pw(uint32_t d) { return bits.data() + (uint64_t)(2 * d) * W; }
  uint64_t *Xw(uint32_t d) { return bits.data() + (uint64_t)(2 * d + 1) * W; }

  // Builds A over S and the outside-X rows; xs = outside X vertex ids (sorted).
  void build(const std::vector<uint32_t> &xs) {
    p = (uint32_t)S.size();
    W = (p + 63) / 64;
    A.assign((uint64_t)p * W, 0);
    for (uint32_t i = 0; i < p; ++i) {
      const uint32_t u = S[i];
      for (uint64_t k = g->split[u]; k < g->rp[u + 1]; ++k) {
        const int j = find(g->oc[k]);
        if (j >= 0) {
          A[(uint64_t)i * W + (j >> 6)] |= 1ull << (j & 63);
          A[(uint64_t)j * W + (i >> 6)] |= 1ull << (i & 63);
        }
      }
    }
    xid.clear();
    XR.clear();
    if (ipx) {
      std::vector<uint64_t> row(W);
      for (uint32_t x : xs) {
        std::fill(row.begin(), row.end(), 0);
        bool any = false;
        for (uint64_t k = g->split[x]; k < g->rp[x + 1]; ++k) {
          const int j = find(g->oc[k]);
          if (j >= 0) {
            row[j >> 6] |= 1ull << (j & 63);
            any = true;
          }
        }
        if (any) {  // a row with no neighbour in S never survives a branch
          xid.push_back((uint32_t)(XR.size() / W));
          XR.insert(XR.end(), row.begin(), row.end());
        }
      }
    } else {
      xid = xs;
    }
    q = (uint32_t)xid.size();
  }

  // Pivot over P|X bits then outside-X rows (IPX). Returns (score, index).
  void pivot(const uint64_t *P, const uint64_t *X, const std::vector<uint32_t> &xe, uint32_t &best,
             uint32_t &bidx) {
    best = 0;
    bidx = UINT32_MAX;
    uint64_t cand_rows = 0;
    for (uint32_t k = 0; k < W; ++k) {
      uint64_t c = P[k] | X[k];
      while (c) {
        const uint32_t u = k * 64 + __builtin_ctzll(c);
        c &= c - 1;
        uint32_t s = 0;
        const uint64_t *row = &A[(uint64_t)u * W];
        for (uint32_t j = 0; j < W; ++j) s += __builtin_popcountll(P[j] & row[j]);
        ++cand_rows;
        if (bidx == UINT32_MAX || s > best) {
          best = s;
          bidx = u;
        }
      }
    }
    if (ipx) {
      for (uint32_t e : xe) {
        uint32_t s = 0;
        const uint64_t *row = &XR[(uint64_t)e * W];
        for (uint32_t j = 0; j < W; ++j) s += __builtin_popcountll(P[j] & row[j]);
        ++cand_rows;
        const uint32_t idx = p + e;
        if (bidx == UINT32_MAX || s > best || (s == best && idx < bidx)) {
          best = s;
          bidx = idx;
        }
      }
    }
    if (work) {
      acc->ops += 2.0 * 2.0 * W * (double)cand_rows;
      acc->lds += 8.0 * W * (double)cand_rows;
    }
  }

  void report(uint32_t size, uint64_t mix) {
    ++cliques;
    acc->count++;
    acc->hash += fmix64(mix + (uint64_t)size * kGolden);
    acc->hist[std::min<uint32_t>(size, kHist - 1)]++;
  }

  // Expands frame d (P, X at depth d; outside list xl[d]); R has rsize vertices.
  void expand(uint32_t d, uint32_t piv, uint32_t rsize, uint64_t rmix) {
    const uint64_t *N = piv < p ? &A[(uint64_t)piv * W] : &XR[(uint64_t)(piv - p) * W];
    for (uint32_t k = 0; k < W; ++k) {
      while (true) {
        uint64_t *P = Pw(d), *X = Xw(d);
        const uint64_t B = P[k] & ~N[k];
        if (!B) break;
        const uint32_t b = __builtin_ctzll(B);
        const uint32_t c = k * 64 + b;
        P[k] &= ~(1ull << b);
        X[k] |= 1ull << b;
        ++nodes;
        uint64_t *Pc = Pw(d + 1), *Xc = Xw(d + 1);
        const uint64_t *Ac = &A[(uint64_t)c * W];
        bool pany = false, xany = false;
        uint32_t pc = 0;
        for (uint32_t j = 0; j < W; ++j) {
          Pc[j] = P[j] & Ac[j];
          Xc[j] = X[j] & Ac[j];
          pany |= Pc[j] != 0;
          xany |= Xc[j] != 0;
          pc += __builtin_popcountll(Pc[j]);
        }
        if (work) {
          acc->ops += 2.0 * 4.0 * W;
          acc->lds += 8.0 * W * 4.0;
        }
        const uint64_t cmix = rmix + splitmix64(S[c]);
        if (!pany) {
          if (xany) continue;
          bool alive = false;
          for (uint32_t e : xl[d])
            if (adj_outside(e, c)) {
              alive = true;
              break;
            }
          if (!alive) report(rsize + 1, cmix);
          continue;
        }
        std::vector<uint32_t> &child = xl[d + 1];
        child.clear();
        for (uint32_t e : xl[d])
          if (adj_outside(e, c)) child.push_back(e);
        uint32_t best, bidx;
        pivot(Pc, Xc, child, best, bidx);
        if (best == pc) continue;  // an X vertex covers P: no maximal clique below
        expand(d + 1, bidx, rsize + 1, cmix);
      }
    }
  }

  // Knuth's random-probe estimate of this subproblem's tree (D. E. Knuth,
  // "Estimating the efficiency of backtrack programs", Math. Comp. 29, 1975):
  // one root-to-leaf path, taking at every node one of its k children
  // uniformly; the product of the k's along the path, times 1 when the path
  // ends in a maximal clique, is an unbiased estimate of the subtree's
  // maximal cliques, and sum over the path of (product so far) counts its
  // nodes the same way. Children are those of expand(): candidate c_i (the
  // i-th bit of P & ~N(pivot) in index order) sees P without c_1..c_{i-1} and
  // X with them. Needs build() and the depth-0 frame of run().
  void probe(uint64_t &rng, double &cl, double &nd) {
    cl = 0;
    nd = 0;
    std::vector<uint64_t> P(Pw(0), Pw(0) + W), X(Xw(0), Xw(0) + W), Pc(W), Xc(W);
    std::vector<uint32_t> xe = xl[0], xc;
    uint32_t best, bidx;
    pivot(P.data(), X.data(), xe, best, bidx);
    if (best == p) return;
    double wgt = 1;
    while (true) {
      const uint64_t *N = bidx < p ? &A[(uint64_t)bidx * W] : &XR[(uint64_t)(bidx - p) * W];
      uint32_t k = 0;
      for (uint32_t j = 0; j < W; ++j) k += __builtin_popcountll(P[j] & ~N[j]);
      if (k == 0) return;
      rng = splitmix64(rng);
      uint32_t pick = (uint32_t)(((unsigned __int128)rng * k) >> 64);
      wgt *= k;
      nd += wgt;
      // move the candidates before the picked one from P to X
      uint32_t c = 0;
      for (uint32_t j = 0; j < W; ++j) {
        uint64_t B = P[j] & ~N[j];
        while (B) {
          const uint32_t b = __builtin_ctzll(B);
          B &= B - 1;
          if (pick == 0) {
            c = j * 64 + b;
            j = W;
            break;
          }
          --pick;
          P[j] &= ~(1ull << b);
          X[j] |= 1ull << b;
        }
      }
      P[c >> 6] &= ~(1ull << (c & 63));
      const uint64_t *Ac = &A[(uint64_t)c * W];
      bool pany = false, xany = false;
      uint32_t pc = 0;
      for (uint32_t j = 0; j < W; ++j) {
        Pc[j] = P[j] & Ac[j];
        Xc[j] = X[j] & Ac[j];
        pany |= Pc[j] != 0;
        xany |= Xc[j] != 0;
        pc += __builtin_popcountll(Pc[j]);
      }
      xc.clear();
      for (uint32_t e : xe)
        if (adj_outside(e, c)) xc.push_back(e);
      if (!pany) {
        if (!xany && xc.empty()) cl = wgt;
        return;
      }
      pivot(Pc.data(), Xc.data(), xc, best, bidx);
      if (best == pc) return;
      P.swap(Pc);
      X.swap(Xc);
      xe.swap(xc);
    }
  }

  // Depth-0 frame of run() without the enumeration (for probe()).
  bool start(const std::vector<uint32_t> &xs) {
    if (S.empty()) return false;
    build(xs);
    bits.assign((uint64_t)2 * (p + 2) * W, 0);
    if (xl.size() < p + 2) xl.resize(p + 2);
    uint64_t *P = Pw(0);
    for (uint32_t i = 0; i < p; ++i) P[i >> 6] |= 1ull << (i & 63);
    xl[0].resize(q);
    for (uint32_t e = 0; e < q; ++e) xl[0][e] = ipx ? e : xid[e];
    return true;
  }

  // Runs BK on (R, P = all of S, X = {} inside S, outside X = xs).
  void run(const std::vector<uint32_t> &xs, uint32_t rsize, uint64_t rmix) {
    nodes = 0;
    cliques = 0;
    if (S.empty()) {
      if (xs.empty() && rsize >= 2) report(rsize, rmix);
      return;
    }
    build(xs);
    // depth never exceeds p + 1 (each level moves one vertex of S into R)
    bits.assign((uint64_t)2 * (p + 2) * W, 0);
    if (xl.size() < p + 2) xl.resize(p + 2);
    uint64_t *P = Pw(0);
    for (uint32_t i = 0; i < p; ++i) P[i >> 6] |= 1ull << (i & 63);
    xl[0].resize(q);
    for (uint32_t e = 0; e < q; ++e) xl[0][e] = ipx ? e : xid[e];
    uint32_t best, bidx;
    pivot(Pw(0), Xw(0), xl[0], best, bidx);
    if (best == p) return;
    expand(0, bidx, rsize, rmix);
  }
};

void intersect(const uint32_t *a, const uint32_t *ae, const uint32_t *b, const uint32_t *be,
               std::vector<uint32_t> &out) {
  out.clear();
  while (a < ae && b < be) {
    if (*a < *b) ++a;
    else if (*b < *a) ++b;
    else {
      out.push_back(*a);
      ++a;
      ++b;
    }
  }
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

typedef struct oracle_result {
  uint64_t count, hash, nodes;
  uint64_t hist[kHist];
  uint32_t max_clique_size, degeneracy, max_degree, max_out_degree;
  double t_order_ms, t_count_ms;
  double b_hbm, ops, b_lds;  // algorithmic work figures
  uint32_t threads;
} oracle_result;

// Batched peeling (see header). prio[n] receives round indices. Returns the
// degeneracy. rounds_out (nullable) receives the number of rounds.
uint32_t oracle_peel(uint32_t n, const uint64_t *rp, const uint32_t *ci, uint32_t *prio,
                     uint32_t *rounds_out) {
  std::vector<int64_t> deg(n);
  std::vector<uint8_t> gone(n, 0);
  std::vector<uint32_t> front, next, alive;
  alive.reserve(n);
  for (uint32_t v = 0; v < n; ++v) {
    deg[v] = (int64_t)(rp[v + 1] - rp[v]);
    alive.push_back(v);
  }
  int64_t level = 0;
  uint32_t round = 0, degeneracy = 0;
  uint64_t removed = 0;
  while (removed < n) {
    // level step: compact alive, find the frontier at the smallest level
    int64_t mn = INT64_MAX;
    size_t w = 0;
    for (uint32_t v : alive)
      if (!gone[v]) {
        alive[w++] = v;
        mn = std::min(mn, deg[v]);
      }
    alive.resize(w);
    level = std::max(level, mn);
    front.clear();
    for (uint32_t v : alive)
      if (deg[v] <= level) front.push_back(v);
    degeneracy = (uint32_t)level;
    while (!front.empty()) {
      for (uint32_t v : front) {
        prio[v] = round;
        gone[v] = 1;
      }
      removed += front.size();
      next.clear();
      for (uint32_t v : front)
        for (uint64_t k = rp[v]; k < rp[v + 1]; ++k) {
          const uint32_t u = ci[k];
          if (gone[u]) continue;
          if (--deg[u] == level) next.push_back(u);
        }
      ++round;
      front.swap(next);
    }
  }
  if (rounds_out) *rounds_out = round;
  return degeneracy;
}

// Batagelj-Zaversnik O(m) core decomposition; returns max core.
uint32_t oracle_core_numbers(uint32_t n, const uint64_t *rp, const uint32_t *ci, uint32_t *core) {
  uint32_t md = 0;
  std::vector<uint32_t> deg(n), pos(n), vert(n);
  for (uint32_t v = 0; v < n; ++v) {
    deg[v] = (uint32_t)(rp[v + 1] - rp[v]);
    md = std::max(md, deg[v]);
  }
  std::vector<uint32_t> bin(md + 2, 0);
  for (uint32_t v = 0; v < n; ++v) bin[deg[v]]++;
  uint32_t start = 0;
  for (uint32_t d = 0; d <= md; ++d) {
    const uint32_t num = bin[d];
    bin[d] = start;
    start += num;
  }
  for (uint32_t v = 0; v < n; ++v) {
    pos[v] = bin[deg[v]];
    vert[pos[v]] = v;
    bin[deg[v]]++;
  }
  for (uint32_t d = md; d >= 1; --d) bin[d] = bin[d - 1];
  bin[0] = 0;
  uint32_t kmax = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t v = vert[i];
    kmax = std::max(kmax, deg[v]);
    for (uint64_t k = rp[v]; k < rp[v + 1]; ++k) {
      const uint32_t u = ci[k];
      if (deg[u] > deg[v]) {
        const uint32_t du = deg[u], pu = pos[u], pw = bin[du], w = vert[pw];
        if (u != w) {
          pos[u] = pw;
          vert[pu] = w;
          pos[w] = pu;
          vert[pw] = u;
        }
        bin[du]++;
        deg[u]--;
      }
    }
  }
  for (uint32_t v = 0; v < n; ++v) core[v] = deg[v];
  return kmax;
}

void oracle_orient(uint32_t n, const uint64_t *rp, const uint32_t *ci, const uint32_t *prio,
                   uint64_t *split, uint32_t *oc) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t u = 0; u < n; ++u) {
    uint64_t w = rp[u];
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      const uint32_t x = ci[k];
      const bool later = prio[u] < prio[x] || (prio[u] == prio[x] && u < x);
      if (!later) oc[w++] = x;
    }
    split[u] = w;
    for (uint64_t k = rp[u]; k < rp[u + 1]; ++k) {
      const uint32_t x = ci[k];
      const bool later = prio[u] < prio[x] || (prio[u] == prio[x] && u < x);
      if (later) oc[w++] = x;
    }
  }
}

// Enumerates maximal cliques. level: 1 (L1) or 2 (L2); ipx: 0 (IP) or 1 (IPX).
// order (nullable) gives the first-level vertices to run, in that order (L1
// and L2 alike); otherwise all vertices. sub_nodes / sub_cliques (nullable)
// receive per-first-level-vertex tree nodes and cliques (L1 only).
// threads <= 0 uses all hardware threads. Returns 0, or -1 on bad input.
int oracle_mce(uint32_t n, const uint64_t *rp, const uint32_t *ci, int level, int ipx,
               const uint32_t *order, uint64_t order_len, int threads, int collect_work,
               uint64_t *sub_nodes, uint64_t *sub_cliques, oracle_result *res) {
  if (!res || (n && (!rp || !ci)) || (level != 1 && level != 2)) return -1;
  memset(res, 0, sizeof(*res));
  if (threads <= 0) threads = omp_get_max_threads();
  res->threads = (uint32_t)threads;
  Graph g;
  g.n = n;
  g.rp = rp;
  g.ci = ci;
  const double t0 = now_ms();
  std::vector<uint32_t> prio(n);
  res->degeneracy = n ? oracle_peel(n, rp, ci, prio.data(), nullptr) : 0;
  g.split.resize(n);
  g.oc.resize(n ? rp[n] : 0);
  if (n) oracle_orient(n, rp, ci, prio.data(), g.split.data(), g.oc.data());
  const double t1 = now_ms();
  uint32_t md = 0, mo = 0;
  for (uint32_t v = 0; v < n; ++v) {
    md = std::max<uint32_t>(md, (uint32_t)(rp[v + 1] - rp[v]));
    mo = std::max<uint32_t>(mo, (uint32_t)(rp[v + 1] - g.split[v]));
  }
  res->max_degree = md;
  res->max_out_degree = mo;
  res->t_order_ms = t1 - t0;
  if (mo > 4096) return -2;
  std::vector<uint32_t> all;
  if (!order) {
    all.resize(n);
    for (uint32_t v = 0; v < n; ++v) all[v] = v;
    order = all.data();
    order_len = n;
  }
  std::vector<Acc> accs(threads);
  double bh = 0;
#pragma omp parallel num_threads(threads) reduction(+ : bh)
  {
    Acc &acc = accs[omp_get_thread_num()];
    Solver s;
    s.g = &g;
    s.ipx = ipx != 0;
    s.work = collect_work != 0;
    s.acc = &acc;
    std::vector<uint32_t> xs, tmp;
#pragma omp for schedule(dynamic, 1)
    for (uint64_t i = 0; i < order_len; ++i) {
      const uint32_t v = order[i];
      const uint64_t *rp_ = g.rp;
      const uint32_t *np = g.oc.data() + g.split[v], *ne = g.oc.data() + rp_[v + 1];
      uint64_t nodes = 0, cl = 0;
      if (level == 1) {
        if (np == ne) continue;
        s.S.assign(np, ne);
        xs.assign(g.oc.data() + rp_[v], g.oc.data() + g.split[v]);
        s.run(xs, 1, splitmix64(v));
        nodes = s.nodes;
        cl = s.cliques;
        if (s.work) {
          double b = 4.0 * (double)(rp_[v + 1] - rp_[v]);
          for (const uint32_t *u = np; u < ne; ++u) b += 4.0 * (double)(rp_[*u + 1] - g.split[*u]);
          if (ipx)
            for (uint32_t x : xs) b += 4.0 * (double)(rp_[x + 1] - g.split[x]);
          bh += b;
        }
      } else {
        for (const uint32_t *wp = np; wp < ne; ++wp) {
          const uint32_t w = *wp;
          intersect(np, ne, g.oc.data() + g.split[w], g.oc.data() + rp_[w + 1], s.S);
          intersect(g.ci + rp_[v], g.ci + rp_[v + 1], g.oc.data() + rp_[w], g.oc.data() + g.split[w], xs);
          s.run(xs, 2, splitmix64(v) + splitmix64(w));
          nodes += s.nodes;
          cl += s.cliques;
        }
      }
      acc.nodes += nodes;
      if (sub_nodes) sub_nodes[v] = nodes;
      if (sub_cliques) sub_cliques[v] = cl;
    }
  }
  const double t2 = now_ms();
  res->t_count_ms = t2 - t1;
  for (auto &a : accs) {
    res->count += a.count;
    res->hash += a.hash;
    res->nodes += a.nodes;
    res->ops += a.ops;
    res->b_lds += a.lds;
    for (int k = 0; k < kHist; ++k) res->hist[k] += a.hist[k];
  }
  res->b_hbm = bh + 4.0 * (4.0 * n + 3.0 * (double)(n ? rp[n] : 0));
  for (int k = kHist - 1; k >= 0; --k)
    if (res->hist[k]) {
      res->max_clique_size = (uint32_t)k;
      break;
    }
  return 0;
}

// Oriented graph for tests: split[n], oc[m] and prio[n] as the oracle builds them.
uint32_t oracle_order_orient(uint32_t n, const uint64_t *rp, const uint32_t *ci, uint32_t *prio,
                             uint64_t *split, uint32_t *oc, uint32_t *rounds) {
  const uint32_t d = oracle_peel(n, rp, ci, prio, rounds);
  oracle_orient(n, rp, ci, prio, split, oc);
  return d;
}

int oracle_hist_len(void) { return kHist; }

// Knuth estimates of the level-1 subtrees of the vertices in `order`: the
// mean over `probes` random probes of each vertex's maximal cliques and tree
// nodes (Solver::probe), written to est_cliques / est_nodes per entry. The
// probe random stream of (vertex v, probe j) is splitmix64 of seed, v and j,
// so results do not depend on the thread count. Returns 0, -1 on bad input,
// -2 when an out-degree exceeds 4096.
int oracle_knuth(uint32_t n, const uint64_t *rp, const uint32_t *ci, int ipx, const uint32_t *order,
                 uint64_t order_len, uint32_t probes, uint64_t seed, int threads, double *est_cliques,
                 double *est_nodes) {
  if ((n && (!rp || !ci)) || (order_len && (!order || !est_cliques || !est_nodes)) || probes == 0) return -1;
  if (threads <= 0) threads = omp_get_max_threads();
  Graph g;
  g.n = n;
  g.rp = rp;
  g.ci = ci;
  std::vector<uint32_t> prio(n);
  if (n) oracle_peel(n, rp, ci, prio.data(), nullptr);
  g.split.resize(n);
  g.oc.resize(n ? rp[n] : 0);
  if (n) oracle_orient(n, rp, ci, prio.data(), g.split.data(), g.oc.data());
  for (uint32_t v = 0; v < n; ++v)
    if (rp[v + 1] - g.split[v] > 4096) return -2;
  for (uint64_t i = 0; i < order_len; ++i)
    if (order[i] >= n) return -1;
#pragma omp parallel num_threads(threads)
  {
    Acc acc;
    Solver s;
    s.g = &g;
    s.ipx = ipx != 0;
    s.work = false;
    s.acc = &acc;
    std::vector<uint32_t> xs;
#pragma omp for schedule(dynamic, 1)
    for (uint64_t i = 0; i < order_len; ++i) {
      const uint32_t v = order[i];
      s.S.assign(g.oc.data() + g.split[v], g.oc.data() + rp[v + 1]);
      xs.assign(g.oc.data() + rp[v], g.oc.data() + g.split[v]);
      double sc = 0, sn = 0;
      if (s.start(xs)) {
        for (uint32_t j = 0; j < probes; ++j) {
          uint64_t rng = splitmix64(seed ^ splitmix64(((uint64_t)v << 32) | j));
          double c, nd;
          s.probe(rng, c, nd);
          sc += c;
          sn += nd;
        }
      }
      est_cliques[i] = sc / probes;
      est_nodes[i] = sn / probes;
    }
  }
  return 0;
}

}  // extern "C"
