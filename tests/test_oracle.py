"""The CPU oracle pinned by what the mathematics fixes and by an independent
library: closed forms, brute force over all vertex subsets, networkx
find_cliques fixtures, and the k-core / orientation invariants."""
import itertools
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
VARIANTS = [(1, False), (1, True), (2, False), (2, True)]


def splitmix64(x):
    x = (x + GOLDEN) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def fmix64(k):
    k ^= k >> 33
    k = (k * 0xFF51AFD7ED558CCD) & M64
    k ^= k >> 33
    k = (k * 0xC4CEB9FE1A85EC53) & M64
    k ^= k >> 33
    return k


def clique_hash(cliques):
    h = 0
    for c in cliques:
        h = (h + fmix64((sum(splitmix64(v) for v in c) + len(c) * GOLDEN) & M64)) & M64
    return h


def graph(mce, n, edges):
    if not edges:
        return mce.from_edges(np.zeros(0, np.uint32), np.zeros(0, np.uint32), n=n)
    e = np.array(edges, dtype=np.uint32)
    return mce.from_edges(e[:, 0], e[:, 1], n=n)


def all_variants(oracle, g):
    res = [oracle.mce(g, level=lvl, ipx=ipx, threads=2) for lvl, ipx in VARIANTS]
    for r in res[1:]:
        assert (r["count"], r["hash"], r["hist"]) == (res[0]["count"], res[0]["hash"], res[0]["hist"])
    assert sum(res[0]["hist"].values()) == res[0]["count"]
    return res[0]


def expect(oracle, g, cliques):
    """Exact clique set given in closed form."""
    r = all_variants(oracle, g)
    hist = {}
    for c in cliques:
        hist[len(c)] = hist.get(len(c), 0) + 1
    assert r["count"] == len(cliques)
    assert r["hist"] == hist
    assert r["hash"] == clique_hash(cliques)
    return r


# ---- closed forms ---------------------------------------------------------

def test_complete_graph(mce, oracle):
    for n in (2, 3, 7, 64, 65, 130):
        r = expect(oracle, graph(mce, n, list(itertools.combinations(range(n), 2))), [tuple(range(n))])
        assert r["degeneracy"] == n - 1


def test_empty_and_isolated(mce, oracle):
    expect(oracle, graph(mce, 10, []), [])
    # isolated vertices and id gaps are never cliques
    expect(oracle, graph(mce, 20, [(3, 9)]), [(3, 9)])


def test_matching_cycle_star(mce, oracle):
    k = 12
    expect(oracle, graph(mce, 2 * k, [(2 * i, 2 * i + 1) for i in range(k)]), [(2 * i, 2 * i + 1) for i in range(k)])
    for n in (4, 5, 17):
        edges = [(i, (i + 1) % n) for i in range(n)]
        expect(oracle, graph(mce, n, edges), [tuple(sorted(e)) for e in edges])
    expect(oracle, graph(mce, 3, [(0, 1), (1, 2), (0, 2)]), [(0, 1, 2)])
    s = 40
    expect(oracle, graph(mce, s + 1, [(0, i) for i in range(1, s + 1)]), [(0, i) for i in range(1, s + 1)])


def test_complete_multipartite_and_turan(mce, oracle):
    for sizes in ((2, 3, 4), (1, 1, 5, 2), (3, 3, 3, 3), (7, 9)):
        parts, v = [], 0
        for sz in sizes:
            parts.append(list(range(v, v + sz)))
            v += sz
        edges = [(a, b) for i, j in itertools.combinations(range(len(parts)), 2) for a in parts[i] for b in parts[j]]
        cl = [tuple(sorted(c)) for c in itertools.product(*parts)]
        expect(oracle, graph(mce, v, edges), cl)


def test_moon_moser(mce, oracle):
    # complement of t disjoint triangles: 3^t maximal cliques of size t
    for t in (2, 4, 6):
        n = 3 * t
        tri = {(3 * i + a, 3 * i + b) for i in range(t) for a, b in ((0, 1), (0, 2), (1, 2))}
        edges = [e for e in itertools.combinations(range(n), 2) if e not in tri]
        cl = [tuple(sorted(c)) for c in itertools.product(*[(3 * i, 3 * i + 1, 3 * i + 2) for i in range(t)])]
        expect(oracle, graph(mce, n, edges), cl)


def test_disjoint_cliques_wheel_petersen(mce, oracle):
    sizes, edges, cl, v = (2, 5, 9, 3), [], [], 0
    for sz in sizes:
        block = list(range(v, v + sz))
        edges += list(itertools.combinations(block, 2))
        cl.append(tuple(block))
        v += sz
    expect(oracle, graph(mce, v, edges), cl)
    n = 9  # wheel: hub 0 + cycle 1..n
    rim = [(1 + i, 1 + (i + 1) % n) for i in range(n)]
    wheel = rim + [(0, i) for i in range(1, n + 1)]
    expect(oracle, graph(mce, n + 1, wheel), [tuple(sorted((0,) + e)) for e in rim])
    outer = [(i, (i + 1) % 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    spokes = [(i, i + 5) for i in range(5)]
    pet = outer + inner + spokes
    expect(oracle, graph(mce, 10, pet), [tuple(sorted(e)) for e in pet])


# ---- brute force ----------------------------------------------------------

def brute_force(n, edges):
    adj = [0] * n
    for a, b in edges:
        adj[a] |= 1 << b
        adj[b] |= 1 << a
    cliques = []
    for mask in range(1, 1 << n):
        vs = [v for v in range(n) if mask >> v & 1]
        if len(vs) < 2 or any((adj[v] | (1 << v)) & mask != mask for v in vs):
            continue
        outside = ~mask & ((1 << n) - 1)
        if any(all(adj[u] >> v & 1 for v in vs) for u in range(n) if outside >> u & 1):
            continue
        cliques.append(tuple(vs))
    return cliques


def test_brute_force_small(mce, oracle):
    rng = np.random.default_rng(2024)
    for t in range(80):
        n = int(rng.integers(1, 13))
        p = float(rng.random())
        edges = [e for e in itertools.combinations(range(n), 2) if rng.random() < p]
        expect(oracle, graph(mce, n, edges), brute_force(n, edges))


# ---- networkx fixtures ----------------------------------------------------

def load_fixtures():
    with open(os.path.join(HERE, "golden", "nx_cliques.json")) as f:
        return json.load(f)


def test_networkx_fixtures(mce, oracle):
    fx = load_fixtures()
    for rec in fx["graphs"]:
        g = graph(mce, rec["n"], [tuple(e) for e in rec["edges"]])
        r = all_variants(oracle, g)
        assert r["count"] == rec["count"], rec["name"]
        assert r["hash"] == rec["hash"], rec["name"]
        assert {str(k): v for k, v in r["hist"].items()} == rec["hist"], rec["name"]
        assert r["degeneracy"] == rec["degeneracy"], rec["name"]


def test_er_config_fixture(mce, oracle):
    """ER n=4096, m=65,536, seed 1: generator digest and clique statistics."""
    rec = load_fixtures()["er_config"]
    g = mce.gen_er(4096, 65536, 1)
    assert (g.n, g.m) == (rec["n"], rec["m"])
    assert g.digest() == rec["digest"]
    r = all_variants(oracle, g)
    assert r["count"] == rec["count"] and r["hash"] == rec["hash"]
    assert {str(k): v for k, v in r["hist"].items()} == rec["hist"]
    assert r["degeneracy"] == rec["degeneracy"]


# ---- k-core and orientation ------------------------------------------------

def test_degeneracy_matches_core_numbers(mce, oracle):
    for seed in range(6):
        g = mce.gen_rmat(11, 8 + 4 * seed, seed=seed)
        prio, d, rounds = oracle.peel(g)
        core, kmax = oracle.core_numbers(g)
        assert d == kmax == core.max()
        assert rounds >= 1


def test_orientation_invariants(mce, oracle):
    g = mce.gen_rmat(12, 16, seed=7)
    prio, split, oc, d = oracle.order_orient(g)
    rp = g.row_ptr
    dplus = rp[1:] - split
    assert dplus.max() <= d
    for u in range(0, g.n, 37):
        row = oc[rp[u]:rp[u + 1]]
        early, late = row[: split[u] - rp[u]], row[split[u] - rp[u]:]
        assert np.all(np.diff(early.astype(np.int64)) > 0) and np.all(np.diff(late.astype(np.int64)) > 0)
        key_u = (int(prio[u]), u)
        assert all((int(prio[w]), int(w)) < key_u for w in early)
        assert all((int(prio[w]), int(w)) > key_u for w in late)
        assert sorted(row.tolist()) == g.col_ind[rp[u]:rp[u + 1]].tolist()


def test_peel_rounds_are_batches(mce, oracle):
    """Each round removes only vertices of remaining degree <= its level, and
    the first round is every vertex of minimum degree."""
    g = mce.gen_rmat(10, 16, seed=2)
    prio, d, rounds = oracle.peel(g)
    deg = np.diff(g.row_ptr).astype(np.int64)
    first = np.nonzero(prio == 0)[0]
    assert set(first.tolist()) == set(np.nonzero(deg == deg.min())[0].tolist())
    # a round belongs to one level of the peel, and that level is the core
    # number of every vertex it removes (Batagelj-Zaversnik, computed
    # independently): each vertex leaves with at most that many neighbours
    core, k = oracle.core_numbers(g)
    assert k == d
    removed = np.zeros(g.n, bool)
    levels = []
    for r in range(rounds):
        batch = np.nonzero(prio == r)[0]
        if batch.size == 0:
            continue
        level = set(core[batch].tolist())
        assert len(level) == 1, (r, level)
        level = level.pop()
        rem = [int(np.sum(~removed[g.col_ind[g.row_ptr[v]:g.row_ptr[v + 1]]])) for v in batch]
        assert max(rem) <= level, (r, level, max(rem))
        levels.append(level)
        removed[batch] = True
    assert removed.all()
    assert levels == sorted(levels) and levels[-1] == d


def test_schedule_invariance_and_labels(mce, oracle):
    g = mce.gen_rmat(11, 16, seed=4)
    base = oracle.mce(g, level=1, ipx=True, sub_nodes=True, sub_cliques=True)
    assert int(base["sub_cliques"].sum()) == base["count"]
    assert int(base["sub_nodes"].sum()) == base["tree_nodes"]
    perm = np.random.default_rng(1).permutation(g.n).astype(np.uint32)
    r = oracle.mce(g, level=1, ipx=True, order=perm, threads=3)
    assert (r["count"], r["hash"], r["hist"], r["tree_nodes"]) == (
        base["count"], base["hash"], base["hist"], base["tree_nodes"])


def test_knuth_estimator(mce, oracle):
    """Knuth's random-probe estimate of a subtree (used to size the R-MAT
    workloads) is unbiased: where the pivot leaves one branch per node it is
    exact, and over many probes it converges to the exact per-vertex clique
    and node counts of the enumeration."""
    # K_8: vertex 0's subproblem is one path down to the single clique
    g = graph(mce, 8, list(itertools.combinations(range(8), 2)))
    ec, en = oracle.knuth(g, np.arange(8, dtype=np.uint32), probes=16)
    assert ec.sum() == 1.0
    # Moon-Moser t = 4: 81 cliques, every vertex subproblem sampled
    t = 4
    tri = {(3 * i + a, 3 * i + b) for i in range(t) for a, b in ((0, 1), (0, 2), (1, 2))}
    g = graph(mce, 3 * t, [e for e in itertools.combinations(range(3 * t), 2) if e not in tri])
    ec, en = oracle.knuth(g, np.arange(3 * t, dtype=np.uint32), probes=20000, seed=3)
    ref = oracle.mce(g, level=1, ipx=True, sub_nodes=True, sub_cliques=True)
    assert abs(ec.sum() - 81) < 0.03 * 81
    # the heaviest R-MAT subtrees, vertex by vertex
    g = mce.gen_rmat(12, 16, seed=3)
    ref = oracle.mce(g, level=1, ipx=True, sub_nodes=True, sub_cliques=True)
    top = np.argsort(-ref["sub_nodes"].astype(np.float64))[:4].astype(np.uint32)
    for ipx in (True, False):
        ec, en = oracle.knuth(g, top, probes=50000, seed=5, ipx=ipx)
        want = oracle.mce(g, level=1, ipx=ipx, sub_nodes=True, sub_cliques=True)
        assert np.all(np.abs(ec - want["sub_cliques"][top]) < 0.05 * want["sub_cliques"][top]), (ipx, ec)
        assert np.all(np.abs(en - want["sub_nodes"][top]) < 0.05 * want["sub_nodes"][top]), (ipx, en)
    # all vertices of the Erdos-Renyi configuration: the total
    g = mce.gen_er(4096, 65536, 1)
    ec, en = oracle.knuth(g, np.arange(g.n, dtype=np.uint32), probes=500, seed=7)
    assert abs(ec.sum() - 56449) < 0.01 * 56449
