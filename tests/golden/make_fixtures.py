"""Writes the golden maximal-clique fixtures used by the CPU tests.

The expected values come from networkx.find_cliques (an independent library
implementation, networkx 3.4.2), not from this repository's oracle:
  - count, size histogram and order-independent hash of the maximal cliques
    with at least 2 vertices (size-1 cliques are isolated vertices, which the
    reference never counts, mce/mce_kernel_nowl.cuh:61-69);
  - the maximum core number (networkx.core_number).
Graphs are small seeded networkx graphs stored as edge lists, plus the
library-generated Erdos-Renyi n=4096, m=65,536, seed 1 graph, pinned by its
CSR digest and its networkx clique statistics.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import networkx as nx

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


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


def clique_hash(c):
    return fmix64((sum(splitmix64(v) for v in c) + len(c) * GOLDEN) & M64)


def stats(G):
    count, h, hist = 0, 0, {}
    for c in nx.find_cliques(G):
        if len(c) < 2:
            continue
        count += 1
        h = (h + clique_hash(c)) & M64
        hist[len(c)] = hist.get(len(c), 0) + 1
    core = max(nx.core_number(G).values()) if G.number_of_nodes() else 0
    return dict(count=count, hash=h, hist={str(k): v for k, v in sorted(hist.items())}, degeneracy=core)


def graphs():
    yield "gnp_60_0.3", nx.gnp_random_graph(60, 0.3, seed=1)
    yield "gnp_120_0.1", nx.gnp_random_graph(120, 0.1, seed=2)
    yield "gnp_40_0.7", nx.gnp_random_graph(40, 0.7, seed=3)
    yield "ba_300_4", nx.barabasi_albert_graph(300, 4, seed=4)
    yield "powerlaw_cluster_250", nx.powerlaw_cluster_graph(250, 5, 0.6, seed=5)
    yield "watts_strogatz_200", nx.connected_watts_strogatz_graph(200, 10, 0.2, seed=6)
    yield "karate", nx.karate_club_graph()
    yield "les_miserables", nx.convert_node_labels_to_integers(nx.les_miserables_graph())
    yield "florentine", nx.convert_node_labels_to_integers(nx.florentine_families_graph())
    yield "random_regular_100_9", nx.random_regular_graph(9, 100, seed=7)
    yield "sbm_4x50", nx.stochastic_block_model([50, 50, 50, 50], [[0.3 if i == j else 0.01 for j in range(4)]
                                                                   for i in range(4)], seed=8)
    yield "caveman_8x6", nx.connected_caveman_graph(8, 6)


def main():
    out = {"generator": "networkx " + nx.__version__, "graphs": []}
    for name, G in graphs():
        n = G.number_of_nodes()
        edges = sorted((min(u, v), max(u, v)) for u, v in G.edges() if u != v)
        rec = dict(name=name, n=n, edges=edges)
        rec.update(stats(G))
        out["graphs"].append(rec)
        print(name, n, len(edges), rec["count"], flush=True)
    # the library's ER n=4096 m=65536 seed 1 graph
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import load_engine

    mce = load_engine()
    g = mce.gen_er(4096, 65536, 1)
    G = nx.Graph()
    G.add_nodes_from(range(g.n))
    for u in range(g.n):
        for w in g.col_ind[g.row_ptr[u]:g.row_ptr[u + 1]]:
            if u < w:
                G.add_edge(u, int(w))
    rec = dict(name="er_4096_65536_seed1", n=g.n, m=g.m, digest=g.digest())
    rec.update(stats(G))
    out["er_config"] = rec
    print("er", rec["count"], hex(rec["digest"]))
    with open(os.path.join(HERE, "nx_cliques.json"), "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))


if __name__ == "__main__":
    main()
