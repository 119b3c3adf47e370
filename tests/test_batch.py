"""Batched engine (one captured launch over a packed batch of graphs) and the
block-settings autotuner: per-graph counts against the oracle run on each
member graph alone, argument checks, and result invariance under tuning."""
import numpy as np
import pytest


def member(g, off, i):
    """Graph i of a packed batch, renumbered from 0."""
    a, b = int(off[i]), int(off[i + 1])
    rp = g.row_ptr[a:b + 1] - g.row_ptr[a]
    ci = g.col_ind[int(g.row_ptr[a]):int(g.row_ptr[b])].astype(np.int64) - a
    return type(g)(rp, ci.astype(np.uint32))


def test_sbm_batch_shape(mce):
    off, g = mce.gen_sbm_batch(16, n_min=32, n_max=200, blocks=4, p_in=0.3, p_out=0.02, seed=3)
    assert off[0] == 0 and off[-1] == g.n and np.all(np.diff(off) >= 32) and np.all(np.diff(off) <= 200)
    g.validate()
    for i in range(16):  # arcs never cross graphs
        a, b = int(off[i]), int(off[i + 1])
        cols = g.col_ind[int(g.row_ptr[a]):int(g.row_ptr[b])]
        assert cols.size == 0 or (cols.min() >= a and cols.max() < b)


def test_batch_rejects_bad_input(mce):
    off, g = mce.gen_sbm_batch(4, n_min=32, n_max=64, seed=1)
    bad = off.copy()
    bad[-1] -= 1
    with pytest.raises(mce.MCEError) as e:
        mce.run_batch(bad, g)
    assert e.value.code == -1
    # an arc between two member graphs
    a0, a1 = int(off[1]) - 1, int(off[1])
    src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(g.row_ptr).astype(np.int64))
    s = np.concatenate([src, [a0, a1]]).astype(np.uint32)
    d = np.concatenate([g.col_ind, [a1, a0]]).astype(np.uint32)
    crossed = mce.from_edges(s, d, g.n)
    with pytest.raises(mce.MCEError) as e:
        mce.run_batch(off, crossed)
    assert e.value.code == -6


@pytest.mark.gpu
@pytest.mark.parametrize("wl", [1, 2])
def test_batch_counts_per_graph(mce, oracle, wl):
    off, g = mce.gen_sbm_batch(96, n_min=64, n_max=600, blocks=4, p_in=0.35, p_out=0.02, seed=7)
    counts, r = mce.run_batch(off, g, level=1, induced=2, worklist=wl)
    ref = oracle.mce(g, level=1, ipx=True, sub_cliques=True)
    assert (r["count"], r["hash"], r["hist"], r["tree_nodes"]) == (
        ref["count"], ref["hash"], ref["hist"], ref["tree_nodes"])
    per = np.add.reduceat(ref["sub_cliques"], off[:-1].astype(np.int64)) if g.n else np.zeros(0)
    # graphs with no vertex would break reduceat; the generator makes none
    assert np.array_equal(counts, per.astype(np.uint64))
    for i in (0, 17, 95):  # member graphs enumerated alone
        assert int(counts[i]) == oracle.mce(member(g, off, i), level=1, ipx=True)["count"]


@pytest.mark.gpu
def test_batch_l2_and_ip(mce, oracle):
    off, g = mce.gen_sbm_batch(40, n_min=64, n_max=400, blocks=3, p_in=0.4, p_out=0.03, seed=11)
    ref = oracle.mce(g, level=1, ipx=True, sub_cliques=True)
    per = np.add.reduceat(ref["sub_cliques"], off[:-1].astype(np.int64))
    c1, r1 = mce.run_batch(off, g, level=1, induced=1)
    assert np.array_equal(c1, per.astype(np.uint64)) and r1["hash"] == ref["hash"]
    # second level: per-root counts are attributed to the arc's tail, in the same graph
    c2, r2 = mce.run_batch(off, g, level=2, induced=2)
    assert r2["count"] == ref["count"] and r2["hash"] == ref["hash"]
    assert np.array_equal(c2, per.astype(np.uint64))


@pytest.mark.gpu
def test_autotune_is_result_invariant(mce, oracle):
    g = mce.gen_rmat(12, 16, seed=5)
    ref = oracle.mce(g, level=1, ipx=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        best, trials = dg.autotune(sample_fraction=0.25, worklist=2)
        assert len(trials) >= 8 and all(t["ms"] >= 0 for t in trials)
        fastest = min(trials, key=lambda t: t["ms"])
        for k in ("waves_per_group", "groups_per_cu", "stack_lds_levels"):
            assert best[k] == fastest[k]
        r = dg.count(level=1, induced=2, **best)
        assert (r["count"], r["hash"], r["hist"]) == (ref["count"], ref["hash"], ref["hist"])


@pytest.mark.gpu
@pytest.mark.parametrize("wl", [1, 2])
def test_graph_launch_on_resident_graph(mce, oracle, wl):
    """The captured (hipGraph) enumeration launch on a resident batch. The graph
    handle keeps the instantiated graph and replays it while the launch
    parameters are unchanged: four replays, with ordinary launches, a launch
    with other settings (a new capture) and host synchronisations in between,
    each checked against the oracle."""
    off, g = mce.gen_sbm_batch(24, n_min=64, n_max=500, blocks=4, p_in=0.3, p_out=0.02, seed=5)
    ref = oracle.mce(g, level=1, ipx=True, sub_cliques=True)
    ref_ip = oracle.mce(g, level=1, ipx=False)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        for k in range(4):
            r = dg.count(level=1, induced=2, worklist=wl, graph_launch=1, sub_cliques=True)
            assert (r["count"], r["hash"], r["hist"], r["tree_nodes"]) == (
                ref["count"], ref["hash"], ref["hist"], ref["tree_nodes"]), k
            assert np.array_equal(r["sub_cliques"], ref["sub_cliques"]), k
            r = dg.count(level=1, induced=2, worklist=wl)
            assert (r["count"], r["hash"]) == (ref["count"], ref["hash"])
            if k == 1:
                r = dg.count(level=1, induced=1, worklist=wl, graph_launch=1)
                assert (r["count"], r["hash"], r["tree_nodes"]) == (ref_ip["count"], ref_ip["hash"],
                                                                    ref_ip["tree_nodes"])
