"""GPU engine vs CPU oracle: bit-exact count, size histogram, clique hash and
BK tree-node count, through the C ABI, for every -p/-i variant."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VARIANTS = [(lvl, ind) for lvl in (1, 2) for ind in (1, 2)]  # (L1|L2, IP|IPX)


def rand_graph(mce, n, p, seed):
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(iu[0].size) < p
    return mce.from_edges(iu[0][keep], iu[1][keep], n=n)


def assert_same(res, ref, what, nodes=True):
    assert res["count"] == ref["count"], (what, res["count"], ref["count"])
    assert res["hist"] == ref["hist"], (what, res["hist"], ref["hist"])
    assert res["hash"] == ref["hash"], what
    if nodes:
        assert res["tree_nodes"] == ref["tree_nodes"], (what, res["tree_nodes"], ref["tree_nodes"])


def run_all(mce, oracle, g, what, nodes=True, **block):
    with mce.DeviceGraph(g) as dg:
        info = dg.order()
        for lvl, ind in VARIANTS:
            ref = oracle.mce(g, level=lvl, ipx=(ind == 2))
            res = dg.count(level=lvl, induced=ind, **block)
            assert info["degeneracy"] == ref["degeneracy"], what
            # maximum degree and out-degree after orientation (Delta, Delta+)
            assert (info["max_degree"], info["max_out_degree"]) == (ref["max_degree"], ref["max_out_degree"]), what
            assert_same(res, ref, (what, lvl, ind), nodes=nodes)


def test_random_small_graphs(mce, oracle):
    for seed in range(60):
        rng = np.random.default_rng(1000 + seed)
        n = int(rng.integers(1, 90))
        p = float(rng.random())
        print(f"case seed={seed} n={n} p={p:.3f}", flush=True)
        run_all(mce, oracle, rand_graph(mce, n, p, seed), f"rand-{seed}")


def test_ordering_and_orientation_match_oracle(mce, oracle):
    for seed in range(8):
        g = rand_graph(mce, 300, 0.05 + 0.05 * seed, seed)
        prio_o, split_o, oc_o, d_o = oracle.order_orient(g)
        with mce.DeviceGraph(g) as dg:
            info = dg.order()
            prio, split, oc = dg.download()
        assert info["degeneracy"] == d_o
        np.testing.assert_array_equal(prio, prio_o)
        np.testing.assert_array_equal(split, split_o)
        np.testing.assert_array_equal(oc, oc_o)


def test_er_4096_config(mce, oracle):
    """Erdős–Rényi n=4096, m=65,536, seed 1 (the small plain-method config)."""
    g = mce.gen_er(4096, 65536, 1)
    run_all(mce, oracle, g, "er4096")


def test_rmat_scale14(mce, oracle):
    """R-MAT scale 14 (127,583,041 maximal cliques at seed 3): level 1 with and
    without the worker list, level 2, against the oracle."""
    g = mce.gen_rmat(14, 16, seed=3)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        for lvl, wl in ((1, 1), (1, 2), (2, 2)):
            ref = oracle.mce(g, level=lvl, ipx=True)
            print(f"case rmat14 level={lvl} worklist={wl}", flush=True)
            res = dg.count(level=lvl, induced=2, worklist=wl)
            assert_same(res, ref, ("rmat14", lvl, wl))


@pytest.mark.parametrize("wl", [1, 2])
def test_deferred_items_second_pass(mce, oracle, wl):
    """A small LDS pool and a 4 KiB HBM arena cap defer the larger subproblems
    to the second launch with arenas sized for the largest; results and tree
    nodes stay those of the oracle in every variant."""
    g = rand_graph(mce, 400, 0.3, 21)
    run_all(mce, oracle, g, f"deferred-wl{wl}", worklist=wl, lds_arena_bytes=1024, hbm_arena_bytes=4096)


def test_rmat_scale12(mce, oracle):
    # R-MAT (.57, .19, .19, .05), edge factor 16: 0.57M maximal cliques at scale 12
    g = mce.gen_rmat(12, 16, seed=3)
    run_all(mce, oracle, g, "rmat12")


def test_multiword_bitmaps(mce, oracle):
    # out-degrees above 64, 128 and 192 exercise W = 2, 3 and 4 words
    # (degeneracy 85, 172 and 208 at seed 1; 0.5M-1.5M maximal cliques each)
    for seed, (n, p) in enumerate([(200, 0.5), (1000, 0.2), (2000, 0.12)]):
        print(f"case dense n={n} p={p}", flush=True)
        run_all(mce, oracle, rand_graph(mce, n, p, 1), f"dense-{n}")


@pytest.mark.parametrize("small", [-1, 8, 33])
def test_small_universe_thresholds(mce, oracle, small):
    # nodes with |P| + |X| + outside rows <= small run one lane per member;
    # -1 keeps every node on the word-per-lane solver, 8 and 33 switch deep
    # in the tree (and 33 at the root of the sparser graph)
    for n, p, seed in [(200, 0.5, 1), (120, 0.3, 2)]:
        print(f"case small={small} n={n} p={p}", flush=True)
        run_all(mce, oracle, rand_graph(mce, n, p, seed), f"small{small}-{n}", small_universe=small)
    print(f"case small={small} hbm arena", flush=True)
    run_all(mce, oracle, rand_graph(mce, 220, 0.55, 7), f"small{small}-hbm", small_universe=small,
            lds_arena_bytes=1024)


@pytest.mark.parametrize("lane", [-1, 4, 16])
def test_lane_universe_thresholds(mce, oracle, lane):
    # subtrees with |P| + |X| <= lane run one per lane (-1: never; 4: only
    # the smallest; 16: every one the small-universe solver meets), with and
    # without the worker list, with per-subproblem labels: a Chung-Lu graph
    # with planted cliques queues thousands of subtrees per heavy subproblem
    # (full queues, runs in the middle of a subtree), the dense random graph
    # few large ones, and a tiny LDS arena moves the bit matrices to HBM
    g = mce.gen_chung_lu(1 << 14, 2.1, 24.0, 8, 20, 40, seed=3)
    ref = oracle.mce(g, level=1, ipx=True, sub_nodes=True, sub_cliques=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        for wl in (1, 2):
            print(f"case lane={lane} chung-lu 2^14 wl={wl}", flush=True)
            res = dg.count(level=1, induced=2, worklist=wl, lane_universe=lane, sub_nodes=True, sub_cliques=True)
            assert_same(res, ref, (lane, wl))
            np.testing.assert_array_equal(res["sub_nodes"], ref["sub_nodes"])
            np.testing.assert_array_equal(res["sub_cliques"], ref["sub_cliques"])
    for n, p, seed in [(200, 0.5, 1), (120, 0.3, 2)]:
        print(f"case lane={lane} n={n} p={p}", flush=True)
        run_all(mce, oracle, rand_graph(mce, n, p, seed), f"lane{lane}-{n}", lane_universe=lane)
    print(f"case lane={lane} hbm arena", flush=True)
    run_all(mce, oracle, rand_graph(mce, 220, 0.55, 7), f"lane{lane}-hbm", lane_universe=lane,
            lds_arena_bytes=1024, worklist=2)


def hub_graph(mce, live, dead, live2=0, two_first=False, core=60, p_core=0.9, hub_core=40, seed=4):
    """A hub with `live` + `dead` + `live2` leaf neighbours and `hub_core`
    neighbours in a dense core of `core` vertices. Leaves and the hub are
    peeled before the core, so the hub's subproblem has P = its core
    neighbours and X = all its leaves; a live leaf also touches one of those
    core neighbours (its outside row has one bit: many repeat), a `live2` leaf
    two of them, a dead leaf only the hub. Outside rows are in vertex order:
    the `live2` leaves come last, or first with two_first."""
    rng = np.random.default_rng(seed)
    hub = 0
    cs = np.arange(1, core + 1)
    iu = np.triu_indices(core, 1)
    keep = rng.random(iu[0].size) < p_core
    src = [cs[iu[0][keep]], np.full(hub_core, hub)]
    dst = [cs[iu[1][keep]], cs[:hub_core]]
    first = core + 1 + (live2 if two_first else 0)
    leaves = np.arange(first, first + live + dead)
    src.append(np.full(leaves.size, hub))
    dst.append(leaves)
    src.append(leaves[:live])
    dst.append(rng.choice(cs[:hub_core], live))
    t0 = core + 1 if two_first else core + 1 + live + dead
    two = np.arange(t0, t0 + live2)
    src += [np.full(two.size, hub), two, two]
    a = rng.integers(0, hub_core, live2)
    b = (a + 1 + rng.integers(0, hub_core - 1, live2)) % hub_core
    dst += [two, cs[a], cs[b]]
    n = core + 1 + live + dead + live2
    return mce.from_edges(np.concatenate(src).astype(np.uint32), np.concatenate(dst).astype(np.uint32), n=n)


@pytest.mark.parametrize("live,dead,live2,two_first", [(3000, 2500, 0, False), (4500, 200, 0, False),
                                                      (3000, 0, 3000, False), (100, 100, 4200, False),
                                                      (500, 0, 4080, False), (500, 0, 4080, True)])
def test_hub_outside_rows_beyond_bitmap(mce, oracle, live, dead, live2, two_first):
    # a hub subproblem with more than 4096 outside rows: when at most 4096 of
    # them are kept (a neighbour in P; a one-neighbour row only if no earlier
    # kept row holds that neighbour) it runs on the bitmap path over them
    # (3000 live; 40 distinct of 4500; 40 + 3000 of 6000; 4080 two-neighbour
    # rows first, covering the 500 one-neighbour rows after them), otherwise
    # on the list path (4238 rows; 40 first one-neighbour rows + 4080)
    g = hub_graph(mce, live, dead, live2, two_first)
    ref = oracle.mce(g, level=1, ipx=True, sub_nodes=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        for wl in (1, 2):
            print(f"case hub live={live} dead={dead} live2={live2} two_first={two_first} wl={wl}", flush=True)
            res = dg.count(level=1, induced=2, worklist=wl, sub_nodes=True)
            assert_same(res, ref, (live, dead, wl))
            np.testing.assert_array_equal(res["sub_nodes"], ref["sub_nodes"])


def test_hbm_arena_and_frame_spill(mce, oracle):
    # a tiny LDS arena forces the HBM arena and HBM stack frames
    g = rand_graph(mce, 220, 0.55, 7)
    run_all(mce, oracle, g, "hbm-arena", lds_arena_bytes=1024)
    # a complete graph K_150: one clique, 149-deep recursion in one subproblem
    n = 150
    iu = np.triu_indices(n, 1)
    k = mce.from_edges(iu[0], iu[1], n=n)
    run_all(mce, oracle, k, "K150", lds_arena_bytes=2048)


def test_degenerate_graphs(mce, oracle):
    # isolated vertices, a single edge, a star, a path, an empty graph
    cases = {
        "edge": ([0], [1], 2),
        "isolated": ([0, 5], [1, 6], 9),
        "star": ([0] * 9, list(range(1, 10)), 10),
        "path": (list(range(0, 30)), list(range(1, 31)), 31),
        "empty": ([], [], 5),
    }
    for name, (s, d, n) in cases.items():
        g = mce.from_edges(np.array(s, np.uint32), np.array(d, np.uint32), n=n)
        run_all(mce, oracle, g, name)


def test_schedule_invariance(mce, oracle):
    """Any order of first-level subproblems gives identical results."""
    g = mce.gen_rmat(12, 16, seed=5)
    ref = oracle.mce(g, level=1, ipx=True)
    rng = np.random.default_rng(9)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        for t in range(3):
            order = rng.permutation(g.n).astype(np.uint32)
            res = dg.count(level=1, induced=2, order=order)
            assert_same(res, ref, f"perm-{t}")


def test_per_subproblem_labels(mce, oracle):
    g = mce.gen_rmat(12, 16, seed=11)
    ref = oracle.mce(g, level=1, ipx=True, sub_nodes=True, sub_cliques=True)
    res = mce.run(g, level=1, induced=2, sub_nodes=True, sub_cliques=True)
    np.testing.assert_array_equal(res["sub_nodes"], ref["sub_nodes"])
    np.testing.assert_array_equal(res["sub_cliques"], ref["sub_cliques"])
    assert int(res["sub_cliques"].sum()) == res["count"]
