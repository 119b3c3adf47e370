"""mce_graph_count_noalloc (static scratch array, no device allocation)
against mce_graph_count on the same graphs: count, hash, histogram and tree
nodes must be identical. Two calls in a row, the second on a smaller graph,
so that the second call runs on the scratch the first one left behind."""
import pytest

from test_gpu_parity import rand_graph

pytestmark = pytest.mark.gpu

KEYS = ("count", "hash", "hist", "tree_nodes", "max_clique_size")


def check(mce, g, what, **block):
    with mce.DeviceGraph(g) as dg:
        dg.order()
        for ind in (mce.IP, mce.IPX):
            print(f"case noalloc {what} induced={ind}", flush=True)
            ref = dg.count(level=mce.L1, induced=ind, worklist=mce.NOWL, **block)
            got = dg.count_noalloc(induced=ind, **block)
            for k in KEYS:
                assert got[k] == ref[k], (what, ind, k)
            assert got["count"] > 0


def test_noalloc_matches_count_two_calls(mce):
    check(mce, mce.gen_rmat(12, 16, seed=3), "rmat12")
    check(mce, rand_graph(mce, 300, 0.2, 5), "rand300")


def test_noalloc_small_lds_uses_scratch_arenas(mce):
    # a 1 KiB LDS share per wave sends the subproblems' bit matrices to the
    # per-wave arenas carved from the scratch array
    check(mce, rand_graph(mce, 400, 0.3, 21), "rand400-hbm", lds_arena_bytes=1024)
    check(mce, mce.gen_er(4096, 65536, 1), "er4096-hbm", lds_arena_bytes=1024)


def test_noalloc_refuses_other_variants(mce):
    g = rand_graph(mce, 100, 0.2, 2)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        o = mce.default_opts(level=mce.L2, device=0)
        r = mce.Result()
        assert mce.lib.mce_graph_count_noalloc(dg.handle, mce.byref(o), mce.byref(r)) == -1
        o = mce.default_opts(level=mce.L1, worklist=mce.WL, device=0)
        assert mce.lib.mce_graph_count_noalloc(dg.handle, mce.byref(o), mce.byref(r)) == -1
