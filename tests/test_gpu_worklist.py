"""Worker-list (donation) engine on the GPU: results and BK tree-node counts
identical to the oracle while subtrees move between waves; the evaluation
variants (lb / donor / bd) report consistent per-wave statistics."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# donate at every poll, down to tiny subtrees, so donation really happens
EAGER = dict(donate_min_p=2, donate_check_nodes=1, donate_age_us=1)


def rand_graph(mce, n, p, seed):
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(iu[0].size) < p
    return mce.from_edges(iu[0][keep], iu[1][keep], n=n)


def same(res, ref, what):
    assert res["count"] == ref["count"], (what, res["count"], ref["count"])
    assert res["hist"] == ref["hist"], what
    assert res["hash"] == ref["hash"], what
    assert res["tree_nodes"] == ref["tree_nodes"], (what, res["tree_nodes"], ref["tree_nodes"])


@pytest.mark.parametrize("level,induced", [(1, 1), (1, 2), (2, 1), (2, 2)])
def test_worklist_parity_rmat(mce, oracle, level, induced):
    g = mce.gen_rmat(12, 16, seed=3)
    ref = oracle.mce(g, level=level, ipx=(induced == 2))
    with mce.DeviceGraph(g) as dg:
        dg.order()
        res = dg.count(level=level, induced=induced, worklist=mce.WL, **EAGER)
        same(res, ref, ("rmat12-wl", level, induced))
        assert res["donations"] > 0, (level, induced)
        res2 = dg.count(level=level, induced=induced, worklist=mce.WL)  # default thresholds
        same(res2, ref, ("rmat12-wl-default", level, induced))


def test_worklist_parity_random(mce, oracle):
    for seed in range(12):
        rng = np.random.default_rng(500 + seed)
        n = int(rng.integers(20, 160))
        p = float(rng.uniform(0.2, 0.8))
        print(f"case seed={seed} n={n} p={p:.3f}", flush=True)
        g = rand_graph(mce, n, p, seed)
        with mce.DeviceGraph(g) as dg:
            dg.order()
            for level, induced in ((1, 2), (1, 1), (2, 2)):
                ref = oracle.mce(g, level=level, ipx=(induced == 2))
                res = dg.count(level=level, induced=induced, worklist=mce.WL, **EAGER)
                same(res, ref, (seed, level, induced))


def test_worklist_with_hbm_arena(mce, oracle):
    g = rand_graph(mce, 260, 0.5, 77)
    ref = oracle.mce(g, level=1, ipx=True)
    res = mce.run(g, level=1, induced=2, worklist=mce.WL, lds_arena_bytes=2048, **EAGER)
    same(res, ref, "wl-hbm")
    assert res["donations"] > 0


def test_labels_with_donation(mce, oracle):
    g = mce.gen_rmat(12, 16, seed=11)
    ref = oracle.mce(g, level=1, ipx=True, sub_nodes=True, sub_cliques=True)
    res = mce.run(g, level=1, induced=2, worklist=mce.WL, sub_nodes=True, sub_cliques=True, **EAGER)
    np.testing.assert_array_equal(res["sub_nodes"], ref["sub_nodes"])
    np.testing.assert_array_equal(res["sub_cliques"], ref["sub_cliques"])


def test_eval_variants(mce, oracle):
    g = mce.gen_rmat(11, 16, seed=2)
    ref = oracle.mce(g, level=1, ipx=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        lb = dg.count(level=1, induced=2, worklist=mce.WL, task=mce.TASK_LB_EVAL, wave_stats_cap=1 << 16, **EAGER)
        same(lb, ref, "lb")
        assert len(lb["wave_stats"]) == lb["num_waves"]
        assert sum(w[2] for w in lb["wave_stats"]) == lb["tree_nodes"]
        lb_nowl = dg.count(level=1, induced=2, worklist=mce.NOWL, task=mce.TASK_LB_EVAL, wave_stats_cap=1 << 16)
        same(lb_nowl, ref, "lb-nowl")
        assert sum(w[2] for w in lb_nowl["wave_stats"]) == lb_nowl["tree_nodes"]
        dn = dg.count(level=1, induced=2, worklist=mce.WL, task=mce.TASK_DONOR_EVAL, wave_stats_cap=1 << 16, **EAGER)
        same(dn, ref, "donor")
        assert sum(w[2] for w in dn["wave_stats"]) == dn["donations"]
        bd = dg.count(level=1, induced=2, worklist=mce.WL, task=mce.TASK_BD_EVAL, wave_stats_cap=1 << 16)
        same(bd, ref, "bd")
        cycles = np.array([w[3] for w in bd["wave_stats"]], dtype=np.float64)
        assert cycles.shape[1] == 7 and cycles.sum() > 0
        with pytest.raises(mce.MCEError):
            dg.count(level=1, induced=2, worklist=mce.NOWL, task=mce.TASK_BD_EVAL)


def test_watchdog_gives_up_cleanly(mce, oracle):
    """With a 1 us limit, idle waves stop waiting for work and the call reports
    the failure instead of hanging; every wave ends normally, and the next
    call on the same graph is correct. Donation is off, so no task goes to a
    wave that stopped waiting."""
    g = mce.gen_rmat(12, 16, seed=3)
    ref = oracle.mce(g, level=1, ipx=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        with pytest.raises(mce.MCEError) as e:
            dg.count(level=1, induced=2, worklist=mce.WL, donate_min_p=1 << 30, watchdog_us=1)
        assert e.value.code == -8 and "watchdog" in e.value.detail
        res = dg.count(level=1, induced=2, worklist=mce.WL)
        same(res, ref, "after-watchdog")
