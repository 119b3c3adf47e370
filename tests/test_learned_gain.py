"""What the trained cost model gains over the method without it, on a graph
it was not trained on. The committed Chung-Lu model (16 -> 64 -> 64 -> 1,
trained on the GPU on seeds 1-8 at n = 2^18, 2^19 and 2^20, beta 2.1, mean
degree 24, n/2048 planted cliques of 20..60) is evaluated on seed 9 at 2^18:

  - ranking: the Spearman correlation between its predictions and the
    measured BK tree size of every first-level subproblem beats that of the
    out-degree, which is the plain schedule's key;
  - time: the learned engine (model order, the variant and donation
    threshold the model sets) runs the enumeration faster than the plain
    engine (out-degree order, default thresholds) with the same variant rule
    and block settings; both give the oracle's result.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "mce-gpu_amd", "cost_model_cl.mlp")


def spearman(a, b):
    ra = np.argsort(np.argsort(a, kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b, kind="stable"), kind="stable").astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / np.sqrt((ra * ra).sum() * (rb * rb).sum()))


def test_model_beats_out_degree_on_held_out_seed(mce, oracle):
    n = 1 << 18
    g = mce.gen_chung_lu(n, 2.1, 24.0, n >> 11, 20, 60, seed=9)
    ref = oracle.mce(g, level=1, ipx=True)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        lab = dg.count(level=1, induced=2, worklist=2, sub_nodes=True)
        assert int(lab["sub_nodes"].sum()) == ref["tree_nodes"]
        F = dg.features()
        roots = F[:, 1] > 0  # out-degree >= 1: the first-level subproblems
        y = np.log1p(lab["sub_nodes"].astype(np.float64))[roots]
        pred = mce.mlp_predict(mce.mlp_load(MODEL), F[roots])
        rho_model, rho_deg = spearman(pred, y), spearman(F[roots, 1], y)
        print(f"Spearman rho on held-out seed 9: model {rho_model:.4f}, out-degree {rho_deg:.4f}", flush=True)
        assert rho_model > rho_deg + 0.01

        times = {}
        for name, kw in (("learned", dict(model_path=MODEL)), ("plain", {})):
            runs = []
            for _ in range(3):
                r = dg.count(level=1, induced=2, worklist=mce.WL_AUTO, **kw)
                assert (r["count"], r["hash"], r["hist"]) == (ref["count"], ref["hash"], ref["hist"]), name
                assert r["fallback_events"] == 0
                runs.append(r["t_count_ms"])
            times[name] = min(runs)
        print(f"enumeration: learned {times['learned']:.1f} ms, plain {times['plain']:.1f} ms, "
              f"{times['plain'] / times['learned']:.2f}x", flush=True)
        assert times["learned"] < times["plain"]
