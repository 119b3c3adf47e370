"""The benchmark configurations on the GPU, at their real sizes where the
oracle or a committed oracle fixture can check them:

  R-MAT scale 20 (edge factor 16): ordering and orientation bit-equal to the
      oracle's (priorities, split offsets, oriented columns), degeneracy equal
      to Batagelj-Zaversnik, degree maxima; a seeded sample of first-level
      subproblems enumerated in every variant and checked against the oracle.
  Chung-Lu (beta 2.1, mean degree 24, planted cliques of 20..60): at n = 2^16
      every level x induced mode x worker-list variant against the oracle; at
      the full n = 4,194,304 with 2,000 planted cliques (evaluation seed 9) the
      ordering bit-equal to the oracle's and a sample of subproblems.
  R-MAT scale 24 (edge factor 32, ~1e9 arcs): digests of the ordering arrays,
      the degeneracy (Batagelj-Zaversnik), the degree maxima and a sample of
      subproblems against tests/golden/configs.json (written by the oracle).

The whole-graph enumerations of the R-MAT configurations are not run: their
clique counts are estimated at 1e12 and beyond.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIGS = json.load(open(os.path.join(HERE, "golden", "configs.json")))

VARIANTS = [(lvl, ind, wl) for lvl in (1, 2) for ind in (1, 2) for wl in (1, 2)]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sample_roots(split, row_ptr, k, seed, lo, hi):
    """Seeded uniform picks among the vertices with out-degree in [lo, hi)."""
    dplus = (row_ptr[1:] - split).astype(np.int64)
    cand = np.nonzero((dplus >= lo) & (dplus < hi))[0]
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(cand, size=min(k, cand.size), replace=False)).astype(np.uint32)


def check_order_against_oracle(mce, oracle, g, dg, what):
    info = dg.order()
    prio, split, oc = dg.download()
    rp, rs, ro, d = oracle.order_orient(g)
    core, k = oracle.core_numbers(g)
    assert info["degeneracy"] == d == k, what
    assert np.array_equal(prio, rp), what
    assert np.array_equal(split, rs), what
    assert np.array_equal(oc, ro), what
    dplus = (g.row_ptr[1:] - split).astype(np.int64)
    assert info["max_out_degree"] == int(dplus.max()) <= d, what
    assert info["max_degree"] == int(np.diff(g.row_ptr).max()), what
    return split


def test_c2_rmat20_order_and_sample(mce, oracle):
    g = mce.gen_rmat(20, 16, seed=1)
    assert f"{g.digest():#018x}" == CONFIGS["c2"]["digest"]
    with mce.DeviceGraph(g) as dg:
        print("case c2 ordering", flush=True)
        split = check_order_against_oracle(mce, oracle, g, dg, "c2")
        roots = sample_roots(split, g.row_ptr, 384, 7, 1, 128)
        for lvl, ind, wl in VARIANTS:
            print(f"case c2 sample level={lvl} induced={ind} worklist={wl}", flush=True)
            ref = oracle.mce(g, level=lvl, ipx=ind == 2, order=roots)
            res = dg.count(level=lvl, induced=ind, worklist=wl, order=roots)
            assert_same(res, ref, ("c2-sample", lvl, ind, wl))


def test_c3_chung_lu_2_16_all_variants(mce, oracle):
    n = 1 << 16
    g = mce.gen_chung_lu(n, 2.1, 24.0, n >> 11, 20, 60, seed=9)
    with mce.DeviceGraph(g) as dg:
        dg.order()
        refs = {}
        for lvl, ind, wl in VARIANTS:
            print(f"case c3-2^16 level={lvl} induced={ind} worklist={wl}", flush=True)
            if (lvl, ind) not in refs:
                refs[(lvl, ind)] = oracle.mce(g, level=lvl, ipx=ind == 2)
            res = dg.count(level=lvl, induced=ind, worklist=wl)
            assert_same(res, refs[(lvl, ind)], ("c3-2^16", lvl, ind, wl))
            if wl == 2:
                assert res["worklist_used"] == 2


def test_c3_full_size_order_and_sample(mce, oracle):
    g = mce.gen_chung_lu(1 << 22, 2.1, 24.0, 2000, 20, 60, seed=9)
    assert f"{g.digest():#018x}" == CONFIGS["c3_seed9"]["digest"]
    with mce.DeviceGraph(g) as dg:
        print("case c3 full ordering", flush=True)
        split = check_order_against_oracle(mce, oracle, g, dg, "c3")
        roots = sample_roots(split, g.row_ptr, 2048, 9, 1, 96)
        for lvl, ind, wl in ((1, 2, 2), (1, 1, 1), (2, 2, 1)):
            print(f"case c3 full sample level={lvl} induced={ind} worklist={wl}", flush=True)
            ref = oracle.mce(g, level=lvl, ipx=ind == 2, order=roots)
            res = dg.count(level=lvl, induced=ind, worklist=wl, order=roots)
            assert_same(res, ref, ("c3-sample", lvl, ind, wl))


def test_c5_rmat24_order_digests_and_sample(mce):
    fx = CONFIGS["c5"]
    g = mce.gen_rmat(24, 32, seed=1)
    assert (g.n, g.m, f"{g.digest():#018x}") == (fx["n"], fx["m"], fx["digest"])
    with mce.DeviceGraph(g) as dg:
        print("case c5 ordering", flush=True)
        info = dg.order()
        assert info["degeneracy"] == fx["degeneracy"] == fx["bz_degeneracy"]
        assert info["max_degree"] == fx["max_degree"]
        assert info["max_out_degree"] == fx["max_out_degree"] <= info["degeneracy"]
        assert info["max_out_degree"] <= mce.MAX_OUT_DEGREE
        prio, split, oc = dg.download()
        assert sha(prio) == fx["prio_sha256"]
        assert sha(split) == fx["split_sha256"]
        del prio
        assert sha(oc) == fx["oc_sha256"]
        del oc
        samp = fx["sample"]
        roots = sample_roots(split, g.row_ptr, samp["size"], samp["seed"], *samp["out_degree"])
        assert roots.tolist() == samp["vertices"]
        for key, ref in samp["results"].items():
            lvl, ind = int(key[1]), 2 if key.endswith("IPX") else 1
            for wl in (1, 2):
                print(f"case c5 sample {key} worklist={wl}", flush=True)
                res = dg.count(level=lvl, induced=ind, worklist=wl, order=roots)
                want = dict(count=ref["count"], hash=int(ref["hash"], 16),
                            hist={int(k): v for k, v in ref["hist"].items()}, tree_nodes=ref["tree_nodes"])
                assert_same(res, want, ("c5-sample", key, wl))
