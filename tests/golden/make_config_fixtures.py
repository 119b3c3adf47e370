"""Writes tests/golden/configs.json: the pinned inputs of the five benchmark
configurations and the oracle's answers that the GPU tests compare against
where the oracle itself is too slow to run inside a test.

For every configuration: the CSR digest of the seeded generator's output
(mce_csr_digest over n, m, row_ptr and col_ind), so that a change to a
generator cannot silently change a workload. For the SBM batch also a digest
of the graph offsets. For R-MAT scale 24 (about 1e9 arcs) also what the CPU
oracle computes once here: the SHA-256 of its batched-peel priorities, split
offsets and oriented column indices, the Batagelj-Zaversnik degeneracy (an
independent algorithm), the degree maxima, and the exact results of a fixed
sample of first-level subproblems (count, histogram, hash and tree nodes for
each level and induced-subgraph mode), which the GPU engine must reproduce
bit for bit.

The sample: seeded uniform picks among the vertices with out-degree in
[1, 96) after orientation, whose subtrees are small, so that the oracle
finishes in seconds.

Run from the repository root:  python tests/golden/make_config_fixtures.py
"""
import hashlib
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def _import(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# The workloads exactly as bench.py builds them (seed 1), plus the Chung-Lu
# evaluation seed 9.
CONFIGS = {
    "c1": ("gen_er", dict(n=4096, m=65536, seed=1)),
    "c2": ("gen_rmat", dict(scale=20, edge_factor=16, seed=1)),
    "c3_seed1": ("gen_chung_lu", dict(n=1 << 22, beta=2.1, mean_degree=24.0, planted=2000, k_min=20, k_max=60,
                                      seed=1)),
    "c3_seed9": ("gen_chung_lu", dict(n=1 << 22, beta=2.1, mean_degree=24.0, planted=2000, k_min=20, k_max=60,
                                      seed=9)),
    "c4": ("gen_sbm_batch", dict(num_graphs=4096, n_min=64, n_max=4096, blocks=4, p_in=0.3, p_out=0.01, seed=1)),
    "c5": ("gen_rmat", dict(scale=24, edge_factor=32, seed=1)),
}

SAMPLE_SEED = 2024
SAMPLE_SIZE = 256
SAMPLE_DPLUS = (1, 96)


def sample_roots(split, row_ptr, k=SAMPLE_SIZE, seed=SAMPLE_SEED, lo=SAMPLE_DPLUS[0], hi=SAMPLE_DPLUS[1]):
    dplus = (row_ptr[1:] - split).astype(np.int64)
    cand = np.nonzero((dplus >= lo) & (dplus < hi))[0]
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(cand, size=min(k, cand.size), replace=False)).astype(np.uint32)


def main(which):
    mce = _import("mce_gpu_amd", os.path.join(ROOT, "mce-gpu_amd", "__init__.py"))
    oracle = _import("mce_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    path = os.path.join(HERE, "configs.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for name, (fn, kw) in CONFIGS.items():
        if which and name not in which:
            continue
        t = time.time()
        res = getattr(mce, fn)(**kw)
        entry = {"generator": fn, "args": kw}
        if fn == "gen_sbm_batch":
            off, g = res
            entry["vtx_off_sha256"] = sha(off.astype(np.uint32))
        else:
            g = res
        entry.update(n=int(g.n), m=int(g.m), digest=f"{g.digest():#018x}")
        print(f"{name}: n={g.n} m={g.m} digest={entry['digest']} ({time.time() - t:.0f}s)", flush=True)
        if name == "c5":
            t = time.time()
            prio, split, oc, d = oracle.order_orient(g)
            core, k = oracle.core_numbers(g)
            deg = np.diff(g.row_ptr)
            dplus = (g.row_ptr[1:] - split).astype(np.int64)
            entry.update(degeneracy=int(d), bz_degeneracy=int(k), max_degree=int(deg.max()),
                         max_out_degree=int(dplus.max()), prio_sha256=sha(prio), split_sha256=sha(split),
                         oc_sha256=sha(oc))
            print(f"  ordering: degeneracy {d} (BZ {k}), max out-degree {dplus.max()} ({time.time() - t:.0f}s)",
                  flush=True)
            roots = sample_roots(split, g.row_ptr)
            samp = {"seed": SAMPLE_SEED, "size": SAMPLE_SIZE, "out_degree": list(SAMPLE_DPLUS),
                    "vertices": roots.tolist(), "results": {}}
            for level in (1, 2):
                for ipx in (False, True):
                    t = time.time()
                    r = oracle.mce(g, level=level, ipx=ipx, order=roots)
                    samp["results"][f"L{level}_{'IPX' if ipx else 'IP'}"] = {
                        "count": r["count"], "hash": f"{r['hash']:#018x}",
                        "hist": {str(k): v for k, v in r["hist"].items()}, "tree_nodes": r["tree_nodes"]}
                    print(f"  sample L{level} ipx={ipx}: {r['count']} cliques, {r['tree_nodes']} nodes "
                          f"({time.time() - t:.0f}s)", flush=True)
            entry["sample"] = samp
        out[name] = entry
        del g, res
        json.dump(out, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main(set(sys.argv[1:]))
