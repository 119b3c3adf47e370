"""Benchmark: maximal cliques per second on one MI355X.

A step is one pass of the hot path over the HBM-resident synthetic graph:
degeneracy ordering (k-core peel) + orientation + Bron-Kerbosch enumeration,
all on the GPU through libmce.so. The timed region starts with the graph
already in HBM. Results are checked bit-exactly: the same bounded sample that
times the CPU baseline is also run on the GPU and compared (count, size
histogram, clique hash), and every timed step must reproduce the first step's
count and hash.

Workloads (--config), all seeded:
  c1  Erdos-Renyi n=4096, m=65,536
  c2  R-MAT scale 20, edge factor 16, (.57,.19,.19,.05)
  c3  Chung-Lu n=4,194,304, beta 2.1, mean degree 24, 2,000 planted cliques of 20..60
  c4  4,096 stochastic-block-model graphs, n ~ U[64,4096], 4 blocks, p_in .3, p_out .01
  c5  R-MAT scale 24, edge factor 32
The default is c4, the largest configuration whose exact enumeration finishes
on one GPU in about a minute. By Knuth's tree-size estimator over the oracle's
BK tree, c2 holds about 4e12 maximal cliques (1e13 tree nodes), the full c3
about 3e12 (9.5e12 nodes) and c5 about 5e21: hours of enumeration for c2 and
c3, out of reach of any engine for c5. c3 runs at smaller scales with
--cl-scale (2^21 takes about ten minutes).

Prints one JSON line (rank 0, N=1).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "maximal cliques/s on 1 MI355X: learned-scheduled vs plain BK, count bit-exact"
HBM_PEAK_GBS = 8000.0
SCHEDULES = {"plain": 0, "learned": 1}


def spearman(a, b):
    ra = np.argsort(np.argsort(a, kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b, kind="stable"), kind="stable").astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / np.sqrt((ra * ra).sum() * (rb * rb).sum()))


def _import(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(mce, name, seed, cl_scale=22):
    """Returns (graph, vtx_off or None, description dict)."""
    if name == "c1":
        return mce.gen_er(4096, 65536, seed), None, dict(workload="er", n=4096, m=65536, seed=seed)
    if name == "c2":
        return mce.gen_rmat(20, 16, seed=seed), None, dict(workload="rmat", scale=20, edge_factor=16, seed=seed)
    if name == "c3":
        # the 4 M-vertex configuration plants 2,000 cliques; smaller scales keep n / 2048 of them
        n = 1 << cl_scale
        planted = 2000 if cl_scale == 22 else n >> 11
        return (mce.gen_chung_lu(n, 2.1, 24.0, planted, 20, 60, seed=seed), None,
                dict(workload="chung-lu", n=n, beta=2.1, mean_degree=24, planted=planted, k=[20, 60], seed=seed))
    if name == "c4":
        off, g = mce.gen_sbm_batch(4096, 64, 4096, 4, 0.3, 0.01, seed=seed)
        return g, off, dict(workload="sbm-batch", graphs=4096, n=[64, 4096], blocks=4, p_in=0.3, p_out=0.01,
                            seed=seed)
    if name == "c5":
        return mce.gen_rmat(24, 32, seed=seed), None, dict(workload="rmat", scale=24, edge_factor=32, seed=seed)
    raise SystemExit(f"unknown config {name}")


def pinned_digest(name, seed, cl_scale, g, off):
    """Checks the generated workload against the CSR digest committed in
    tests/golden/configs.json for this configuration and seed, if there is
    one: a generator change must not silently change the benchmark."""
    key = {"c3": f"c3_seed{seed}" if cl_scale == 22 else None}.get(name, name if seed == 1 else None)
    path = os.path.join(ROOT, "tests", "golden", "configs.json")
    if key is None or not os.path.exists(path):
        return None
    fx = json.load(open(path)).get(key)
    if fx is None or fx["args"].get("seed") != seed:
        return None
    got = f"{g.digest():#018x}"
    if (g.n, g.m, got) != (fx["n"], fx["m"], fx["digest"]):
        raise SystemExit(f"{name} seed {seed}: CSR digest {got} differs from the committed {fx['digest']}")
    if off is not None:
        import hashlib
        if hashlib.sha256(off.astype(np.uint32).tobytes()).hexdigest() != fx["vtx_off_sha256"]:
            raise SystemExit(f"{name} seed {seed}: graph offsets differ from the committed ones")
    return got


def sample_workload(mce, g, off, frac):
    """Bounded sample of the same workload. A batch keeps every k-th member
    graph (cliques never cross graphs); a single graph keeps every k-th
    first-level vertex (the cliques whose earliest vertex it is).
    Returns (graph, first-level order or None, description)."""
    k = max(1, int(round(1.0 / frac)))
    if off is not None:
        sel = np.arange(0, off.size - 1, k)
        parts_rp, parts_ci, base = [np.zeros(1, np.uint64)], [], 0
        for i in sel:
            a, b = int(off[i]), int(off[i + 1])
            rp = g.row_ptr[a:b + 1] - g.row_ptr[a]
            ci = g.col_ind[g.row_ptr[a]:g.row_ptr[b]].astype(np.int64) - a + base
            parts_rp.append(rp[1:] + parts_rp[-1][-1])
            parts_ci.append(ci.astype(np.uint32))
            base += b - a
        rp = np.concatenate(parts_rp)
        ci = np.concatenate(parts_ci) if parts_ci else np.zeros(0, np.uint32)
        return mce.Graph(rp, ci), None, f"every {k}th member graph ({sel.size} of {off.size - 1} graphs)"
    order = np.arange(0, g.n, k, dtype=np.uint32)
    return g, order, f"every {k}th first-level vertex ({order.size} of {g.n})"


def counter_summary(name, workload_key):
    """key = value lines of a committed rocprofv3 counter summary under
    profiles/, when it was taken on this workload."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            kv[k.strip()] = v.strip()
    return kv if kv.get("workload") == workload_key else None


def algorithmic_bytes(g, split, ipx=True):
    """HBM bytes the enumeration must stream at least once per launch: every
    first-level vertex reads its row and, to build its induced bit matrix, the
    out-lists of its out-neighbours (and, with IPX, of its in-neighbours);
    plus the ordering/orientation passes over the CSR. Per vertex u, its
    out-list is read once per in-neighbour (d-(u) times) and, with IPX, once
    per out-neighbour (d+(u) times)."""
    rp = g.row_ptr.astype(np.float64)
    deg = np.diff(rp)
    dplus = rp[1:] - split.astype(np.float64)
    dminus = deg - dplus
    b = 4.0 * deg.sum() + 4.0 * (dplus * dminus).sum()
    if ipx:
        b += 4.0 * (dplus * dplus).sum()
    return float(b + 4.0 * (4.0 * g.n + 3.0 * rp[-1]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--level", type=int, default=1)
    ap.add_argument("--induced", type=int, default=2)
    ap.add_argument("--worklist", type=int, default=3, help="1 nowl, 2 wl, 3 auto (picked by the cost model)")
    ap.add_argument("--schedule", choices=["learned", "plain"], default="learned")
    ap.add_argument("--model", default=None, help="cost model file (default: the committed model of the family)")
    ap.add_argument("--cl-scale", type=int, default=22, help="c3: log2 of the vertex count")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-schedule comparison step")
    ap.add_argument("--autotune", choices=["auto", "on", "off"], default="auto",
                    help="tune block settings before timing (auto: for batches, whose many small subproblems a "
                         "sample represents well; a sample misjudges a heavy tail)")
    ap.add_argument("--no-autotune", action="store_true", help="same as --autotune off")
    ap.add_argument("--tune-sample", type=float, default=1.0 / 64,
                    help="fraction of first-level subproblems the autotuner times each trial on")
    ap.add_argument("--cpu-sample", type=float, default=1.0 / 64,
                    help="fraction of member graphs (batch) or first-level vertices timed on the CPU")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--write-bel", default=None, metavar="PATH",
                    help="only write the selected workload (--config, --seed; c1, c2, c3, c5) as a .bel file for "
                         "the parallel_mce_on_gpus command line and exit")
    ap.add_argument("--rmat-scale", type=int, default=0,
                    help="with --write-bel: an R-MAT graph of this scale (edge factor 16, seed --seed) instead")
    ap.add_argument("--rank-sample", type=int, default=0, metavar="N",
                    help="only score the cost model's ranking: exact tree sizes of N first-level subproblems drawn "
                         "uniformly (seeded) from the workload, by one given-order enumeration, and the Spearman rho "
                         "of the model's predictions and of out-degree against them; prints JSON and exits (for "
                         "graphs too large to label whole)")
    ap.add_argument("--peaks", action="store_true",
                    help="only measure the device's VALU lane-op/s and LDS byte/s peaks (mce_device_peaks), print "
                         "them as JSON and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (the reference kernel for the HBM "
                         "traffic counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(args.copy_calibration // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": a.numel() * 8}))
        return
    if args.write_bel:
        mce = _import("mce_gpu_amd", os.path.join(ROOT, "mce-gpu_amd", "__init__.py"))
        if args.rmat_scale:
            g = mce.gen_rmat(args.rmat_scale, 16, seed=args.seed)
        else:
            g, off, _ = make_workload(mce, args.config, args.seed, args.cl_scale)
            if off is not None:
                raise SystemExit("a batch has no single-graph .bel form")
        g.write_bel(args.write_bel)
        print(json.dumps({"bel": args.write_bel, "n": g.n, "arcs": g.m, "digest": f"{g.digest():#018x}"}))
        return
    if args.peaks:
        mce = _import("mce_gpu_amd", os.path.join(ROOT, "mce-gpu_amd", "__init__.py"))
        runs = [mce.device_peaks() for _ in range(3)]
        print(json.dumps({"valu_lane_ops_per_s": max(r["valu_lane_ops_per_s"] for r in runs),
                          "lds_bytes_per_s": max(r["lds_bytes_per_s"] for r in runs), "runs": runs}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the engine pins the device count to 1",
              file=sys.stderr)
        sys.exit(2)

    mce = _import("mce_gpu_amd", os.path.join(ROOT, "mce-gpu_amd", "__init__.py"))
    if mce.device_count() < 1:
        raise SystemExit("no GPU visible to libmce.so")
    t0 = time.time()
    g, off, desc = make_workload(mce, args.config, args.seed, args.cl_scale)
    if args.rank_sample:
        fam = {"c3": "cl", "c4": "sbm"}.get(args.config, "sbm")
        model = args.model or os.path.join(ROOT, "mce-gpu_amd", f"cost_model_{fam}.mlp")
        with mce.DeviceGraph(g) as dg:
            dg.order()
            F = dg.features()
            roots = np.nonzero(F[:, 1] > 0)[0]
            rng = np.random.default_rng(args.seed * 7919 + 1)
            pick = np.sort(rng.choice(roots, size=min(args.rank_sample, roots.size), replace=False)).astype(np.uint32)
            t = time.perf_counter()
            lab = dg.count(level=1, induced=2, worklist=2, order=pick, sub_nodes=True, progress_s=30)
            y = np.log1p(lab["sub_nodes"][pick].astype(np.float64))
            pred = mce.mlp_predict(mce.mlp_load(model), F[pick])
            print(json.dumps({"config": desc, "model": os.path.basename(model), "roots": int(roots.size),
                              "sampled": int(pick.size), "sample_tree_nodes": lab["tree_nodes"],
                              "sample_cliques": lab["count"], "enum_s": time.perf_counter() - t,
                              "spearman_model": spearman(pred, y), "spearman_out_degree": spearman(F[pick, 1], y)}))
        return
    if args.model is None:
        fam = {"c3": "cl", "c4": "sbm"}.get(args.config, "sbm")
        args.model = os.path.join(ROOT, "mce-gpu_amd", f"cost_model_{fam}.mlp")
    log(f"[bench] generated {args.config}: n={g.n} arcs={g.m} in {time.time() - t0:.1f}s")
    desc["digest"] = pinned_digest(args.config, args.seed, args.cl_scale, g, off)

    model_path = args.model if args.schedule == "learned" else None
    # progress_s: an enumeration launch longer than half a minute prints its
    # own counters (tickets claimed, tasks outstanding) while it runs
    kw = dict(level=args.level, induced=args.induced, worklist=args.worklist,
              schedule=SCHEDULES[args.schedule], model_path=model_path, progress_s=30)
    if off is not None:
        # the batched engine: one persistent enumeration launch captured into a hipGraph
        kw["graph_launch"] = 1
    with mce.DeviceGraph(g) as dg:
        for w in range(args.warmup):
            info = dg.order()
            r = dg.count(**kw)
            log(f"[bench] warmup {w}: count={r['count']} order={info['t_order_ms']:.1f}ms "
                f"count={r['t_count_ms']:.1f}ms")
        tuned = None
        do_tune = args.autotune == "on" or (args.autotune == "auto" and off is not None)
        if do_tune and not args.no_autotune:
            # block settings (and, left to it, the variant) timed on a sample
            # of this graph; untimed set-up like the warmup
            t = time.perf_counter()
            if not args.warmup:
                dg.order()  # the tuner runs on the ordered graph
            best, trials = dg.autotune(sample_fraction=args.tune_sample, level=args.level, induced=args.induced,
                                       worklist=args.worklist)
            tuned = dict(best, trials=len(trials), seconds=time.perf_counter() - t)
            kw.update(best)
            log(f"[bench] autotuned in {tuned['seconds']:.1f}s over {len(trials)} trials: {best}")
        step_ms, kern_ms, order_ms = [], [], []
        first = None
        for k in range(args.steps):
            t = time.perf_counter()
            info = dg.order()
            r = dg.count(sub_cliques=off is not None, **kw)
            step_ms.append((time.perf_counter() - t) * 1e3)
            kern_ms.append(r["t_count_ms"])
            order_ms.append(info["t_order_ms"] + info["t_orient_ms"])
            log(f"[bench] step {k}: {step_ms[-1]:.1f} ms (order+orient {order_ms[-1]:.1f}, enum kernel "
                f"{kern_ms[-1]:.1f}) count={r['count']} donations={r['donations']} waves={r['num_waves']} "
                f"fallbacks={r['fallback_events']}")
            if first is None:
                first = r
            elif (r["count"], r["hash"]) != (first["count"], first["hash"]):
                raise SystemExit("steps disagree: results are not reproducible")
        _, split, _ = dg.download(prio=False, split=True, col=False)
        plain = None
        if not args.no_plain:
            # The plain-schedule engine on the same resident graph with the same
            # settings (tuned block settings, variant rule, graph launch, warm-up
            # and steps): only the schedule changes. Under -w auto the learned
            # engine picks the variant from its predictions while the plain one
            # takes the worker list, so when the learned run picked another
            # variant the plain engine is also timed with that one.
            pkw = dict(kw, schedule=SCHEDULES["plain"], model_path=None)

            def time_plain(extra):
                for _ in range(min(args.warmup, 1)):
                    dg.order()
                    dg.count(**pkw, **extra)
                ms_, kern_, last = [], [], None
                for _ in range(args.steps):
                    t = time.perf_counter()
                    dg.order()
                    last = dg.count(**pkw, **extra)
                    ms_.append((time.perf_counter() - t) * 1e3)
                    kern_.append(last["t_count_ms"])
                    if (last["count"], last["hash"]) != (first["count"], first["hash"]):
                        raise SystemExit("plain and learned schedules disagree")
                return float(np.mean(ms_)), float(np.mean(kern_)), last

            plain_ms, plain_kern, rp_ = time_plain({})
            plain = {"value": rp_["count"] / (plain_ms / 1e3), "ms_per_step": plain_ms,
                     "enum_kernel_ms": plain_kern, "schedule": "plain",
                     "worklist": {1: "nowl", 2: "wl"}[rp_["worklist_used"]]}
            if rp_["worklist_used"] != first["worklist_used"]:
                sv_ms, sv_kern, _ = time_plain({"worklist": first["worklist_used"]})
                plain["same_variant"] = {"worklist": {1: "nowl", 2: "wl"}[first["worklist_used"]],
                                         "ms_per_step": sv_ms, "enum_kernel_ms": sv_kern,
                                         "speedup_learned": sv_ms / float(np.mean(step_ms))}
            # per-subproblem tree sizes score the cost model's ranking (untimed)
            lab = dg.count(**pkw, sub_nodes=True)
            F = dg.features()
            roots = F[:, 1] > 0
            y = np.log1p(lab["sub_nodes"].astype(np.float64))[roots]
            plain["spearman_out_degree"] = spearman(F[roots, 1], y)
            if model_path and os.path.exists(model_path):
                pred = mce.mlp_predict(mce.mlp_load(model_path), F[roots])
                plain["spearman_model"] = spearman(pred, y)
            log(f"[bench] plain engine: {plain_ms:.1f} ms ({plain})")
    count = first["count"]
    ms = float(np.mean(step_ms))
    value = count / (ms / 1e3)
    algo_b = algorithmic_bytes(g, split, ipx=args.induced == 2)
    kavg = float(np.mean(kern_ms))
    achieved = algo_b / (kavg / 1e3) / 1e9
    traffic, valu = None, None
    if args.config == "c4":
        key = (f"sbm-batch graphs=4096 n=64..4096 blocks=4 p_in=0.3 p_out=0.01 seed={args.seed} "
               f"level={args.level} induced={args.induced}")
        t = counter_summary("bk_kernel_sbm4096_traffic.txt", key)
        traffic = float(t["traffic_bytes"]) if t else None
        v = counter_summary("bk_kernel_sbm4096_valu_lds.txt", key)
        if v:
            # the kernel's binding resource: VALU issue (bitwise set algebra)
            valu = {"achieved": float(v["valu_rate_ginst_s"]), "peak": float(v["valu_peak_ginst_s"]),
                    "unit": "G wave64 VALU inst/s", "frac": float(v["valu_issue_frac"]),
                    "busy_frac": float(v["valu_busy_frac"]), "insts_per_node": float(v["valu_insts_per_node"]),
                    "lds_bank_conflict_frac": float(v["lds_bank_conflict_frac"]),
                    "source": "profiles/bk_kernel_sbm4096_valu_lds.txt"}
    per_graph = None
    if off is not None:
        sc = first["sub_cliques"].astype(np.int64)
        per_graph = np.add.reduceat(sc, off[:-1].astype(np.int64)) if g.n else np.zeros(0)
        assert int(per_graph.sum()) == count
    out = {
        "metric": METRIC, "value": value, "unit": "maximal cliques/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": dict(desc, level=f"L{args.level}", induced="IPX" if args.induced == 2 else "IP",
                       worklist={1: "nowl", 2: "wl", 3: "auto"}[args.worklist], schedule=args.schedule,
                       parallelism="1 GPU"),
        "count": count, "hash": f"{first['hash']:#018x}", "hist": first["hist"], "degeneracy": first["degeneracy"],
        "tree_nodes": first["tree_nodes"], "enum_kernel_ms": kavg, "order_orient_ms": float(np.mean(order_ms)),
        "fallback_events": first["fallback_events"], "waves": first["num_waves"], "donations": first["donations"],
        "worklist_used": {1: "nowl", 2: "wl"}[first["worklist_used"]],
        "plain_engine": plain, "speedup_vs_plain": (plain["ms_per_step"] / ms) if plain else None,
        "autotune": tuned,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": algo_b, "kernel": "bk_kernel (enumeration)", "valu": valu},
    }
    if not args.no_cpu:
        oracle = _import("mce_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
        sg, order, what = sample_workload(mce, g, off, args.cpu_sample)
        # the host share this process may use (OMP_NUM_THREADS where the host sets it)
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1
        t = time.perf_counter()
        ref = oracle.mce(sg, level=args.level, ipx=args.induced == 2, threads=threads, order=order)
        cpu_s = time.perf_counter() - t
        with mce.DeviceGraph(sg) as dgs:
            dgs.order()
            gpu_s = dgs.count(order=order, **kw) if order is not None else dgs.count(**kw)
        ok = (gpu_s["count"], gpu_s["hash"], gpu_s["hist"]) == (ref["count"], ref["hash"], ref["hist"])
        log(f"[bench] cpu sample ({what}): {ref['count']} cliques in {cpu_s:.2f}s on {threads} threads; "
            f"GPU parity on the sample: {'ok' if ok else 'MISMATCH'}")
        if not ok:
            raise SystemExit("GPU result differs from the oracle on the CPU sample")
        out["cpu_baseline"] = {"value": ref["count"] / cpu_s, "unit": "maximal cliques/s", "cores": threads,
                               "kind": "port", "sample": f"{args.config} {what}: {ref['count']} cliques, "
                                                         f"order+orient+count timed together"}
        out["parity_sample"] = {"count": ref["count"], "hash": f"{ref['hash']:#018x}", "bit_exact": ok}
        # The full roofline t_roof = max(B_hbm / HBM peak, OPS / VALU peak,
        # B_lds / LDS peak). OPS and B_lds are the algorithmic work of the
        # deterministic BK tree (per node: 2 32-bit lane-ops for the AND and
        # the POPCNT of every 64-bit word of every candidate pivot row, 4 words
        # of set updates per child; 8 bytes of LDS per word read), counted by
        # the oracle on the same sample (untimed) and scaled by the GPU's tree
        # nodes; the VALU and LDS peaks are measured on this device.
        wk = oracle.mce(sg, level=args.level, ipx=args.induced == 2, threads=threads, order=order, work=True)
        peaks = mce.device_peaks()
        ops = wk["ops"] / max(wk["tree_nodes"], 1) * first["tree_nodes"]
        b_lds = wk["b_lds"] / max(wk["tree_nodes"], 1) * first["tree_nodes"]
        parts = {"hbm": algo_b / (HBM_PEAK_GBS * 1e9), "valu": ops / peaks["valu_lane_ops_per_s"],
                 "lds": b_lds / peaks["lds_bytes_per_s"]}
        t_roof = max(parts.values())
        out["roofline"]["t_roof"] = {
            "seconds": parts, "binding": max(parts, key=parts.get), "t_roof_s": t_roof,
            "t_kernel_s": kavg / 1e3, "frac": t_roof / (kavg / 1e3), "ops": ops, "lds_bytes": b_lds,
            "hbm_bytes": algo_b, "ops_per_node": wk["ops"] / max(wk["tree_nodes"], 1),
            "lds_bytes_per_node": wk["b_lds"] / max(wk["tree_nodes"], 1),
            "valu_peak_lane_ops_per_s": peaks["valu_lane_ops_per_s"], "lds_peak_bytes_per_s": peaks["lds_bytes_per_s"],
            "source": "oracle work counters on the CPU sample, scaled by tree nodes; peaks from mce_device_peaks"}
        log(f"[bench] roofline: {parts} -> binding {out['roofline']['t_roof']['binding']}, "
            f"t_roof/t_kernel = {out['roofline']['t_roof']['frac']:.4f}")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
