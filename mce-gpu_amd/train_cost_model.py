"""Trains the subproblem cost model on data the engine generates.

For each training seed the script builds a graph of the chosen family, runs
the exact enumeration with per-subproblem tree-node labels (one BK tree per
first-level vertex), and pairs the 16 device-computed features of every
vertex with log(1 + tree nodes). The MLP (16 -> 64 -> 64 -> 1) is trained on
the GPU by the engine's own forward/backward/Adam kernels, evaluated on a
held-out seed (Spearman rank correlation of predicted vs. measured tree
size), and written in the engine's model format.

    python mce-gpu_amd/train_cost_model.py --family sbm --train-seeds 2-9 \\
        --eval-seed 1 --out mce-gpu_amd/cost_model_sbm.mlp

Families: sbm (batches of SBM graphs, n in [64, 4096], 4 blocks, p_in .3,
p_out .01; --graphs per batch) and chung-lu (beta 2.1, mean degree 24,
planted cliques of 20..60 vertices, n / 2048 of them as in the 4 M-vertex
configuration; --n and --planted, or --scales for several sizes: every
training seed at every scale, evaluated at --eval-scale).

The committed Chung-Lu model was made with

    python mce-gpu_amd/train_cost_model.py --family chung-lu --scales 18,19,20 \\
        --train-seeds 1-8 --eval-seed 9 --out mce-gpu_amd/cost_model_cl.mlp
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _engine():
    if "mce_gpu_amd" in sys.modules:
        return sys.modules["mce_gpu_amd"]
    spec = importlib.util.spec_from_file_location("mce_gpu_amd", os.path.join(HERE, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mce_gpu_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def make_graph(mce, family, seed, args, n=None):
    if family == "sbm":
        _, g = mce.gen_sbm_batch(args.graphs, 64, 4096, 4, 0.3, 0.01, seed=seed)
        return g
    n = n or args.n
    return mce.gen_chung_lu(n, 2.1, 24.0, args.planted or n // 2048, 20, 60, seed=seed)


def labeled(mce, g, worklist, device=0):
    """Features and log(1 + tree nodes) of every first-level subproblem."""
    with mce.DeviceGraph(g, device=device) as dg:
        dg.order()
        F = dg.features()
        # long-running tasks only give work away: labels do not depend on it
        r = dg.count(level=mce.L1, induced=mce.IPX, worklist=worklist, sub_nodes=True, donate_age_us=10000)
    keep = F[:, 1] > 0  # out-degree >= 1: vertices that root a subproblem
    y = np.log1p(r["sub_nodes"].astype(np.float64)).astype(np.float32)
    return F[keep], y[keep], r


def spearman(a, b):
    ra = np.argsort(np.argsort(a, kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b, kind="stable"), kind="stable").astype(np.float64)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / np.sqrt((ra * ra).sum() * (rb * rb).sum()))


def parse_seeds(text):
    out = []
    for part in text.split(","):
        if "-" in part:
            a, b = part.split("-")
            out.extend(range(int(a), int(b) + 1))
        elif part:
            out.append(int(part))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--family", choices=["sbm", "chung-lu"], default="sbm")
    ap.add_argument("--train-seeds", default="2-9")
    ap.add_argument("--eval-seed", type=int, default=1)
    ap.add_argument("--graphs", type=int, default=256, help="sbm: graphs per batch")
    ap.add_argument("--n", type=int, default=1 << 18, help="chung-lu: vertices")
    ap.add_argument("--planted", type=int, default=0, help="chung-lu: planted cliques (default n / 2048)")
    ap.add_argument("--scales", default="", help="chung-lu: comma-separated log2 sizes to train on (overrides --n)")
    ap.add_argument("--eval-scale", type=int, default=0, help="chung-lu: log2 size of the held-out graph "
                                                                  "(default: the largest training size)")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1, help="initialisation and shuffling")
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    mce = _engine()
    # labels do not depend on the variant; SBM batches have no heavy tail
    wl = mce.NOWL if args.family == "sbm" else mce.WL

    sizes = [1 << int(k) for k in args.scales.split(",") if k] if args.scales else [args.n]
    eval_n = (1 << args.eval_scale) if args.eval_scale else max(sizes)
    Fs, ys = [], []
    t0 = time.time()
    for sd in parse_seeds(args.train_seeds):
        for n in sizes:
            F, y, r = labeled(mce, make_graph(mce, args.family, sd, args, n), wl)
            Fs.append(F)
            ys.append(y)
            print(f"[train] seed {sd} n {n}: {F.shape[0]} subproblems, {r['tree_nodes']} tree nodes",
                  file=sys.stderr, flush=True)
    F = np.concatenate(Fs)
    y = np.concatenate(ys)
    model = mce.mlp_init(F, seed=args.seed)
    loss = mce.mlp_train(model, F, y, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    _, train_mse = mce.mlp_grad(model, F, y)
    Fe, ye, _ = labeled(mce, make_graph(mce, args.family, args.eval_seed, args, eval_n), wl)
    pe = mce.mlp_predict(model, Fe)
    rho = spearman(pe, ye)
    # the plain schedule's implicit predictor: out-degree
    rho_plain = spearman(Fe[:, 1], ye)
    mce.mlp_save(args.out, model)
    print(json.dumps({"family": args.family, "train_seeds": args.train_seeds, "eval_seed": args.eval_seed,
                      "train_sizes": sizes, "eval_n": eval_n,
                      "train_rows": int(F.shape[0]), "last_epoch_loss": float(loss), "train_mse": float(train_mse),
                      "eval_rows": int(Fe.shape[0]), "spearman_model": rho, "spearman_out_degree": rho_plain,
                      "seconds": time.time() - t0, "out": args.out}))


if __name__ == "__main__":
    main()
