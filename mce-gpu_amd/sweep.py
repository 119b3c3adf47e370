"""Sweep driver with the reference harness's interface and output
(reference run.py:8-22 arguments, :28-129 parsing, :131-222 sweeps): runs
parallel_mce_on_gpus over the -p/-i/-w matrix for each graph, keeps each
run's stdout as <p>-i<i>-<w>-1gpu.txt under <output_dir>/<task>/<graph>/, and
prints the same markdown tables.

Differences from the reference harness: any .bel file may be swept
(--graph names a file stem under --input_dir, or 'all' for every .bel
there; the reference lists twelve downloadable graphs), and only one device
id is accepted (this engine uses one GPU). Extra engine options after `--`
are passed to every run, e.g.  -- --schedule learned --model cost_model_cl.mlp

Usage:
  python mce-gpu_amd/sweep.py --graph g --input_dir dir --output_dir out --task mce --devices 0
"""
import argparse
import glob
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "bin", "parallel_mce_on_gpus")

# The reference's scheme order (run.py:136-138 and its siblings).
MATRIX = [(p, i, w) for p in ("l1", "l2") for i in ("p", "px") for w in ("nowl", "wl")]


def scheme_of(text):
    """'L1 trees + IPX + Worker List' as run.py:28-38 builds it."""
    ps = re.findall(r"Parallelization Scheme = (.*)", text)
    ind = re.findall(r"Induced Subgraphs Scheme = (.*)", text)
    ws = re.findall(r"Worker List Scheme = (.*)", text)
    if len(ps) != 1 or len(ind) != 1 or len(ws) != 1:
        raise ValueError("run output lacks the scheme lines")
    return f"{ps[0]} + {ind[0]}{' ' * (3 - len(ind[0]))} + {ws[0]}"


def total_count(text):
    """The 'Found N maximal cliques in total.' count, digits possibly
    comma-grouped (run.py:47-51); None when absent."""
    c = re.findall(r"Found (.*) maximal cliques in total.", text)
    return int(c[0].replace(",", "")) if len(c) == 1 else None


def rows(text, k):
    """Per-wave rows: k comma-separated integers (run.py:74, :99, :121)."""
    return [tuple(int(x) for x in m) for m in re.findall(r"([0-9]+)" + r", ([0-9]+)" * (k - 1), text)]


def table_mce(scheme, text):
    degen = re.findall(r"Degeneracy ordering time: (.*) s", text)
    count = total_count(text)
    timer = re.findall(r"count time (.*) s", text)
    if len(degen) != 1:
        return f"| {scheme:32} | {'OOM':>15} | {'OOM':>15} | {'OOM':>12} |"
    d = float(degen[0])
    if not count or len(timer) != 1:
        return f"| {scheme:32} | {d:>15.3f} | {'OOM':>15} | {'OOM':>12} |"
    return f"| {scheme:32} | {d:>15.3f} | {float(timer[0]):>15.3f} | {count:>12} |"


def table_lb(scheme, text):
    if not total_count(text):
        return f"| {'Scheme':32} |" + f" {'OOM':>8} |" * 5
    peak = {}
    for _, cu, load in rows(text, 3):
        peak[cu] = max(peak.get(cu, 0), load)
    data = np.array([peak.get(c, 0) for c in range(max(peak) + 1)], dtype=float) if peak else np.zeros(1)
    data /= max(data.mean(), 1e-300)
    q = [np.percentile(data, x) for x in (0, 25, 50, 75, 100)]
    return f"| {scheme:32} |" + "".join(f" {v:>8.6f} |" for v in q)


def table_donor(scheme, text):
    if not total_count(text):
        return f"| {'Scheme':32} | {'OOM':>19} |"
    return f"| {scheme:32} | {sum(r[2] for r in rows(text, 3)):>19} |"


def table_bd(scheme, text):
    if not total_count(text):
        return f"| {scheme:32} |" + "".join(f" {'OOM':>{w}} |" for w in (26, 11, 16, 14, 18, 26, 5))
    tot = np.zeros(7)
    for r in rows(text, 7):
        tot += np.array(r, dtype=float)
    tot /= max(tot.sum(), 1e-300)
    cols = [tot[1], tot[2], tot[3], tot[4], tot[5], tot[6], tot[0]]
    return f"| {scheme:32} |" + "".join(f" {v:>{w}.1%} |" for v, w in zip(cols, (26, 11, 16, 14, 18, 26, 5)))


HEADERS = {
    "mce": (f"| {'Scheme':32} | {'Degen Time (s)':15} | {'Count Time (s)':15} | {'MCE Count':12} |", table_mce),
    "mce-lb-eval": (f"| {'Scheme':32} |" + "".join(f" {h:>8} |" for h in ("0%", "25%", "50%", "75%", "100%")),
                    table_lb),
    "mce-donor-eval": (f"| {'Scheme':32} | {'Number of Donations':>19} |", table_donor),
    "mce-bd-eval": (f"| {'Scheme':32} | {'Building Induced Subgraphs':>26} | {'Worker List':>11} | "
                    f"{'Selecting Pivots':>16} | {'Set Operations':>14} | {'Testing Maximality':>18} | "
                    f"{'Branching and Backtracking':>26} | {'Other':>5} | ", table_bd),
}


def sweep(graph_path, out_dir, task, device, extra, exe=EXE, echo=print):
    """Runs the matrix on one .bel; returns {(p, i, w): stdout text}."""
    os.makedirs(out_dir, exist_ok=True)
    header, fmt = HEADERS[task]
    echo("")
    echo("Graph: " + os.path.splitext(os.path.basename(graph_path))[0])
    echo(header)
    outs = {}
    for p, i, w in MATRIX:
        if task in ("mce-bd-eval", "mce-donor-eval") and w == "nowl":
            continue  # no breakdown or donation without the worker list (include/config.h:255-264)
        fp = os.path.join(out_dir, f"{p}-i{i}-{w}-1gpu.txt")
        with open(fp, "w") as f:
            subprocess.run([exe, "-m", task, "-d", str(device), "-p", p, "-i", i, "-w", w, "-g", graph_path] + extra,
                           stdout=f, check=False)
        text = open(fp).read()
        outs[(p, i, w)] = text
        try:
            echo(fmt(scheme_of(text), text))
        except ValueError:
            echo(f"| {p + ' ' + i + ' ' + w:32} | failed, see {fp} |")
    return outs


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    extra = []
    if "--" in argv:
        k = argv.index("--")
        argv, extra = argv[:k], argv[k + 1:]
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--graph", required=True, help="graph file stem under --input_dir, or 'all'")
    ap.add_argument("--input_dir", required=True)
    ap.add_argument("--output_dir", required=True)
    ap.add_argument("--task", required=True, choices=list(HEADERS))
    ap.add_argument("--devices", required=True, help="one device id")
    args = ap.parse_args(argv)
    print(args)
    if "," in args.devices:
        raise SystemExit("this engine uses exactly one GPU: give one device id")
    if not os.path.exists(EXE):
        raise SystemExit(f"{EXE} is not built")
    names = ([os.path.splitext(os.path.basename(p))[0] for p in sorted(glob.glob(os.path.join(args.input_dir, "*.bel")))]
             if args.graph == "all" else [args.graph])
    for name in names:
        sweep(os.path.join(args.input_dir, name + ".bel"), os.path.join(args.output_dir, args.task, name),
              args.task, int(args.devices), extra)


if __name__ == "__main__":
    main()
