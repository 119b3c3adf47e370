"""The drop-in process boundary on the GPU: the parallel_mce_on_gpus binary on
.bel files, driven through the -p/-i/-w matrix by the run.py-compatible
sweep (mce-gpu_amd/sweep.py), its stdout parsed with the reference
harness's patterns (run.py:28-129), and the counts compared with the oracle.
The evaluation tasks print one row per wave; their totals must agree with
the engine's own counters."""
import importlib.util
import os
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_sweep():
    spec = importlib.util.spec_from_file_location("mce_sweep", os.path.join(ROOT, "mce-gpu_amd", "sweep.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def graphs(mce, tmp_path_factory):
    d = tmp_path_factory.mktemp("bel")
    out = {}
    for name, g in (("rmat12", mce.gen_rmat(12, 16, seed=3)), ("er4096", mce.gen_er(4096, 65536, 1))):
        path = str(d / f"{name}.bel")
        g.write_bel(path)
        out[name] = (path, mce.read_bel(path))
    return out


def test_sweep_mce_matrix_counts(mce, oracle, graphs, tmp_path):
    sw = load_sweep()
    for name, (path, g) in graphs.items():
        ref = oracle.mce(g, level=1, ipx=True)
        lines = []
        print(f"case cli sweep mce {name}", flush=True)
        outs = sw.sweep(path, str(tmp_path / "mce" / name), "mce", 0, [], echo=lines.append)
        assert len(outs) == 8
        for (p, i, w), text in outs.items():
            scheme = sw.scheme_of(text)
            want = (f"{p.upper()} trees + {'IP ' if i == 'p' else 'IPX'} + "
                    f"{'No Worker List' if w == 'nowl' else 'Worker List'}")
            assert scheme == want
            assert sw.total_count(text) == ref["count"], (name, p, i, w)
            assert len(re.findall(r"Degeneracy ordering time: (.*) s", text)) == 1
            assert len(re.findall(r"count time (.*) s", text)) == 1
            assert re.findall(r"Max Core = (\d+)", text) == [str(ref["degeneracy"])]
            assert re.findall(r"Clique hash = 0x([0-9a-f]+)", text) == [f"{ref['hash']:016x}"]
        # one table row per run, each with the count
        rows = [ln for ln in lines if ln.startswith("| L")]
        assert len(rows) == 8 and all(ln.rstrip(" |").endswith(str(ref["count"])) for ln in rows)


def test_sweep_eval_tasks(mce, oracle, graphs, tmp_path):
    sw = load_sweep()
    path, g = graphs["rmat12"]
    ref = oracle.mce(g, level=1, ipx=True)
    nodes = {"p": oracle.mce(g, level=1, ipx=False)["tree_nodes"], "px": ref["tree_nodes"]}
    for task in ("mce-lb-eval", "mce-donor-eval", "mce-bd-eval"):
        print(f"case cli sweep {task}", flush=True)
        lines = []
        outs = sw.sweep(path, str(tmp_path / task), task, 0, [], echo=lines.append)
        assert len(outs) == (8 if task == "mce-lb-eval" else 4)
        for key, text in outs.items():
            assert sw.total_count(text) == ref["count"], (task, key)
            k = 7 if task == "mce-bd-eval" else 3
            rows = sw.rows(text.split("Finished Launching Instance 0.")[1].split("Instance 0 found")[0], k)
            assert len(rows) > 0
            if task == "mce-lb-eval":
                # tree nodes per wave add up to the oracle's level-1 tree of
                # the same induced-subgraph mode
                if key[0] == "l1":
                    assert sum(r[2] for r in rows) == nodes[key[1]], key
            if task == "mce-bd-eval":
                assert all(sum(r) > 0 for r in rows)
        assert all("OOM" not in ln for ln in lines)


def test_cli_learned_schedule_on_bel(mce, oracle, graphs, tmp_path):
    sw = load_sweep()
    path, g = graphs["rmat12"]
    ref = oracle.mce(g, level=1, ipx=True)
    model = os.path.join(ROOT, "mce-gpu_amd", "cost_model_sbm.mlp")
    outs = sw.sweep(path, str(tmp_path / "learned"), "mce", 0, ["--schedule", "learned", "--model", model],
                    echo=lambda s: None)
    for key, text in outs.items():
        assert sw.total_count(text) == ref["count"], key
