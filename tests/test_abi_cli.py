"""The C ABI library loads and exports every symbol include/mce.h declares;
the parallel_mce_on_gpus CLI keeps the reference's argument rules. No GPU
compute is done here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mce.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(mce_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_exports_every_declared_symbol(mce):
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(mce.LIB_PATH)
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(mce.SIGNATURES)


def test_struct_layout_matches_header(mce, tmp_path):
    """sizeof/offsetof of the ABI structs in C equal the ctypes mirror."""
    src = tmp_path / "probe.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "mce.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(mce_csr), sizeof(mce_batch),"
        " sizeof(mce_opts), sizeof(mce_result), sizeof(mce_wave_stat), offsetof(mce_result, wave_stats),"
        " offsetof(mce_opts, sub_cliques));return 0;}\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    want = [ctypes.sizeof(mce.Csr), ctypes.sizeof(mce.Batch), ctypes.sizeof(mce.Opts), ctypes.sizeof(mce.Result),
            ctypes.sizeof(mce.WaveStat), mce.Result.wave_stats.offset, mce.Opts.sub_cliques.offset]
    assert got == want


def test_version_errors_defaults(mce):
    assert mce.lib.mce_abi_version() == 1
    assert mce.lib.mce_strerror(-1) == b"invalid argument"
    assert mce.lib.mce_strerror(-5) == b"engine limit exceeded"
    o = mce.default_opts()
    assert (o.level, o.induced, o.worklist, o.task, o.schedule, o.donate_min_p) == (1, 2, 1, 0, 0, 10)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-GPU error path")
def test_no_gpu_is_an_error_not_a_crash(mce):
    assert mce.device_count() == 0
    g = mce.gen_er(50, 100, 1)
    with pytest.raises(mce.MCEError) as e:
        mce.DeviceGraph(g)
    assert e.value.code == -4


def cli(*args):
    return subprocess.run([os.path.join(ROOT, "mce-gpu_amd", "bin", "parallel_mce_on_gpus"), *args],
                          capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args,msg", [
    ([], "Must specify -m option"),
    (["-m", "mce"], "Must specify -g option"),
    (["-m", "mce", "-g", "x.bel"], "Must specify -d option"),
    (["-m", "mce", "-g", "x.bel", "-d", "0"], "Must specify -p option"),
    (["-m", "mce", "-g", "x.bel", "-d", "0", "-p", "l1"], "Must specify -i option"),
    (["-m", "mce", "-g", "x.bel", "-d", "0", "-p", "l1", "-i", "p"], "Must specify -w option"),
    (["-m", "mce-bd-eval", "-g", "x.bel", "-d", "0", "-p", "l1", "-i", "p", "-w", "nowl"], "No breakdown"),
    (["-m", "mce-donor-eval", "-g", "x.bel", "-d", "0", "-p", "l1", "-i", "p", "-w", "nowl"], "No donoation"),
    (["-m", "mce-lb-eval", "-g", "x.bel", "-d", "0,1", "-p", "l1", "-i", "p", "-w", "wl"], "only one GPU"),
    (["-m", "mce", "-g", "x.bel", "-d", "0,1", "-p", "l1", "-i", "p", "-w", "wl"], "exactly one GPU"),
    (["-m", "bogus"], "Unrecognized -m option"),
    (["-m", "mce", "-p", "l3"], "Unrecognized -p option"),
    (["-m", "mce", "-i", "q"], "Unrecognized -i option"),
    (["-m", "mce", "-w", "xx"], "Unrecognized -w option"),
    (["-m", "convert", "-g", "in.txt"], "Must specify -r option"),
])
def test_cli_argument_rules(args, msg):
    r = cli(*args)
    assert r.returncode == 0  # the reference exits with status 0 on argument errors
    assert msg in r.stderr


def test_cli_auto_variant_and_engine_options():
    # -w auto and the engine's long options parse; the run then stops at the
    # missing graph file, not at the options
    r = cli("-m", "mce", "-g", "/nonexistent.bel", "-d", "0", "-p", "l1", "-i", "px", "-w", "auto",
            "--donate-age", "5000", "--waves-per-group", "8", "--schedule", "learned", "--small-universe", "32")
    assert r.returncode == 0
    assert "Unrecognized" not in r.stderr and "Must specify" not in r.stderr
    assert "Worker List Scheme = Auto (cost model)" in r.stdout
    assert "unable to open /nonexistent.bel" in r.stdout + r.stderr


def test_cli_help_and_convert(tmp_path):
    r = cli("-h")
    assert r.returncode == 0 and "Usage:" in r.stderr and "--lds-arena" in r.stderr and "--donate-age" in r.stderr
    assert "--small-universe" in r.stderr
    txt = tmp_path / "e.txt"
    txt.write_text("0 1\n1 2\n2 0\n2 3\n")
    out = tmp_path / "e.bel"
    r = cli("-m", "convert", "-g", str(txt), "-r", str(out))
    assert r.returncode == 0
    assert "Destination Graph = " in r.stdout and "n = 4 and m = 4" in r.stdout
    assert out.stat().st_size == 24 * 8
