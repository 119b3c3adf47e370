"""Host graph layer of libmce.so (no GPU): .bel IO with the reference layout,
text conversion, CSR validation, and the seeded generators."""
import os
import struct

import numpy as np
import pytest


def test_bel_layout_and_roundtrip(mce, tmp_path):
    g = mce.gen_er(200, 900, 5)
    p = str(tmp_path / "g.bel")
    g.write_bel(p)
    raw = open(p, "rb").read()
    assert len(raw) == 24 * g.m
    # record = {dst u64, src u64, 0 u64}, sorted by (src, dst) (reference fileop.h:66-68)
    recs = np.frombuffer(raw, dtype=np.uint64).reshape(-1, 3)
    assert np.all(recs[:, 2] == 0)
    src = np.repeat(np.arange(g.n, dtype=np.uint64), np.diff(g.row_ptr).astype(np.int64))
    np.testing.assert_array_equal(recs[:, 1], src)
    np.testing.assert_array_equal(recs[:, 0], g.col_ind.astype(np.uint64))
    h = mce.read_bel(p)
    assert (h.n, h.m) == (g.n, g.m) and h.digest() == g.digest()


def test_read_bel_errors(mce, tmp_path):
    with pytest.raises(mce.MCEError) as e:
        mce.read_bel(str(tmp_path / "missing.bel"))
    assert e.value.code == -2
    p = tmp_path / "trunc.bel"
    p.write_bytes(b"\0" * 25)
    with pytest.raises(mce.MCEError) as e:
        mce.read_bel(str(p))
    assert e.value.code == -2
    # an asymmetric edge list is refused
    q = tmp_path / "asym.bel"
    q.write_bytes(struct.pack("<3Q", 1, 0, 0))
    with pytest.raises(mce.MCEError) as e:
        mce.read_bel(str(q))
    assert e.value.code == -6


def test_convert_text(mce, tmp_path):
    # comments skipped, self loops dropped, both directions, duplicates merged
    txt = tmp_path / "g.txt"
    txt.write_text("# comment\n% other\n0 1\n1 0\n2 2\n1 3\n3 1\n5 4\n")
    bel = str(tmp_path / "g.bel")
    n, m = mce.convert_text_to_bel(str(txt), bel)
    assert (n, m) == (6, 3)
    g = mce.read_bel(bel)
    assert g.n == 6 and g.m == 6
    assert g.col_ind[g.row_ptr[1]:g.row_ptr[2]].tolist() == [0, 3]


def test_validate(mce):
    g = mce.gen_er(100, 300, 1)
    g.validate()
    bad = mce.Graph(g.row_ptr.copy(), g.col_ind.copy())
    r = int(g.row_ptr[10])
    bad.col_ind[r], bad.col_ind[r + 1] = bad.col_ind[r + 1], bad.col_ind[r]
    with pytest.raises(mce.MCEError):
        bad.validate()


def test_er_generator(mce):
    g = mce.gen_er(4096, 65536, 1)
    assert g.n == 4096 and g.m == 2 * 65536
    g.validate()
    assert mce.gen_er(4096, 65536, 1).digest() == g.digest()
    assert mce.gen_er(4096, 65536, 2).digest() != g.digest()


def test_rmat_generator(mce):
    g = mce.gen_rmat(12, 16, seed=1)
    g.validate()
    assert g.n == 4096 and g.m <= 2 * 16 * 4096
    assert mce.gen_rmat(12, 16, seed=1).digest() == g.digest()
    deg = g.degrees()
    # (a, b, c, d) = (.57, .19, .19, .05) without permutation: vertex 0 is the hub
    assert deg.argmax() == 0 and deg[0] > 20 * deg.mean()
    # quadrant masses: the low half of the ids holds most arcs
    lo = g.row_ptr[2048]
    assert lo > 0.65 * g.m


def test_chung_lu_generator(mce):
    n = 1 << 14
    g = mce.gen_chung_lu(n, 2.1, 24.0, 20, 20, 30, seed=3)
    g.validate()
    deg = g.degrees()
    assert 16 < deg.mean() < 26
    assert deg.max() <= int(np.sqrt(24 * n) * 3)
    assert mce.gen_chung_lu(n, 2.1, 24.0, 20, 20, 30, seed=3).digest() == g.digest()


def test_sbm_batch_generator(mce):
    off, g = mce.gen_sbm_batch(16, 64, 256, 4, 0.3, 0.01, seed=1)
    g.validate()
    assert off.size == 17 and off[0] == 0 and off[-1] == g.n
    sizes = np.diff(off)
    assert sizes.min() >= 64 and sizes.max() <= 256
    src = np.repeat(np.arange(g.n), np.diff(g.row_ptr).astype(np.int64))
    gs = np.searchsorted(off, src, side="right") - 1
    gd = np.searchsorted(off, g.col_ind, side="right") - 1
    assert np.all(gs == gd)  # no arc crosses graphs
    # block densities near p_in / p_out
    inb = outb = pin = pout = 0
    for k in range(16):
        nk, b = int(sizes[k]), int(off[k])
        blk = (np.arange(nk) * 4) // nk
        same = blk[:, None] == blk[None, :]
        pin += (same.sum() - nk) / 2
        pout += (~same).sum() / 2
        mask = gs == k
        s, d = src[mask] - b, g.col_ind[mask].astype(np.int64) - b
        e = same[s, d]
        inb += e.sum() / 2
        outb += (~e).sum() / 2
    assert abs(inb / pin - 0.3) < 0.02 and abs(outb / pout - 0.01) < 0.003


def test_from_edges_normalizes(mce):
    g = mce.from_edges([0, 1, 1, 2, 3], [1, 0, 1, 0, 2], n=5)
    assert g.m == 6
    g.validate()


def test_config_digests(mce):
    """The benchmark workloads are pinned: the generators' CSR digests for
    the Erdos-Renyi, R-MAT scale 20 and full-size Chung-Lu (evaluation seed 9)
    configurations equal those committed in tests/golden/configs.json (the
    SBM batch and R-MAT scale 24 are checked where they are generated: the
    GPU tests and bench.py)."""
    import json
    import os
    cfg = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "configs.json")))
    for name in ("c1", "c2", "c3_seed9"):
        fx = cfg[name]
        g = getattr(mce, fx["generator"])(**fx["args"])
        assert (g.n, g.m, f"{g.digest():#018x}") == (fx["n"], fx["m"], fx["digest"]), name
