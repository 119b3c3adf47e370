"""ctypes wrapper of the CPU oracle (liboracle.so) — test infrastructure only.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product. See mce_oracle.cpp for what it restates and how each
function is pinned.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
HIST = 4098


class OracleResult(ctypes.Structure):
    _fields_ = [
        ("count", c_uint64), ("hash", c_uint64), ("nodes", c_uint64), ("hist", c_uint64 * HIST),
        ("max_clique_size", c_uint32), ("degeneracy", c_uint32), ("max_degree", c_uint32),
        ("max_out_degree", c_uint32), ("t_order_ms", c_double), ("t_count_ms", c_double),
        ("b_hbm", c_double), ("ops", c_double), ("b_lds", c_double), ("threads", c_uint32),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"oracle not built at {LIB_PATH}")
    lib = ctypes.CDLL(LIB_PATH)
    lib.oracle_mce.restype = c_int
    lib.oracle_mce.argtypes = [c_uint32, c_void_p, c_void_p, c_int, c_int, c_void_p, c_uint64, c_int, c_int,
                               c_void_p, c_void_p, POINTER(OracleResult)]
    lib.oracle_peel.restype = c_uint32
    lib.oracle_peel.argtypes = [c_uint32, c_void_p, c_void_p, c_void_p, c_void_p]
    lib.oracle_core_numbers.restype = c_uint32
    lib.oracle_core_numbers.argtypes = [c_uint32, c_void_p, c_void_p, c_void_p]
    lib.oracle_knuth.restype = c_int
    lib.oracle_knuth.argtypes = [c_uint32, c_void_p, c_void_p, c_int, c_void_p, c_uint64, c_uint32, c_uint64, c_int,
                                 c_void_p, c_void_p]
    lib.oracle_order_orient.restype = c_uint32
    lib.oracle_order_orient.argtypes = [c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    return lib


lib = _load()


def _p(a: np.ndarray):
    return a.ctypes.data_as(c_void_p)


def mce(g, level: int = 1, ipx: bool = True, order=None, threads: int = 0, work: bool = False,
        sub_nodes: bool = False, sub_cliques: bool = False) -> dict:
    n = g.n
    r = OracleResult()
    ordp, olen, keep = None, 0, None
    if order is not None:
        keep = np.ascontiguousarray(order, dtype=np.uint32)
        ordp, olen = _p(keep), keep.size
    sn = np.zeros(max(n, 1), np.uint64) if sub_nodes else None
    sc = np.zeros(max(n, 1), np.uint64) if sub_cliques else None
    rc = lib.oracle_mce(n, _p(g.row_ptr), _p(g.col_ind), level, 1 if ipx else 0, ordp, olen, threads,
                        1 if work else 0, _p(sn) if sn is not None else None, _p(sc) if sc is not None else None,
                        byref(r))
    if rc != 0:
        raise RuntimeError(f"oracle_mce failed with {rc}")
    hist = {k: int(r.hist[k]) for k in range(HIST) if r.hist[k]}
    out = dict(count=int(r.count), hash=int(r.hash), hist=hist, tree_nodes=int(r.nodes),
               max_clique_size=int(r.max_clique_size), degeneracy=int(r.degeneracy), max_degree=int(r.max_degree),
               max_out_degree=int(r.max_out_degree), t_order_ms=float(r.t_order_ms), t_count_ms=float(r.t_count_ms),
               b_hbm=float(r.b_hbm), ops=float(r.ops), b_lds=float(r.b_lds), threads=int(r.threads))
    if sn is not None:
        out["sub_nodes"] = sn[:n]
    if sc is not None:
        out["sub_cliques"] = sc[:n]
    return out


def peel(g):
    """Batched-peel priorities, degeneracy and round count."""
    n = g.n
    prio = np.zeros(max(n, 1), np.uint32)
    rounds = c_uint32()
    d = lib.oracle_peel(n, _p(g.row_ptr), _p(g.col_ind), _p(prio), byref(rounds))
    return prio[:n], int(d), int(rounds.value)


def core_numbers(g):
    n = g.n
    core = np.zeros(max(n, 1), np.uint32)
    k = lib.oracle_core_numbers(n, _p(g.row_ptr), _p(g.col_ind), _p(core))
    return core[:n], int(k)


def order_orient(g):
    n, m = g.n, g.m
    prio = np.zeros(max(n, 1), np.uint32)
    split = np.zeros(max(n, 1), np.uint64)
    oc = np.zeros(max(m, 1), np.uint32)
    rounds = c_uint32()
    d = lib.oracle_order_orient(n, _p(g.row_ptr), _p(g.col_ind), _p(prio), _p(split), _p(oc), byref(rounds))
    return prio[:n], split[:n], oc[:m], int(d)


def knuth(g, order, probes: int = 64, seed: int = 1, ipx: bool = True, threads: int = 0):
    """Knuth random-probe estimates of the level-1 subtrees of `order`'s
    vertices: (mean maximal cliques, mean tree nodes) per vertex, each an
    unbiased estimate of that subproblem's exact figure."""
    o = np.ascontiguousarray(order, dtype=np.uint32)
    ec = np.zeros(max(o.size, 1), np.float64)
    en = np.zeros(max(o.size, 1), np.float64)
    rc = lib.oracle_knuth(g.n, _p(g.row_ptr), _p(g.col_ind), 1 if ipx else 0, _p(o), o.size, probes, seed, threads,
                          _p(ec), _p(en))
    if rc != 0:
        raise RuntimeError(f"oracle_knuth failed with {rc}")
    return ec[:o.size], en[:o.size]
