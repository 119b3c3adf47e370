"""Python binding of libmce.so (include/mce.h) through ctypes.

The engine itself is native: host C++ and hand-written HIP kernels for gfx950
behind a C ABI. This module only marshals arguments. It mirrors the reference
pipeline's surface (yen-hsiang-chang/mce-gpu src/main.cu:18-187): read a .bel
graph, order it, count maximal cliques with a chosen -p/-i/-w variant.

There is no fallback: if libmce.so is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmce.so")
CLI_PATH = os.path.join(_HERE, "bin", "parallel_mce_on_gpus")

L1, L2 = 1, 2
IP, IPX = 1, 2
NOWL, WL, WL_AUTO = 1, 2, 3
TASK_MCE, TASK_LB_EVAL, TASK_BD_EVAL, TASK_DONOR_EVAL = 0, 1, 2, 3
SCHED_PLAIN, SCHED_LEARNED, SCHED_GIVEN = 0, 1, 2
MAX_OUT_DEGREE = 4096
HIST_LEN = MAX_OUT_DEGREE + 2

ERRORS = {
    0: "MCE_OK", -1: "MCE_EINVAL", -2: "MCE_EIO", -3: "MCE_ENOMEM", -4: "MCE_EDEVICE",
    -5: "MCE_ELIMIT", -6: "MCE_EGRAPH", -7: "MCE_EMODEL", -8: "MCE_EINTERNAL",
}


class MCEError(RuntimeError):
    def __init__(self, code: int, detail: str):
        super().__init__(f"{ERRORS.get(code, code)}: {detail}")
        self.code = code
        self.detail = detail


class Csr(ctypes.Structure):
    _fields_ = [("n", c_uint32), ("m", c_uint64), ("row_ptr", POINTER(c_uint64)), ("col_ind", POINTER(c_uint32))]


class Batch(ctypes.Structure):
    _fields_ = [("num_graphs", c_uint32), ("vtx_off", POINTER(c_uint32)), ("csr", Csr)]


MLP_FEATURES, MLP_HIDDEN = 16, 64
MLP_PARAMS = 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1


class Mlp(ctypes.Structure):
    _fields_ = [("mu", ctypes.c_float * MLP_FEATURES), ("sigma", ctypes.c_float * MLP_FEATURES),
                ("params", ctypes.c_float * MLP_PARAMS)]


class TrainOpts(ctypes.Structure):
    _fields_ = [("device", c_int), ("epochs", c_int), ("batch", c_int), ("lr", ctypes.c_float),
                ("weight_decay", ctypes.c_float), ("seed", c_uint64)]


class Opts(ctypes.Structure):
    _fields_ = [
        ("level", c_int), ("induced", c_int), ("worklist", c_int), ("task", c_int), ("device", c_int),
        ("waves_per_group", c_int), ("groups_per_cu", c_int), ("lds_arena_bytes", c_int),
        ("hbm_arena_bytes", c_int), ("stack_lds_levels", c_int), ("donate_min_p", c_int),
        ("donate_check_nodes", c_int), ("donate_age_us", c_int), ("split_heavy", c_int), ("graph_launch", c_int), ("small_universe", c_int),
        ("watchdog_us", c_int), ("progress_s", c_int), ("lane_universe", c_int), ("schedule", c_int), ("order", POINTER(c_uint32)),
        ("order_len", c_uint64), ("model_path", c_char_p), ("model", POINTER(Mlp)), ("sub_nodes", POINTER(c_uint64)),
        ("sub_cliques", POINTER(c_uint64)),
    ]


class WaveStat(ctypes.Structure):
    _fields_ = [("wave", c_uint32), ("cu", c_uint32), ("value", c_uint64), ("bd", c_uint64 * 7)]


class Result(ctypes.Structure):
    _fields_ = [
        ("count", c_uint64), ("hash", c_uint64), ("hist", POINTER(c_uint64)), ("hist_len", c_uint32),
        ("max_clique_size", c_uint32), ("degeneracy", c_uint32), ("max_degree", c_uint32),
        ("max_out_degree", c_uint32), ("num_waves", c_uint32), ("worklist_used", c_uint32), ("tree_nodes", c_uint64),
        ("donations", c_uint64), ("fallback_events", c_uint64), ("split_vertices", c_uint64), ("t_order_ms", c_double),
        ("t_orient_ms", c_double), ("t_count_ms", c_double), ("wave_stats", POINTER(WaveStat)),
        ("wave_stats_cap", c_uint32), ("wave_stats_len", c_uint32),
    ]


# Every exported symbol of include/mce.h with (restype, argtypes).
SIGNATURES = {
    "mce_abi_version": (c_int, []),
    "mce_strerror": (c_char_p, [c_int]),
    "mce_last_error": (c_char_p, []),
    "mce_opts_default": (None, [POINTER(Opts)]),
    "mce_device_count": (c_int, []),
    "mce_device_peaks": (c_int, [c_int, POINTER(c_double), POINTER(c_double)]),
    "mce_csr_read_bel": (c_int, [c_char_p, POINTER(Csr)]),
    "mce_csr_write_bel": (c_int, [c_char_p, POINTER(Csr)]),
    "mce_convert_text_to_bel": (c_int, [c_char_p, c_char_p, POINTER(c_uint64), POINTER(c_uint64)]),
    "mce_csr_from_edges": (c_int, [c_uint32, POINTER(c_uint32), POINTER(c_uint32), c_uint64, POINTER(Csr)]),
    "mce_csr_validate": (c_int, [POINTER(Csr)]),
    "mce_csr_free": (None, [POINTER(Csr)]),
    "mce_batch_free": (None, [POINTER(Batch)]),
    "mce_gen_er": (c_int, [c_uint32, c_uint64, c_uint64, POINTER(Csr)]),
    "mce_gen_rmat": (c_int, [c_uint32, c_uint32, c_double, c_double, c_double, c_uint64, POINTER(Csr)]),
    "mce_gen_chung_lu": (c_int, [c_uint32, c_double, c_double, c_uint32, c_uint32, c_uint32, c_uint64, POINTER(Csr)]),
    "mce_gen_sbm_batch": (c_int, [c_uint32, c_uint32, c_uint32, c_uint32, c_double, c_double, c_uint64, POINTER(Batch)]),
    "mce_csr_digest": (c_uint64, [POINTER(Csr)]),
    "mce_graph_create": (c_int, [POINTER(Csr), c_int, POINTER(c_void_p)]),
    "mce_graph_order": (c_int, [c_void_p, POINTER(Result)]),
    "mce_graph_count": (c_int, [c_void_p, POINTER(Opts), POINTER(Result)]),
    "mce_graph_count_noalloc": (c_int, [c_void_p, POINTER(Opts), POINTER(Result)]),
    "mce_graph_download": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint64), POINTER(c_uint32)]),
    "mce_graph_destroy": (None, [c_void_p]),
    "mce_graph_features": (c_int, [c_void_p, POINTER(ctypes.c_float)]),
    "mce_mlp_init": (c_int, [POINTER(Mlp), POINTER(ctypes.c_float), c_uint64, c_uint64]),
    "mce_mlp_train": (c_int, [POINTER(Mlp), POINTER(ctypes.c_float), POINTER(ctypes.c_float), c_uint64,
                              POINTER(TrainOpts), POINTER(c_double)]),
    "mce_mlp_grad": (c_int, [POINTER(Mlp), POINTER(ctypes.c_float), POINTER(ctypes.c_float), c_uint64, c_int,
                             POINTER(ctypes.c_float), POINTER(c_double)]),
    "mce_mlp_predict": (c_int, [POINTER(Mlp), POINTER(ctypes.c_float), c_uint64, c_int, POINTER(ctypes.c_float)]),
    "mce_mlp_save": (c_int, [c_char_p, POINTER(Mlp)]),
    "mce_mlp_load": (c_int, [c_char_p, POINTER(Mlp)]),
    "mce_run": (c_int, [POINTER(Csr), POINTER(Opts), POINTER(Result)]),
    "mce_run_batch": (c_int, [POINTER(Batch), POINTER(Opts), POINTER(c_uint64), POINTER(Result)]),
    "mce_autotune": (c_int, [c_void_p, POINTER(Opts), c_double, POINTER(Opts), ctypes.c_char_p, c_uint64]),
}


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libmce.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _check(rc: int) -> None:
    if rc != 0:
        raise MCEError(rc, (lib.mce_last_error() or b"").decode())


def _u64p(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_uint64))


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_uint32))


def _f32p(a: np.ndarray):
    return a.ctypes.data_as(POINTER(ctypes.c_float))


class Graph:
    """Host CSR owned by numpy arrays (row_ptr uint64[n+1], col_ind uint32[m])."""

    def __init__(self, row_ptr: np.ndarray, col_ind: np.ndarray):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        self.col_ind = np.ascontiguousarray(col_ind, dtype=np.uint32)
        if self.row_ptr.size == 0:
            self.row_ptr = np.zeros(1, np.uint64)
        if self.col_ind.size == 0:
            self.col_ind = np.zeros(1, np.uint32)[:0]

    @property
    def n(self) -> int:
        return int(self.row_ptr.size - 1)

    @property
    def m(self) -> int:
        return int(self.row_ptr[-1])

    def csr(self) -> Csr:
        c = Csr()
        c.n = self.n
        c.m = self.m
        c.row_ptr = _u64p(self.row_ptr)
        c.col_ind = _u32p(self.col_ind) if self.col_ind.size else ctypes.cast(0, POINTER(c_uint32))
        return c

    @staticmethod
    def _from_lib_csr(c: Csr) -> "Graph":
        n, m = c.n, c.m
        rp = np.ctypeslib.as_array(c.row_ptr, shape=(n + 1,)).copy() if c.row_ptr else np.zeros(1, np.uint64)
        ci = np.ctypeslib.as_array(c.col_ind, shape=(m,)).copy() if (m and c.col_ind) else np.zeros(0, np.uint32)
        lib.mce_csr_free(byref(c))
        return Graph(rp, ci)

    def digest(self) -> int:
        c = self.csr()
        return int(lib.mce_csr_digest(byref(c)))

    def validate(self) -> None:
        c = self.csr()
        _check(lib.mce_csr_validate(byref(c)))

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr).astype(np.int64)

    def write_bel(self, path: str) -> None:
        c = self.csr()
        _check(lib.mce_csr_write_bel(path.encode(), byref(c)))


def read_bel(path: str) -> Graph:
    c = Csr()
    _check(lib.mce_csr_read_bel(path.encode(), byref(c)))
    return Graph._from_lib_csr(c)


def convert_text_to_bel(text_path: str, bel_path: str) -> tuple[int, int]:
    n, m = c_uint64(), c_uint64()
    _check(lib.mce_convert_text_to_bel(text_path.encode(), bel_path.encode(), byref(n), byref(m)))
    return n.value, m.value


def from_edges(src, dst, n: int = 0) -> Graph:
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    c = Csr()
    _check(lib.mce_csr_from_edges(n, _u32p(s), _u32p(d), s.size, byref(c)))
    return Graph._from_lib_csr(c)


def gen_er(n: int, m: int, seed: int) -> Graph:
    c = Csr()
    _check(lib.mce_gen_er(n, m, seed, byref(c)))
    return Graph._from_lib_csr(c)


def gen_rmat(scale: int, edge_factor: int, a=0.57, b=0.19, c_=0.19, seed: int = 1) -> Graph:
    c = Csr()
    _check(lib.mce_gen_rmat(scale, edge_factor, a, b, c_, seed, byref(c)))
    return Graph._from_lib_csr(c)


def gen_chung_lu(n: int, beta=2.1, mean_degree=24.0, planted=2000, k_min=20, k_max=60, seed: int = 1) -> Graph:
    c = Csr()
    _check(lib.mce_gen_chung_lu(n, beta, mean_degree, planted, k_min, k_max, seed, byref(c)))
    return Graph._from_lib_csr(c)


def gen_sbm_batch(num_graphs: int, n_min=64, n_max=4096, blocks=4, p_in=0.3, p_out=0.01, seed: int = 1):
    b = Batch()
    _check(lib.mce_gen_sbm_batch(num_graphs, n_min, n_max, blocks, p_in, p_out, seed, byref(b)))
    off = np.ctypeslib.as_array(b.vtx_off, shape=(b.num_graphs + 1,)).copy()
    n, m = b.csr.n, b.csr.m
    rp = np.ctypeslib.as_array(b.csr.row_ptr, shape=(n + 1,)).copy()
    ci = np.ctypeslib.as_array(b.csr.col_ind, shape=(m,)).copy() if m else np.zeros(0, np.uint32)
    lib.mce_batch_free(byref(b))
    return off, Graph(rp, ci)


def device_count() -> int:
    return int(lib.mce_device_count())


def device_peaks(device: int = 0) -> dict:
    """Measured VALU lane-op/s and LDS byte/s peaks of the device."""
    v, l = c_double(), c_double()
    _check(lib.mce_device_peaks(device, byref(v), byref(l)))
    return {"valu_lane_ops_per_s": v.value, "lds_bytes_per_s": l.value}


def default_opts(**kw) -> Opts:
    o = Opts()
    lib.mce_opts_default(byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class DeviceGraph:
    """A graph resident in HBM (mce_graph_create/order/count/destroy)."""

    def __init__(self, g: Graph, device: int = 0):
        self._host = g
        self._csr = g.csr()
        h = c_void_p()
        _check(lib.mce_graph_create(byref(self._csr), device, byref(h)))
        self.handle = h
        self.device = device
        self.order_info: dict | None = None

    def order(self) -> dict:
        r = Result()
        _check(lib.mce_graph_order(self.handle, byref(r)))
        self.order_info = dict(degeneracy=r.degeneracy, max_degree=r.max_degree, max_out_degree=r.max_out_degree,
                               t_order_ms=r.t_order_ms, t_orient_ms=r.t_orient_ms)
        return self.order_info

    def download(self, prio=True, split=True, col=True):
        """Ordering results: (priority[n], split[n], oriented col_ind[m]); a
        part not asked for is returned as None and not copied."""
        n, m = self._host.n, self._host.m
        P = np.zeros(max(n, 1), np.uint32) if prio else None
        S = np.zeros(max(n, 1), np.uint64) if split else None
        C = np.zeros(max(m, 1), np.uint32) if col else None
        null32, null64 = ctypes.cast(0, POINTER(c_uint32)), ctypes.cast(0, POINTER(c_uint64))
        _check(lib.mce_graph_download(self.handle, _u32p(P) if prio else null32, _u64p(S) if split else null64,
                                      _u32p(C) if col else null32))
        return (P[:n] if prio else None), (S[:n] if split else None), (C[:m] if col else None)

    def features(self) -> np.ndarray:
        """Cost-model features of every vertex (n x 16 float32); needs order()."""
        n = self._host.n
        F = np.zeros((max(n, 1), MLP_FEATURES), np.float32)
        _check(lib.mce_graph_features(self.handle, _f32p(F)))
        return F[:n]

    def count(self, level=L1, induced=IPX, worklist=NOWL, task=TASK_MCE, order=None, sub_nodes=False,
              sub_cliques=False, wave_stats_cap=0, model=None, model_path=None, **block) -> dict:
        o = default_opts(level=level, induced=induced, worklist=worklist, task=task, device=self.device, **block)
        keep = []
        if model is not None or model_path is not None:
            o.schedule = SCHED_LEARNED
            if model is not None:
                keep.append(model)
                o.model = ctypes.pointer(model)
            if model_path is not None:
                keep.append(model_path.encode())
                o.model_path = keep[-1]
        if order is not None:
            arr = np.ascontiguousarray(order, dtype=np.uint32)
            keep.append(arr)
            o.schedule = SCHED_GIVEN
            o.order = _u32p(arr)
            o.order_len = arr.size
        n = self._host.n
        sn = sc = None
        if sub_nodes:
            sn = np.zeros(max(n, 1), np.uint64)
            o.sub_nodes = _u64p(sn)
        if sub_cliques:
            sc = np.zeros(max(n, 1), np.uint64)
            o.sub_cliques = _u64p(sc)
        r = Result()
        hist = np.zeros(HIST_LEN, np.uint64)
        r.hist = _u64p(hist)
        r.hist_len = HIST_LEN
        ws = None
        if wave_stats_cap:
            ws = (WaveStat * wave_stats_cap)()
            r.wave_stats = ws
            r.wave_stats_cap = wave_stats_cap
        _check(lib.mce_graph_count(self.handle, byref(o), byref(r)))
        out = _result_dict(r, hist)
        if sn is not None:
            out["sub_nodes"] = sn[:n]
        if sc is not None:
            out["sub_cliques"] = sc[:n]
        if ws is not None:
            out["wave_stats"] = [(ws[i].wave, ws[i].cu, ws[i].value, tuple(ws[i].bd)) for i in range(r.wave_stats_len)]
        return out

    def count_noalloc(self, induced=IPX, **block) -> dict:
        """mce_graph_count_noalloc: level 1, no worker list, plain schedule,
        with every device buffer taken from the library's static scratch array."""
        o = default_opts(level=L1, induced=induced, worklist=NOWL, device=self.device, **block)
        r = Result()
        hist = np.zeros(HIST_LEN, np.uint64)
        r.hist = _u64p(hist)
        r.hist_len = HIST_LEN
        _check(lib.mce_graph_count_noalloc(self.handle, byref(o), byref(r)))
        return _result_dict(r, hist)

    def autotune(self, sample_fraction=0.05, level=L1, induced=IPX, worklist=WL, **block):
        """Fastest block settings on a stride sample; returns (settings, trials)."""
        import json as _json
        base = default_opts(level=level, induced=induced, worklist=worklist, device=self.device, **block)
        best = Opts()
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib.mce_autotune(self.handle, byref(base), sample_fraction, byref(best), buf, len(buf)))
        keys = ("worklist", "waves_per_group", "groups_per_cu", "stack_lds_levels", "donate_min_p",
                "donate_check_nodes", "donate_age_us", "small_universe", "lds_arena_bytes",
                "lane_universe")
        return {k: getattr(best, k) for k in keys}, _json.loads(buf.value.decode() or "[]")

    def close(self) -> None:
        if self.handle:
            lib.mce_graph_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _result_dict(r: Result, hist: np.ndarray) -> dict:
    nz = np.nonzero(hist)[0]
    return dict(
        count=int(r.count), hash=int(r.hash), hist={int(k): int(hist[k]) for k in nz},
        max_clique_size=int(r.max_clique_size), degeneracy=int(r.degeneracy), max_degree=int(r.max_degree),
        max_out_degree=int(r.max_out_degree), num_waves=int(r.num_waves), worklist_used=int(r.worklist_used),
        tree_nodes=int(r.tree_nodes),
        donations=int(r.donations), fallback_events=int(r.fallback_events), split_vertices=int(r.split_vertices),
        t_order_ms=float(r.t_order_ms),
        t_orient_ms=float(r.t_orient_ms), t_count_ms=float(r.t_count_ms),
    )


def run_batch(off: np.ndarray, g: Graph, device: int = 0, **kw):
    """All graphs of a packed batch in one launch; returns (per-graph counts, result)."""
    off = np.ascontiguousarray(off, dtype=np.uint32)
    b = Batch()
    b.num_graphs = off.size - 1
    b.vtx_off = _u32p(off)
    b.csr = g.csr()
    o = default_opts(device=device, **kw)
    counts = np.zeros(max(off.size - 1, 1), np.uint64)
    r = Result()
    hist = np.zeros(HIST_LEN, np.uint64)
    r.hist = _u64p(hist)
    r.hist_len = HIST_LEN
    _check(lib.mce_run_batch(byref(b), byref(o), _u64p(counts), byref(r)))
    return counts[:off.size - 1], _result_dict(r, hist)


def run(g: Graph, device: int = 0, **kw) -> dict:
    """Order + count on one GPU (mce_graph_create/order/count)."""
    with DeviceGraph(g, device) as dg:
        info = dg.order()
        res = dg.count(**kw)
        res["t_order_ms"] = info["t_order_ms"]
        res["t_orient_ms"] = info["t_orient_ms"]
        return res


# ---- learned cost model -----------------------------------------------------

def mlp_init(x: np.ndarray, seed: int = 1) -> Mlp:
    x = np.ascontiguousarray(x, dtype=np.float32)
    m = Mlp()
    _check(lib.mce_mlp_init(byref(m), _f32p(x), x.shape[0], seed))
    return m


def mlp_params(m: Mlp) -> dict:
    p = np.ctypeslib.as_array(m.params).copy()
    o = [0, 16 * 64, 16 * 64 + 64, 16 * 64 + 64 + 64 * 64, 16 * 64 + 64 + 64 * 64 + 64, MLP_PARAMS - 1]
    return dict(W1=p[o[0]:o[1]].reshape(16, 64), b1=p[o[1]:o[2]], W2=p[o[2]:o[3]].reshape(64, 64), b2=p[o[3]:o[4]],
                w3=p[o[4]:o[5]], b3=float(p[o[5]]), mu=np.ctypeslib.as_array(m.mu).copy(),
                sigma=np.ctypeslib.as_array(m.sigma).copy())


def mlp_train(m: Mlp, x: np.ndarray, y: np.ndarray, epochs=20, batch=8192, lr=1e-3, weight_decay=0.0, seed=1,
              device=0) -> float:
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    o = TrainOpts(device, epochs, batch, lr, weight_decay, seed)
    loss = c_double()
    _check(lib.mce_mlp_train(byref(m), _f32p(x), _f32p(y), x.shape[0], byref(o), byref(loss)))
    return loss.value


def mlp_grad(m: Mlp, x: np.ndarray, y: np.ndarray, device=0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = c_double()
    _check(lib.mce_mlp_grad(byref(m), _f32p(x), _f32p(y), x.shape[0], device, _f32p(g), byref(loss)))
    return g, loss.value


def mlp_predict(m: Mlp, x: np.ndarray, device=0) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(max(x.shape[0], 1), np.float32)
    _check(lib.mce_mlp_predict(byref(m), _f32p(x), x.shape[0], device, _f32p(out)))
    return out[:x.shape[0]]


def mlp_save(path: str, m: Mlp) -> None:
    _check(lib.mce_mlp_save(path.encode(), byref(m)))


def mlp_load(path: str) -> Mlp:
    m = Mlp()
    _check(lib.mce_mlp_load(path.encode(), byref(m)))
    return m
