"""ctypes binding of the CPU oracle (test infrastructure only).

Loads oracle/liboracle.so (the plain-C restatement of the reference's CPU
DAWN, dawn_oracle.c) and, when present, oracle/_ref/libdawn_ref.so (the
reference's own CPU sources compiled by oracle/Makefile).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module; it is the checker and the CPU baseline, never a product code path.
"""
from __future__ import annotations

import ctypes
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libdawn_ref.so")
INF = 2147483647

_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)
_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)
_lib = None
_ref = None


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        l = ctypes.CDLL(ORACLE_LIB)
        c = ctypes
        for name, res, args in [
            ("oracle_bfs_dense", c.c_int, [c.c_int32, _I32P, _I32P, c.c_int32, _I32P, _I32P]),
            ("oracle_bfs_sparse", c.c_int, [c.c_int32, _I32P, _I32P, c.c_int32, _I32P, _I32P]),
            ("oracle_ref_to_inf", None, [c.c_int32, _I32P, c.c_int32]),
            ("oracle_sssp_dense", c.c_int, [c.c_int32, _I32P, _I32P, _I32P, c.c_int32, _I32P, _I32P]),
            ("oracle_sssp_sparse", c.c_int, [c.c_int32, _I32P, _I32P, _I32P, c.c_int32, _I32P, _I32P]),
            ("oracle_dijkstra", c.c_int, [c.c_int32, _I32P, _I32P, _I32P, c.c_int32, _I32P]),
            ("oracle_mssp", c.c_int, [c.c_int32, _I32P, _I32P, _I32P, _I32P, c.c_int32, _I32P, _I64P,
                                      _I32P, c.c_int32, _F64P]),
            ("oracle_time_bfs_dense", c.c_double, [c.c_int32, _I32P, _I32P, c.c_int32, c.c_int32, _I32P]),
            ("oracle_closeness", c.c_double, [c.c_int32, _I32P, _I64P, _F32P]),
            ("oracle_dobfs_modes", c.c_int, [c.c_int32, _I32P, _I32P, c.c_int32, c.c_int32, c.c_int32,
                                             _I32P, c.c_int32]),
            ("oracle_certify", c.c_int, [c.c_int32, _I32P, _I32P, _I32P, c.c_int32, _I32P, c.c_int32]),
            ("oracle_mlp_forward", None, [_F64P, _F64P, c.c_int32, _F64P]),
            ("oracle_mlp_loss_grad", c.c_double, [_F64P, _F64P, _F64P, c.c_int32, _F64P]),
        ]:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    """The reference's own CPU DAWN (oracle/_ref), or None if not built."""
    global _ref
    if _ref is None and ref_available():
        l = ctypes.CDLL(REF_LIB)
        c = ctypes
        l.ref_bfs.restype = c.c_int
        l.ref_bfs.argtypes = [c.c_int, _I32P, _I32P, c.c_int, c.c_int, _I32P, c.c_char_p]
        l.ref_sssp.restype = c.c_int
        l.ref_sssp.argtypes = [c.c_int, _I32P, _I32P, _F32P, c.c_int, _F32P, c.c_char_p, c.c_int]
        l.ref_closeness.restype = c.c_float
        l.ref_closeness.argtypes = [c.c_int, _I32P, _I32P, _F32P, c.c_int, c.c_int]
        l.ref_mssp_time.restype = c.c_double
        l.ref_mssp_time.argtypes = [c.c_int, c.c_longlong, _I32P, _I32P, _F32P, _I32P, c.c_int,
                                    c.c_int, _F32P]
        _ref = l
    return _ref


def _csr(row_ptr, col, w=None):
    rp = np.ascontiguousarray(row_ptr, np.int32)
    cc = np.ascontiguousarray(col, np.int32)
    ww = None if w is None else np.ascontiguousarray(w, np.int32)
    return rp, cc, ww


def bfs(row_ptr, col, source, dense=True, encoding="inf"):
    """CPU DAWN BFS; encoding 'inf' (engine) or 'ref' (0 = unvisited)."""
    rp, cc, _ = _csr(row_ptr, col)
    n = len(rp) - 1
    dist = np.zeros(n, np.int32)
    steps = ctypes.c_int32(0)
    fn = lib().oracle_bfs_dense if dense else lib().oracle_bfs_sparse
    rc = fn(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32), int(source), _p(dist, ctypes.c_int32),
            ctypes.byref(steps))
    if rc != 0:
        raise ValueError(f"oracle bfs status {rc}")
    if encoding == "inf":
        lib().oracle_ref_to_inf(n, _p(dist, ctypes.c_int32), int(source))
    return dist


def sssp(row_ptr, col, w, source, dense=True, overflow_ok=False):
    """CPU DAWN min-plus SSSP.  A path whose weight reaches INT32_MAX is
    skipped (status -8); overflow_ok accepts that and returns the row."""
    rp, cc, ww = _csr(row_ptr, col, w)
    n = len(rp) - 1
    dist = np.zeros(n, np.int32)
    steps = ctypes.c_int32(0)
    fn = lib().oracle_sssp_dense if dense else lib().oracle_sssp_sparse
    rc = fn(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32), _p(ww, ctypes.c_int32), int(source),
            _p(dist, ctypes.c_int32), ctypes.byref(steps))
    if rc != 0 and not (overflow_ok and rc == -8):
        raise ValueError(f"oracle sssp status {rc}")
    return dist


def dijkstra(row_ptr, col, w, source, overflow_ok=False):
    rp, cc, ww = _csr(row_ptr, col, w)
    n = len(rp) - 1
    dist = np.zeros(n, np.int32)
    rc = lib().oracle_dijkstra(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32),
                               None if ww is None else _p(ww, ctypes.c_int32), int(source),
                               _p(dist, ctypes.c_int32))
    if rc != 0 and not (overflow_ok and rc == -8):
        raise ValueError(f"oracle dijkstra status {rc}")
    return dist


def mssp(row_ptr, col, w, sources, want_dist=True, threads=0, overflow_ok=False):
    rp, cc, ww = _csr(row_ptr, col, w)
    n = len(rp) - 1
    src = np.ascontiguousarray(sources, np.int32)
    S = len(src)
    dist = np.zeros((S, n), np.int32) if want_dist else None
    sums = np.zeros(S, np.int64)
    skipped = np.zeros(S, np.int32)
    secs = ctypes.c_double(0)
    rc = lib().oracle_mssp(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32),
                           None if ww is None else _p(ww, ctypes.c_int32), _p(src, ctypes.c_int32), S,
                           None if dist is None else _p(dist, ctypes.c_int32), _p(sums, ctypes.c_int64),
                           _p(skipped, ctypes.c_int32), threads, ctypes.byref(secs))
    if rc != 0 and not (overflow_ok and rc == -8):
        raise ValueError(f"oracle mssp status {rc}")
    return dist, sums, skipped, secs.value


def certify(row_ptr, col, w, source, dist, threads=0):
    """True iff dist is the shortest-path distance array from source
    (optimality conditions, O(m)); w None = unit weights (BFS hops)."""
    rp, cc, ww = _csr(row_ptr, col, w)
    d = np.ascontiguousarray(dist, np.int32)
    n = len(rp) - 1
    assert len(d) == n
    rc = lib().oracle_certify(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32),
                              None if ww is None else _p(ww, ctypes.c_int32), int(source),
                              _p(d, ctypes.c_int32), int(threads))
    if rc < 0:
        raise ValueError(f"oracle certify status {rc}")
    return rc == 1


def closeness(dist_row):
    d = np.ascontiguousarray(dist_row, np.int32)
    s = ctypes.c_int64(0)
    r = ctypes.c_float(0)
    c = lib().oracle_closeness(len(d), _p(d, ctypes.c_int32), ctypes.byref(s), ctypes.byref(r))
    return c, s.value, r.value


def dobfs_modes(row_ptr, col, source, alpha=15, beta=18, max_steps=4096):
    rp, cc, _ = _csr(row_ptr, col)
    n = len(rp) - 1
    modes = np.zeros(max_steps, np.int32)
    k = lib().oracle_dobfs_modes(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32), int(source), alpha,
                                 beta, _p(modes, ctypes.c_int32), max_steps)
    return modes[:k].copy()


def time_bfs_dense(row_ptr, col, source, threads=0):
    rp, cc, _ = _csr(row_ptr, col)
    n = len(rp) - 1
    dist = np.zeros(n, np.int32)
    return lib().oracle_time_bfs_dense(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32), int(source),
                                       threads, _p(dist, ctypes.c_int32))


def mlp_forward(params, x):
    p = np.ascontiguousarray(params, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    out = np.zeros((len(xx), 3), np.float64)
    lib().oracle_mlp_forward(_p(p, ctypes.c_double), _p(xx, ctypes.c_double), len(xx),
                             _p(out, ctypes.c_double))
    return out


def mlp_loss_grad(params, x, y):
    p = np.ascontiguousarray(params, np.float64)
    xx = np.ascontiguousarray(x, np.float64)
    yy = np.ascontiguousarray(y, np.float64)
    g = np.zeros(len(p), np.float64)
    loss = lib().oracle_mlp_loss_grad(_p(p, ctypes.c_double), _p(xx, ctypes.c_double),
                                      _p(yy, ctypes.c_double), len(xx), _p(g, ctypes.c_double))
    return loss, g


# ---- the reference's own CPU code (oracle/_ref) ---------------------------------

def ref_bfs(row_ptr, col, source, sparse=False):
    """Distances from DAWN::BFS_CPU::BFS / BFS_kernel (reference encoding)."""
    r = ref()
    rp, cc, _ = _csr(row_ptr, col)
    n = len(rp) - 1
    dist = np.zeros(n, np.int32)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "out.txt").encode()
        rc = r.ref_bfs(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32), int(source), int(sparse),
                       _p(dist, ctypes.c_int32), path)
    if rc != 0:
        raise ValueError(f"ref_bfs status {rc}")
    return dist


def ref_sssp(row_ptr, col, w, source, threads=1):
    """Distances from DAWN::SSSP_CPU::SSSP with float weights (FLT_MAX
    unreached).  One thread by default: with more, the reference's
    unsynchronised min can lose updates."""
    r = ref()
    rp, cc, ww = _csr(row_ptr, col, w)
    n = len(rp) - 1
    wf = ww.astype(np.float32)
    dist = np.zeros(n, np.float32)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "out.txt").encode()
        rc = r.ref_sssp(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32), _p(wf, ctypes.c_float),
                        int(source), _p(dist, ctypes.c_float), path, int(threads))
    if rc != 0:
        raise ValueError(f"ref_sssp status {rc}")
    return dist


def ref_closeness(row_ptr, col, w, source, threads=1):
    """DAWN::CC_CPU::kernel / kernel_Weighted of the reference build (one
    OpenMP thread by default: the weighted min is unsynchronised)."""
    r = ref()
    rp, cc, ww = _csr(row_ptr, col, w)
    n = len(rp) - 1
    wf = None if ww is None else ww.astype(np.float32)
    return r.ref_closeness(n, _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32),
                           None if wf is None else _p(wf, ctypes.c_float), int(source), int(threads))


def weighted_closeness_ref(dist_row):
    """The reference's weighted closeness value from an engine-encoded int32
    row (src/algorithm/cpu/cc.cpp:132-133): float distances with FLT_MAX
    for unreached vertices, accumulated in order from 0.0f in float, then
    (row - 1) / sum in double, returned as float."""
    d = np.asarray(dist_row)
    f = np.where(d == INF, np.finfo(np.float32).max, d).astype(np.float32)
    with np.errstate(over="ignore"):
        acc = np.add.accumulate(f, dtype=np.float32)[-1] if len(f) else np.float32(0)
        return np.float32((1.0 * len(d) - 1.0) / np.float64(acc))


def ref_mssp_time(row_ptr, col, sources, threads, w=None):
    """Wall seconds of DAWN::MSSP_CPU::run over `sources`; also returns the
    reference's own reported value."""
    r = ref()
    rp, cc, ww = _csr(row_ptr, col, w)
    n = len(rp) - 1
    wf = None if ww is None else ww.astype(np.float32)
    src = np.ascontiguousarray(sources, np.int32)
    rep = ctypes.c_float(0)
    t = r.ref_mssp_time(n, len(cc), _p(rp, ctypes.c_int32), _p(cc, ctypes.c_int32),
                        None if wf is None else _p(wf, ctypes.c_float), _p(src, ctypes.c_int32),
                        len(src), int(threads), ctypes.byref(rep))
    return t, rep.value
