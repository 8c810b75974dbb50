"""Host binding of the MI355X-native DAWN frontier engine.

A thin ctypes layer over ``libdawn_amd.so`` (the C ABI declared in
``include/dawn_capi.h``).  The compute lives in hand-written gfx950 HIP
kernels; this module only moves numpy arrays across the boundary.  There is
no CPU fallback: if the shared library is missing, importing the compute
entry points raises.

Reference interface mirrored (lxrzlyr/GAL-DAWN):
  DAWN::Graph::createGraph / Graph_t      -> HostGraph.read_mtx, Graph
  DAWN::BFS_GPU::run   (src/algorithm/gpu/bfs.cu:107-129)   -> Graph.bfs
  DAWN::SSSP_GPU::run  (src/algorithm/gpu/sssp.cu:131-157)  -> Graph.sssp
  DAWN::MSSP_GPU::run  (src/algorithm/gpu/mssp.cu:9-102)    -> Graph.mssp
  DAWN::APSP_GPU::run  (src/algorithm/gpu/apsp.cu:9-111)    -> Graph.apsp_block
  DAWN::CC_GPU::run    (src/algorithm/gpu/cc.cu:9-202)      -> Graph.closeness
  DAWN::IO::readList / outfile (src/io.cpp:194-281)         -> read_list, write_dist
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdawn_amd.so")

INF = 2147483647
OK = 0
ERR_ISOLATED_SOURCE = -2
ERR_NO_DEVICE = -3
ERR_SCRATCH_LIMIT = -11

MODE_PUSH, MODE_PULL, MODE_BITMAP = 0, 1, 2
MODE_NAMES = ("push", "pull", "bitmap")
POLICY_FIXED, POLICY_LEARNED, POLICY_PUSH, POLICY_PULL, POLICY_BITMAP = 0, 1, 2, 3, 4
POLICIES = {"fixed": 0, "learned": 1, "push": 2, "pull": 3, "bitmap": 4}
SCHED_FEATURES = 8
SCHED_PARAMS = 1443


class DawnError(RuntimeError):
    def __init__(self, status: int, what: str, detail: str = ""):
        self.status = status
        msg = f"{what}: {status_string(status)} ({status})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


class _Csr(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("m", ctypes.c_int64),
        ("directed", ctypes.c_int32),
        ("weighted", ctypes.c_int32),
        ("row_ptr", ctypes.POINTER(ctypes.c_int32)),
        ("col", ctypes.POINTER(ctypes.c_int32)),
        ("weight", ctypes.POINTER(ctypes.c_int32)),
    ]


class Frontier(ctypes.Structure):
    """dawn_frontier: device pointers of one step (include/dawn_capi.h)."""
    _fields_ = [
        ("alpha", ctypes.c_void_p),
        ("beta", ctypes.c_void_p),
        ("dist", ctypes.c_void_p),
        ("changed", ctypes.c_void_p),
        ("level", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("device_ms", ctypes.c_double),
        ("steps", ctypes.c_int64),
        ("mode_steps", ctypes.c_int64 * 3),
        ("edges_reached", ctypes.c_int64),
        ("vertices_reached", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("skipped_sources", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("frontier_edges", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {
            "device_ms": self.device_ms,
            "steps": self.steps,
            "mode_steps": {MODE_NAMES[i]: self.mode_steps[i] for i in range(3)},
            "edges_reached": self.edges_reached,
            "vertices_reached": self.vertices_reached,
            "batches": self.batches,
            "skipped_sources": self.skipped_sources,
            "frontier_edges": self.frontier_edges,
        }


_lib = None
_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)
_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)


def _sig(lib):
    c = ctypes
    P = c.c_void_p
    defs = {
        "dawn_device_count": (c.c_int, [_I32P]),
        "dawn_status_string": (c.c_char_p, [c.c_int]),
        "dawn_last_error": (c.c_char_p, []),
        "dawn_gen_er": (c.c_int, [c.c_int32, c.c_double, c.c_uint64, c.c_int32, c.c_int32, c.c_int32, c.POINTER(_Csr)]),
        "dawn_gen_rmat": (c.c_int, [c.c_int32, c.c_int32, c.c_double, c.c_double, c.c_double, c.c_uint64,
                                    c.c_int32, c.c_int32, c.c_int32, c.POINTER(_Csr)]),
        "dawn_gen_grid": (c.c_int, [c.c_int32, c.c_double, c.c_uint64, c.c_int32, c.c_int32, c.c_int32, c.POINTER(_Csr)]),
        "dawn_gen_sources": (c.c_int, [c.c_int32, c.c_int32, c.c_uint64, _I32P]),
        "dawn_csr_from_edges": (c.c_int, [c.c_int32, c.c_int64, _I32P, _I32P, _I32P, c.c_int32, c.c_int32, c.POINTER(_Csr)]),
        "dawn_read_mtx": (c.c_int, [c.c_char_p, c.c_int32, c.POINTER(_Csr)]),
        "dawn_read_mtx_as": (c.c_int, [c.c_char_p, c.c_int32, c.c_int32, c.POINTER(_Csr)]),
        "dawn_write_mtx": (c.c_int, [c.c_char_p, c.POINTER(_Csr)]),
        "dawn_csr_free": (None, [c.POINTER(_Csr)]),
        "dawn_read_list": (c.c_int, [c.c_char_p, c.POINTER(_I32P), _I64P]),
        "dawn_free": (None, [P]),
        "dawn_write_dist": (c.c_int, [c.c_char_p, c.c_int32, _I32P, c.c_int32]),
        "dawn_graph_create": (c.c_int, [c.POINTER(_Csr), c.POINTER(P)]),
        "dawn_graph_destroy": (c.c_int, [P]),
        "dawn_graph_info": (c.c_int, [P, _I32P, _I64P, _I32P, _I32P]),
        "dawn_graph_from_edges": (c.c_int, [c.c_int32, c.c_int64, P, P, P, c.c_int32, c.c_int32,
                                            c.POINTER(P)]),
        "dawn_graph_gen_er": (c.c_int, [c.c_int32, c.c_double, c.c_uint64, c.c_int32, c.c_int32,
                                        c.c_int32, c.POINTER(P)]),
        "dawn_graph_gen_rmat": (c.c_int, [c.c_int32, c.c_int32, c.c_double, c.c_double, c.c_double,
                                          c.c_uint64, c.c_int32, c.c_int32, c.c_int32, c.POINTER(P)]),
        "dawn_graph_gen_grid": (c.c_int, [c.c_int32, c.c_double, c.c_uint64, c.c_int32, c.c_int32,
                                          c.c_int32, c.POINTER(P)]),
        "dawn_graph_download": (c.c_int, [P, c.POINTER(_Csr)]),
        "dawn_bfs": (c.c_int, [P, c.c_int32, P, c.c_int32, c.POINTER(Stats)]),
        "dawn_sssp": (c.c_int, [P, c.c_int32, P, c.c_int32, c.POINTER(Stats)]),
        "dawn_bfs_noalloc": (c.c_int, [P, c.c_int32, P, c.c_int32, c.POINTER(Stats)]),
        "dawn_mssp": (c.c_int, [P, _I32P, c.c_int32, P, _I64P, c.c_int32, c.POINTER(Stats)]),
        "dawn_apsp_block": (c.c_int, [P, c.c_int32, c.c_int32, P, _I64P, c.c_int32, c.POINTER(Stats)]),
        "dawn_closeness": (c.c_int, [P, _I32P, c.c_int32, _F64P, _I64P, _F32P, c.c_int32, c.POINTER(Stats)]),
        "dawn_batch_mssp": (c.c_int, [c.POINTER(P), c.c_int32, _I32P, c.c_int32, P, _I64P, c.POINTER(Stats)]),
        "dawn_probe_copy": (c.c_int, [c.c_int64, c.c_int32, _F64P]),
        "dawn_mssp_profile": (c.c_int, [P, _I32P, c.c_int32, c.c_int32, _F64P, _I64P, _I32P]),
        "dawn_step": (c.c_int, [P, c.POINTER(Frontier), c.c_int32, c.c_void_p]),
        "dawn_device_malloc": (c.c_int, [c.c_int64, c.POINTER(c.c_void_p)]),
        "dawn_device_free": (c.c_int, [c.c_void_p]),
        "dawn_memcpy": (c.c_int, [c.c_void_p, c.c_void_p, c.c_int64]),
        "dawn_memset": (c.c_int, [c.c_void_p, c.c_int32, c.c_int64]),
        "dawn_graph_synchronize": (c.c_int, [P]),
        "dawn_sched_set_mlp": (c.c_int, [_F32P, _F32P, _F32P]),
        "dawn_sched_record": (c.c_int, [P, _I32P, c.c_int32, c.c_int32, _F32P, _F32P, _I32P, c.c_int32,
                                        _I32P]),
        "dawn_sched_trace": (c.c_int, [P, _I32P, c.c_int32, c.c_int32, _I32P, _F32P, _F32P, c.c_int32, _I32P]),
        "dawn_sched_set_mlp_single": (c.c_int, [_F32P, _F32P, _F32P]),
        "dawn_sched_record_single": (c.c_int, [P, _I32P, c.c_int32, c.c_int32, _F32P, _F32P, _I32P,
                                               c.c_int32, _I32P]),
        "dawn_sched_trace_single": (c.c_int, [P, _I32P, c.c_int32, c.c_int32, _I32P, _F32P, _F32P,
                                              c.c_int32, _I32P]),
        "dawn_mlp_forward": (c.c_int, [_F32P, _F32P, c.c_int32, _F32P]),
        "dawn_mlp_loss_grad": (c.c_int, [_F32P, _F32P, _F32P, c.c_int32, _F32P, _F32P]),
        "dawn_mlp_train": (c.c_int, [_F32P, _F32P, _F32P, c.c_int32, c.c_int32, c.c_int32, c.c_float, c.c_uint64, _F32P]),
        "dawn_tune": (c.c_int, [P, c.c_int32]),
        "dawn_tune_load": (c.c_int, [c.c_char_p]),
        "dawn_tune_save": (c.c_int, [c.c_char_p]),
        "dawn_graph_get_knobs": (c.c_int, [P, c.POINTER(c.c_int32)]),
        "dawn_graph_set_knobs": (c.c_int, [P, c.POINTER(c.c_int32)]),
    }
    for name, (res, args) in defs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return defs


def lib():
    """The loaded C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python __graft_entry__.py build` "
                "(or `make -C gal-dawn_amd`); there is no CPU fallback")
        l = ctypes.CDLL(LIB_PATH)
        _sig(l)
        _lib = l
    return _lib


def exported_symbols() -> list:
    """Names of every C-ABI function this binding declares."""
    return sorted(_sig(lib()).keys())


def status_string(status: int) -> str:
    try:
        return lib().dawn_status_string(status).decode()
    except Exception:  # library missing while formatting an error
        return str(status)


def _check(rc: int, what: str):
    if rc != OK:
        detail = ""
        try:
            detail = lib().dawn_last_error().decode()
        except Exception:
            pass
        raise DawnError(rc, what, detail)


def device_count() -> int:
    n = ctypes.c_int32(0)
    _check(lib().dawn_device_count(ctypes.byref(n)), "dawn_device_count")
    return int(n.value)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class HostGraph:
    """Host CSR (int32 offsets/columns, optional int32 weights).

    Arrays are numpy views valid while this object lives.  Built by the
    seeded generators, the MatrixMarket reader, an edge list, or wrapped
    around caller arrays (borrowed)."""

    def __init__(self):
        self._csr = _Csr()
        self._owned = False
        self._keep = []

    # -- constructors ------------------------------------------------------
    @classmethod
    def _from_c(cls, fn, *args, what="generator") -> "HostGraph":
        g = cls()
        _check(fn(*args, ctypes.byref(g._csr)), what)
        g._owned = True
        return g

    @classmethod
    def er(cls, n: int, avg_degree: float, seed: int, weighted: bool = False,
           wmin: int = 1, wmax: int = 1000) -> "HostGraph":
        return cls._from_c(lib().dawn_gen_er, n, avg_degree, seed, int(weighted), wmin, wmax,
                           what="dawn_gen_er")

    @classmethod
    def rmat(cls, scale: int, edge_factor: int, seed: int, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, weighted: bool = False, wmin: int = 1, wmax: int = 1000) -> "HostGraph":
        return cls._from_c(lib().dawn_gen_rmat, scale, edge_factor, a, b, c, seed, int(weighted),
                           wmin, wmax, what="dawn_gen_rmat")

    @classmethod
    def grid(cls, side: int, extra_fraction: float, seed: int, weighted: bool = False,
             wmin: int = 1, wmax: int = 1000) -> "HostGraph":
        return cls._from_c(lib().dawn_gen_grid, side, extra_fraction, seed, int(weighted), wmin,
                           wmax, what="dawn_gen_grid")

    @classmethod
    def from_edges(cls, n: int, u, v, w=None, symmetrize: bool = True,
                   dedup: bool = True) -> "HostGraph":
        u = _i32(u)
        v = _i32(v)
        wp = None
        if w is not None:
            w = _i32(w)
            wp = _ptr(w, ctypes.c_int32)
        return cls._from_c(lib().dawn_csr_from_edges, n, len(u), _ptr(u, ctypes.c_int32),
                           _ptr(v, ctypes.c_int32), wp, int(symmetrize), int(dedup),
                           what="dawn_csr_from_edges")

    @classmethod
    def read_mtx(cls, path: str, weighted: bool = False, directed: Optional[bool] = None) -> "HostGraph":
        """MatrixMarket file; directed None follows the header's symmetry
        field, True / False read it as the reference's directed / undirected
        readers do."""
        if directed is None:
            return cls._from_c(lib().dawn_read_mtx, path.encode(), int(weighted), what="dawn_read_mtx")
        return cls._from_c(lib().dawn_read_mtx_as, path.encode(), int(weighted), int(directed),
                           what="dawn_read_mtx_as")

    @classmethod
    def from_csr(cls, row_ptr, col, weight=None, directed: bool = False) -> "HostGraph":
        g = cls()
        rp, cc = _i32(row_ptr), _i32(col)
        g._keep = [rp, cc]
        g._csr.n = len(rp) - 1
        g._csr.m = len(cc)
        g._csr.directed = int(directed)
        g._csr.row_ptr = _ptr(rp, ctypes.c_int32)
        g._csr.col = _ptr(cc, ctypes.c_int32)
        if weight is not None:
            ww = _i32(weight)
            g._keep.append(ww)
            g._csr.weight = _ptr(ww, ctypes.c_int32)
            g._csr.weighted = 1
        return g

    def write_mtx(self, path: str):
        _check(lib().dawn_write_mtx(path.encode(), ctypes.byref(self._csr)), "dawn_write_mtx")

    # -- views ---------------------------------------------------------------
    @property
    def n(self) -> int:
        return int(self._csr.n)

    @property
    def m(self) -> int:
        return int(self._csr.m)

    @property
    def directed(self) -> bool:
        return bool(self._csr.directed)

    @property
    def weighted(self) -> bool:
        return bool(self._csr.weight)

    @property
    def row_ptr(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._csr.row_ptr, shape=(self.n + 1,))

    @property
    def col(self) -> np.ndarray:
        if self.m == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(self._csr.col, shape=(self.m,))

    @property
    def weight(self) -> Optional[np.ndarray]:
        if not self._csr.weight:
            return None
        if self.m == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(self._csr.weight, shape=(self.m,))

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr)

    def __del__(self):
        if self._owned and _lib is not None:
            _lib.dawn_csr_free(ctypes.byref(self._csr))
            self._owned = False


def probe_copy(nbytes: int, iters: int = 5) -> float:
    """Achieved GB/s (read + write) of a plain device copy of nbytes."""
    g = ctypes.c_double(0)
    _check(lib().dawn_probe_copy(nbytes, iters, ctypes.byref(g)), "dawn_probe_copy")
    return g.value


def gen_sources(n: int, count: int, seed: int) -> np.ndarray:
    out = np.zeros(count, np.int32)
    _check(lib().dawn_gen_sources(n, count, seed, _ptr(out, ctypes.c_int32)), "dawn_gen_sources")
    return out


def read_list(path: str) -> np.ndarray:
    p = _I32P()
    cnt = ctypes.c_int64(0)
    _check(lib().dawn_read_list(path.encode(), ctypes.byref(p), ctypes.byref(cnt)), "dawn_read_list")
    try:
        return np.ctypeslib.as_array(p, shape=(max(cnt.value, 1),))[: cnt.value].copy()
    finally:
        lib().dawn_free(p)


def write_dist(path: str, dist: np.ndarray, source: int):
    d = _i32(dist)
    _check(lib().dawn_write_dist(path.encode(), len(d), _ptr(d, ctypes.c_int32), source),
           "dawn_write_dist")


def _policy(p) -> int:
    return POLICIES[p] if isinstance(p, str) else int(p)


def _out_arg(out):
    """Host numpy array or torch tensor (device or host) -> void pointer."""
    if out is None:
        return None
    if isinstance(out, np.ndarray):
        assert out.dtype == np.int32 and out.flags["C_CONTIGUOUS"]
        return ctypes.c_void_p(out.ctypes.data)
    return ctypes.c_void_p(int(out.data_ptr()))  # torch tensor


class DeviceArray:
    """A device buffer from dawn_device_malloc, viewed as a numpy dtype and
    shape; data_ptr() is what the entry points take for device outputs."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(np.asarray(shape, np.int64)).tolist())
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().dawn_device_malloc(self.nbytes, ctypes.byref(p)), "dawn_device_malloc")
        self._p = p.value

    def data_ptr(self) -> int:
        return self._p

    def upload(self, arr):
        a = np.ascontiguousarray(arr, self.dtype)
        assert a.nbytes == self.nbytes
        _check(lib().dawn_memcpy(ctypes.c_void_p(self._p), ctypes.c_void_p(a.ctypes.data), self.nbytes),
               "dawn_memcpy")
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(lib().dawn_memcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self._p), self.nbytes),
               "dawn_memcpy")
        return out

    def fill_bytes(self, value: int = 0):
        _check(lib().dawn_memset(ctypes.c_void_p(self._p), int(value), self.nbytes), "dawn_memset")
        return self

    def close(self):
        if self._p and _lib is not None:
            _lib.dawn_device_free(ctypes.c_void_p(self._p))
        self._p = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Device-resident graph (dawn_graph): CSR (+ CSC when directed)."""

    def __init__(self, host: Optional[HostGraph] = None):
        self._h = ctypes.c_void_p()
        if host is None:  # filled by a device-side builder
            return
        _check(lib().dawn_graph_create(ctypes.byref(host._csr), ctypes.byref(self._h)),
               "dawn_graph_create")
        self.n = host.n
        self.m = host.m
        self.directed = host.directed
        self.weighted = host.weighted

    # -- built on the device (edge list / generators -> CSR on the GPU) ----
    @classmethod
    def _built(cls, fn, *args, what) -> "Graph":
        g = cls()
        _check(fn(*args, ctypes.byref(g._h)), what)
        n, m, d, w = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        _check(lib().dawn_graph_info(g._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(d),
                                     ctypes.byref(w)), "dawn_graph_info")
        g.n, g.m, g.directed, g.weighted = n.value, m.value, bool(d.value), bool(w.value)
        return g

    @classmethod
    def from_edges_device(cls, n: int, u, v, w=None, symmetrize: bool = True,
                          dedup: bool = True) -> "Graph":
        """Edge list (numpy host arrays or device buffers) -> CSR built on the
        device, identical to HostGraph.from_edges + Graph."""
        keep = []  # host arrays stay alive through the call

        def arg(a):
            if a is None:
                return None
            if isinstance(a, DeviceArray):
                return ctypes.c_void_p(a.data_ptr())
            x = _i32(a)
            keep.append(x)
            return ctypes.c_void_p(x.ctypes.data)
        pu, pv, pw = arg(u), arg(v), arg(w)
        count = u.shape[0] if isinstance(u, DeviceArray) else len(u)
        return cls._built(lib().dawn_graph_from_edges, n, count, pu, pv, pw, int(symmetrize),
                          int(dedup), what="dawn_graph_from_edges")

    @classmethod
    def er_device(cls, n: int, avg_degree: float, seed: int, weighted: bool = False,
                  wmin: int = 1, wmax: int = 1000) -> "Graph":
        return cls._built(lib().dawn_graph_gen_er, n, avg_degree, seed, int(weighted), wmin, wmax,
                          what="dawn_graph_gen_er")

    @classmethod
    def rmat_device(cls, scale: int, edge_factor: int, seed: int, a: float = 0.57,
                    b: float = 0.19, c: float = 0.19, weighted: bool = False, wmin: int = 1,
                    wmax: int = 1000) -> "Graph":
        return cls._built(lib().dawn_graph_gen_rmat, scale, edge_factor, a, b, c, seed,
                          int(weighted), wmin, wmax, what="dawn_graph_gen_rmat")

    @classmethod
    def grid_device(cls, side: int, extra_fraction: float, seed: int, weighted: bool = False,
                    wmin: int = 1, wmax: int = 1000) -> "Graph":
        return cls._built(lib().dawn_graph_gen_grid, side, extra_fraction, seed, int(weighted),
                          wmin, wmax, what="dawn_graph_gen_grid")

    def download(self) -> HostGraph:
        """The device CSR as a host CSR (owned)."""
        h = HostGraph()
        _check(lib().dawn_graph_download(self._h, ctypes.byref(h._csr)), "dawn_graph_download")
        h._owned = True
        return h

    def close(self):
        if self._h and self._h.value and _lib is not None:
            _lib.dawn_graph_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def bfs(self, source: int, policy="fixed", out=None, stats: bool = False):
        dist = np.empty(self.n, np.int32) if out is None else out
        st = Stats()
        _check(lib().dawn_bfs(self._h, int(source), _out_arg(dist), _policy(policy),
                              ctypes.byref(st)), "dawn_bfs")
        return (dist, st.as_dict()) if stats else dist

    def bfs_noalloc(self, source: int, policy="fixed", out=None, stats: bool = False):
        """dawn_bfs without device allocation (fixed scratch region; graphs
        of at most 4,194,304 vertices)."""
        dist = np.empty(self.n, np.int32) if out is None else out
        st = Stats()
        _check(lib().dawn_bfs_noalloc(self._h, int(source), _out_arg(dist), _policy(policy),
                                      ctypes.byref(st)), "dawn_bfs_noalloc")
        return (dist, st.as_dict()) if stats else dist

    def sssp(self, source: int, policy="fixed", out=None, stats: bool = False):
        dist = np.empty(self.n, np.int32) if out is None else out
        st = Stats()
        _check(lib().dawn_sssp(self._h, int(source), _out_arg(dist), _policy(policy),
                               ctypes.byref(st)), "dawn_sssp")
        return (dist, st.as_dict()) if stats else dist

    def mssp(self, sources: Sequence[int], policy="fixed", want_dist: bool = True, out=None,
             stats: bool = False):
        src = _i32(sources)
        S = len(src)
        dist = out
        if dist is None and want_dist:
            dist = np.empty((S, self.n), np.int32)
        sums = np.zeros(S, np.int64)
        st = Stats()
        _check(lib().dawn_mssp(self._h, _ptr(src, ctypes.c_int32), S, _out_arg(dist),
                               _ptr(sums, ctypes.c_int64), _policy(policy), ctypes.byref(st)),
               "dawn_mssp")
        res = (dist, sums)
        return (res + (st.as_dict(),)) if stats else res

    def apsp_block(self, first: int, count: int, policy="fixed", want_dist: bool = True):
        dist = np.empty((count, self.n), np.int32) if want_dist else None
        sums = np.zeros(count, np.int64)
        st = Stats()
        _check(lib().dawn_apsp_block(self._h, first, count, _out_arg(dist),
                                     _ptr(sums, ctypes.c_int64), _policy(policy), ctypes.byref(st)),
               "dawn_apsp_block")
        return dist, sums

    def closeness(self, sources: Sequence[int], policy="fixed", stats: bool = False):
        """Closeness (n-1)/sum per source, the exact int64 sums, and the
        reference's integer-division value; unweighted graphs compute only
        the sums (no distance rows)."""
        src = _i32(sources)
        S = len(src)
        cc = np.zeros(S, np.float64)
        sums = np.zeros(S, np.int64)
        ref = np.zeros(S, np.float32)
        st = Stats()
        _check(lib().dawn_closeness(self._h, _ptr(src, ctypes.c_int32), S, _ptr(cc, ctypes.c_double),
                                    _ptr(sums, ctypes.c_int64), _ptr(ref, ctypes.c_float),
                                    _policy(policy), ctypes.byref(st)), "dawn_closeness")
        return (cc, sums, ref, st.as_dict()) if stats else (cc, sums, ref)

    PROF_NAMES = ("fill", "seed", "decide", "expand_push", "expand_pull", "expand_bitmap",
                  "finalize", "final")

    def mssp_profile(self, sources, policy="fixed"):
        """Per-kernel HIP-event times of an unweighted MSSP (profiling pass)."""
        src = _i32(sources)
        ms = np.zeros(8, np.float64)
        cnt = np.zeros(8, np.int64)
        levels = np.zeros((len(src) + 63) // 64, np.int32)
        _check(lib().dawn_mssp_profile(self._h, _ptr(src, ctypes.c_int32), len(src), _policy(policy),
                                       _ptr(ms, ctypes.c_double), _ptr(cnt, ctypes.c_int64),
                                       _ptr(levels, ctypes.c_int32)), "dawn_mssp_profile")
        # One entry per batch of 64*W sources, -1 past the last batch.
        return ({k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.PROF_NAMES)},
                levels[levels >= 0])

    def synchronize(self):
        """Waits for the work queued on the graph's stream."""
        _check(lib().dawn_graph_synchronize(self._h), "dawn_graph_synchronize")

    def step(self, alpha, beta, dist, changed, level: int = 0, mode="push", stream=None):
        """One frontier step (dawn_step) on device buffers (DeviceArray):
        alpha/beta uint32 bitmaps of ceil(n/32) words, dist int32[n],
        changed int64[1].  stream: a hipStream_t handle, or None for the
        graph's own stream (then synchronize() before reading results)."""
        f = Frontier(alpha.data_ptr(), beta.data_ptr(), dist.data_ptr(), changed.data_ptr(),
                     int(level), 0)
        md = {"push": MODE_PUSH, "pull": MODE_PULL}[mode] if isinstance(mode, str) else int(mode)
        _check(lib().dawn_step(self._h, ctypes.byref(f), md, ctypes.c_void_p(stream or 0)),
               "dawn_step")

    KNOB_NAMES = ("dense_blocks", "heavy_degree", "pull_chunk", "ss_chunk", "pull_variant",
                  "lane_words", "hub_order")

    def knobs(self) -> dict:
        """The graph's launch knobs (dawn_graph_get_knobs)."""
        k = np.zeros(len(self.KNOB_NAMES), np.int32)
        _check(lib().dawn_graph_get_knobs(self._h, _ptr(k, ctypes.c_int32)), "dawn_graph_get_knobs")
        return dict(zip(self.KNOB_NAMES, (int(x) for x in k)))

    def set_knobs(self, **kw):
        """Changes some launch knobs (time only, never results)."""
        cur = self.knobs()
        for name, val in kw.items():
            if name not in cur:
                raise KeyError(name)
            cur[name] = int(val)
        k = np.array([cur[n] for n in self.KNOB_NAMES], np.int32)
        _check(lib().dawn_graph_set_knobs(self._h, _ptr(k, ctypes.c_int32)), "dawn_graph_set_knobs")

    def sched_record(self, sources, repeats: int = 3, max_rows: int = 4096, single: bool = False):
        """Per-step features, time of every mode and the fixed rule's choice
        (batched engine, or with single=True the single-source engine)."""
        src = _i32(sources)
        feat = np.zeros((max_rows, SCHED_FEATURES), np.float32)
        times = np.zeros((max_rows, 3), np.float32)
        fixed = np.zeros(max_rows, np.int32)
        rows = ctypes.c_int32(0)
        fn = lib().dawn_sched_record_single if single else lib().dawn_sched_record
        _check(fn(self._h, _ptr(src, ctypes.c_int32), len(src), repeats,
                  _ptr(feat, ctypes.c_float), _ptr(times, ctypes.c_float),
                  _ptr(fixed, ctypes.c_int32), max_rows, ctypes.byref(rows)),
               "dawn_sched_record_single" if single else "dawn_sched_record")
        r = rows.value
        return feat[:r].copy(), times[:r].copy(), fixed[:r].copy()

    def sched_trace(self, sources, policy="fixed", max_rows: int = 4096, single: bool = False):
        src = _i32(sources)
        modes = np.zeros(max_rows, np.int32)
        feat = np.zeros((max_rows, SCHED_FEATURES), np.float32)
        times = np.zeros((max_rows, 3), np.float32)
        rows = ctypes.c_int32(0)
        fn = lib().dawn_sched_trace_single if single else lib().dawn_sched_trace
        _check(fn(self._h, _ptr(src, ctypes.c_int32), len(src), _policy(policy),
                  _ptr(modes, ctypes.c_int32), _ptr(feat, ctypes.c_float),
                  _ptr(times, ctypes.c_float), max_rows, ctypes.byref(rows)),
               "dawn_sched_trace_single" if single else "dawn_sched_trace")
        r = rows.value
        return modes[:r].copy(), feat[:r].copy(), times[:r].copy()


def batch_mssp(graphs: Sequence[Graph], sources, count_per_graph: int, want_dist: bool = True):
    """Many graphs of uneven size in one persistent launch."""
    G = len(graphs)
    handles = (ctypes.c_void_p * G)(*[g.handle.value for g in graphs])
    src = _i32(sources)
    sizes = np.array([g.n for g in graphs], np.int64) * count_per_graph
    offs = np.zeros(G, np.int64)
    offs[1:] = np.cumsum(sizes)[:-1]
    dist = np.empty(int(sizes.sum()), np.int32) if want_dist else None
    st = Stats()
    _check(lib().dawn_batch_mssp(handles, G, _ptr(src, ctypes.c_int32), count_per_graph,
                                 _out_arg(dist), _ptr(offs, ctypes.c_int64), ctypes.byref(st)),
           "dawn_batch_mssp")
    rows = []
    if dist is not None:
        for i, g in enumerate(graphs):
            rows.append(dist[offs[i]: offs[i] + sizes[i]].reshape(count_per_graph, g.n))
    return rows, st.as_dict()


# ---- scheduler network --------------------------------------------------------

def tune(graph: "Graph", verbose: bool = False):
    """Autotune the launch knobs on `graph` and cache them for its class."""
    _check(lib().dawn_tune(graph._h, int(verbose)), "dawn_tune")


def tune_load(path: str):
    _check(lib().dawn_tune_load(path.encode()), "dawn_tune_load")


def tune_save(path: str):
    _check(lib().dawn_tune_save(path.encode()), "dawn_tune_save")


def sched_set_mlp(params: Optional[np.ndarray], mean=None, inv_std=None, single: bool = False):
    """Installs (None: clears) the batched engine's scheduler network, or with
    single=True the single-source engine's."""
    fn = lib().dawn_sched_set_mlp_single if single else lib().dawn_sched_set_mlp
    what = "dawn_sched_set_mlp_single" if single else "dawn_sched_set_mlp"
    if params is None:
        _check(fn(None, None, None), what)
        return
    p = np.ascontiguousarray(params, np.float32)
    assert p.size == SCHED_PARAMS
    mu = np.ascontiguousarray(mean if mean is not None else np.zeros(8), np.float32)
    isd = np.ascontiguousarray(inv_std if inv_std is not None else np.ones(8), np.float32)
    _check(fn(_ptr(p, ctypes.c_float), _ptr(mu, ctypes.c_float), _ptr(isd, ctypes.c_float)), what)


def mlp_forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32)
    out = np.zeros((len(xx), 3), np.float32)
    _check(lib().dawn_mlp_forward(_ptr(p, ctypes.c_float), _ptr(xx, ctypes.c_float), len(xx),
                                  _ptr(out, ctypes.c_float)), "dawn_mlp_forward")
    return out


def mlp_loss_grad(params: np.ndarray, x: np.ndarray, y: np.ndarray):
    p = np.ascontiguousarray(params, np.float32)
    xx = np.ascontiguousarray(x, np.float32)
    yy = np.ascontiguousarray(y, np.float32)
    loss = np.zeros(1, np.float32)
    grad = np.zeros(SCHED_PARAMS, np.float32)
    _check(lib().dawn_mlp_loss_grad(_ptr(p, ctypes.c_float), _ptr(xx, ctypes.c_float),
                                    _ptr(yy, ctypes.c_float), len(xx), _ptr(loss, ctypes.c_float),
                                    _ptr(grad, ctypes.c_float)), "dawn_mlp_loss_grad")
    return float(loss[0]), grad


def mlp_train(params: np.ndarray, x: np.ndarray, y: np.ndarray, epochs: int, batch: int,
              lr: float, seed: int):
    p = np.ascontiguousarray(params, np.float32).copy()
    xx = np.ascontiguousarray(x, np.float32)
    yy = np.ascontiguousarray(y, np.float32)
    loss = np.zeros(1, np.float32)
    _check(lib().dawn_mlp_train(_ptr(p, ctypes.c_float), _ptr(xx, ctypes.c_float),
                                _ptr(yy, ctypes.c_float), len(xx), epochs, batch, lr, seed,
                                _ptr(loss, ctypes.c_float)), "dawn_mlp_train")
    return p, float(loss[0])


def mlp_init(seed: int) -> np.ndarray:
    """He-uniform initial parameters in the layout of DAWN_SCHED_PARAMS."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((8, 32), (32, 32), (32, 3)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, size=(fan_out, fan_in)).ravel())
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)
