"""Device-side graph build (dawn_graph_from_edges, dawn_graph_gen_*) through
the C ABI against the host build (dawn_csr_from_edges / dawn_gen_* +
dawn_graph_create): the CSR downloaded from the device equals the host's
array for array -- generators at small and benchmark sizes (the host
build's digests are committed in tests/golden/generator_digests.json),
edge lists with self-loops and duplicates of different weights, directed
and undirected, dedup on and off, host and device input arrays -- and
traversals on device-built graphs (the pull step reads the device-built
CSC of directed graphs) equal the oracle's."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import ROOT, dawn

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "generator_digests.json")))


def log(msg):
    print(msg, flush=True)


def digest(h):
    m = hashlib.sha256()
    m.update(h.row_ptr.tobytes())
    m.update(h.col.tobytes())
    if h.weighted:
        m.update(h.weight.tobytes())
    return m.hexdigest()[:16]


def same_csr(a, b):
    assert (a.n, a.m, a.directed, a.weighted) == (b.n, b.m, b.directed, b.weighted)
    assert np.array_equal(a.row_ptr, b.row_ptr)
    assert np.array_equal(a.col, b.col)
    if a.weighted:
        assert np.array_equal(a.weight, b.weight)


@pytest.mark.parametrize("what", ["er", "er_w", "rmat", "rmat_w", "grid_w", "grid"])
def test_small_generators_match_host(orc, what):
    log(f"case: device generator {what}")
    if what.startswith("er"):
        args = (3000, 6.0, 7)
        h, g = dawn.HostGraph.er(*args, weighted=what.endswith("_w")), \
            dawn.Graph.er_device(*args, weighted=what.endswith("_w"))
    elif what.startswith("rmat"):
        args = (12, 8, 5)
        h, g = dawn.HostGraph.rmat(*args, weighted=what.endswith("_w")), \
            dawn.Graph.rmat_device(*args, weighted=what.endswith("_w"))
    else:
        args = (40, 0.02, 6)
        h, g = dawn.HostGraph.grid(*args, weighted=what.endswith("_w")), \
            dawn.Graph.grid_device(*args, weighted=what.endswith("_w"))
    same_csr(g.download(), h)
    s = int(np.argmax(h.degrees()))
    if h.weighted:
        assert np.array_equal(g.sssp(s), orc.dijkstra(h.row_ptr, h.col, h.weight, s))
    else:
        assert np.array_equal(g.bfs(s), orc.bfs(h.row_ptr, h.col, s))


def test_golden_digests_small():
    assert digest(dawn.Graph.er_device(4096, 8.0, 1).download()) == GOLD["er_n4096_d8_s1"]
    assert digest(dawn.Graph.rmat_device(12, 16, 3).download()) == GOLD["rmat_s12_ef16_s3"]
    assert digest(dawn.Graph.grid_device(64, 0.01, 2, weighted=True).download()) == \
        GOLD["grid_64_1pct_s2_w"]


@pytest.mark.parametrize("sym,dedup,weighted", [(s, d, w) for s in (True, False)
                                                 for d in (True, False) for w in (False, True)])
def test_edge_lists_match_host(orc, sym, dedup, weighted):
    """Self-loops, duplicate pairs with different weights, isolated vertices."""
    log(f"case: device edge list sym={sym} dedup={dedup} weighted={weighted}")
    rng = np.random.default_rng(61)
    n = 700
    u = rng.integers(0, 650, 4000).astype(np.int32)
    v = rng.integers(0, 650, 4000).astype(np.int32)
    u[:40] = v[:40]  # self-loops
    u = np.concatenate([u, u[100:400]])
    v = np.concatenate([v, v[100:400]])
    w = rng.integers(0, 50, len(u)).astype(np.int32) if weighted else None
    h = dawn.HostGraph.from_edges(n, u, v, w, symmetrize=sym, dedup=dedup)
    g = dawn.Graph.from_edges_device(n, u, v, w, symmetrize=sym, dedup=dedup)
    same_csr(g.download(), h)
    # Device input arrays give the same graph.
    du = dawn.DeviceArray(len(u), np.int32).upload(u)
    dv = dawn.DeviceArray(len(v), np.int32).upload(v)
    dw = dawn.DeviceArray(len(w), np.int32).upload(w) if weighted else None
    same_csr(dawn.Graph.from_edges_device(n, du, dv, dw, symmetrize=sym, dedup=dedup).download(), h)
    # Traversals: every mode, pull over the device-built CSC when directed.
    deg = h.degrees()
    for s in [int(np.argmax(deg)), int(np.flatnonzero(deg > 0)[3])]:
        if weighted:
            # The reference's seeding with duplicate edges (sssp.cpp:40-46).
            want = orc.sssp(h.row_ptr, h.col, h.weight, s, dense=False)
            for p in ("fixed", "push", "pull"):
                assert np.array_equal(g.sssp(s, policy=p), want), (s, p)
        if not weighted:
            want = orc.bfs(h.row_ptr, h.col, s)
            for p in ("fixed", "push", "pull", "bitmap"):
                assert np.array_equal(g.bfs(s, policy=p), want), (s, p)
    if not weighted:
        srcs = np.flatnonzero(deg > 0)[:100].astype(np.int32)
        dist, sums = g.mssp(srcs, policy="pull")
        odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
        assert np.array_equal(dist, odist) and np.array_equal(sums, osums)


def test_invalid_edges_rejected():
    u = np.array([0, 1, 5], np.int32)
    v = np.array([1, 2, 3], np.int32)
    with pytest.raises(dawn.DawnError) as e:
        dawn.Graph.from_edges_device(5, u, v)
    assert e.value.status == -1
    with pytest.raises(dawn.DawnError):
        dawn.Graph.from_edges_device(6, u, v, np.array([1, -2, 3], np.int32))
    g = dawn.Graph.from_edges_device(6, u, v)  # the same edges are fine with n = 6
    assert g.m == 6


def test_empty_and_loop_only_edge_lists():
    g = dawn.Graph.from_edges_device(4, np.zeros(0, np.int32), np.zeros(0, np.int32))
    h = g.download()
    assert h.m == 0 and list(h.row_ptr) == [0] * 5
    g = dawn.Graph.from_edges_device(4, np.array([2, 3], np.int32), np.array([2, 3], np.int32))
    assert g.m == 0


@pytest.mark.parametrize("what", ["rmat20", "grid2048_w", "rmat24"])
def test_benchmark_graphs_match_host_digests(orc, what):
    """The benchmark graphs built on the device equal the host build (the
    committed digests of the host generators' arrays)."""
    log(f"case: device build of {what}")
    if what == "rmat20":
        g, key = dawn.Graph.rmat_device(20, 16, 3), "rmat_s20_ef16_s3"
    elif what == "grid2048_w":
        g, key = dawn.Graph.grid_device(2048, 0.01, 2, weighted=True), "grid_2048_1pct_s2_w"
    else:
        g, key = dawn.Graph.rmat_device(24, 16, 99), "rmat_s24_ef16_s99"
    h = g.download()
    assert digest(h) == GOLD[key]
    src = dawn.gen_sources(h.n, 4, 3)
    s = int(src[np.argmax(h.degrees()[src])])
    if h.weighted:
        assert orc.certify(h.row_ptr, h.col, h.weight, s, g.sssp(s))
    else:
        assert orc.certify(h.row_ptr, h.col, None, s, g.bfs(s))
    g.close()
