"""Uneven multi-graph batch (one persistent launch) vs the CPU oracle,
per instance, bit-exact."""
import importlib

import numpy as np
import pytest

from conftest import dawn

pytestmark = pytest.mark.gpu
wl = importlib.import_module("gal_dawn_amd.workloads")


def check_batch(orc, graphs, src, spg, weighted):
    devs = [dawn.Graph(h) for h in graphs]
    rows, st = dawn.batch_mssp(devs, src, spg)
    for i, h in enumerate(graphs):
        s = src[i * spg:(i + 1) * spg]
        odist, _, _, _ = orc.mssp(h.row_ptr, h.col, h.weight if weighted else None, s)
        assert np.array_equal(rows[i], odist), (i, h.n)
    for g in devs:
        g.close()
    return st


@pytest.mark.parametrize("weighted", [False, True])
def test_batch_small_uneven(orc, weighted):
    print(f"case: 24 uneven graphs n in [1k, 20k], weighted={weighted}", flush=True)
    graphs, src, _ = wl.uneven(24, seed=7, weighted=weighted, nmax=20000)
    st = check_batch(orc, graphs, src, 8, weighted)
    assert st["steps"] > 0


@pytest.mark.parametrize("spg", [3, 12, 24, 40, 64])
def test_batch_lane_widths(orc, spg):
    """Tasks keep their lanes in the narrowest word that holds them (1, 2, 4
    or 8 bytes per vertex); every width gives the oracle's rows."""
    print(f"case: 10 uneven graphs n in [1k, 8k], {spg} sources each", flush=True)
    graphs, src, _ = wl.uneven(10, seed=11, spg=spg, nmax=8000)
    check_batch(orc, graphs, src, spg, False)


def test_batch_levels_beyond_16_bits(orc):
    """A path of 70,000 vertices: levels past 65,534 do not fit the 16-bit
    level stage and go to the rows directly; rows are |i - s|."""
    n = 70000
    print(f"case: path of {n} vertices, 3 sources, plus a small ER graph", flush=True)
    u = np.arange(n - 1, dtype=np.int32)
    path = dawn.HostGraph.from_edges(n, u, u + 1)
    er = dawn.HostGraph.er(500, 3.0, 9)
    src = np.array([0, n - 1, 20000, 1, 2, 3], np.int32)
    check_batch(orc, [path, er], src, 3, False)
    g = dawn.Graph(path)
    want = np.abs(np.arange(n)[None, :] - src[:3, None]).astype(np.int32)
    got, _ = dawn.batch_mssp([g], src[:3], 3)
    assert np.array_equal(got[0], want) and got[0].max() == n - 1
    g.close()


def test_batch_isolated_and_duplicate_sources(orc):
    h = dawn.HostGraph.from_edges(50, np.arange(0, 20, dtype=np.int32), np.arange(1, 21, dtype=np.int32))
    h2 = dawn.HostGraph.er(300, 3.0, 3)
    src = np.array([30, 0, 0, 5, 49, 2, 2, 2], np.int32)  # isolated 30/49, duplicates
    src2 = dawn.gen_sources(300, 8, 3)
    check_batch(orc, [h, h2], np.concatenate([src, src2]), 8, False)


def test_batch_c4_full_unweighted(orc):
    print("case: 512 uneven graphs (seed 4), 8 sources each, unweighted", flush=True)
    graphs, src, _ = wl.uneven(512, seed=4)
    check_batch(orc, graphs, src, 8, False)


def test_batch_c4_full_weighted(orc):
    print("case: 512 uneven graphs (seed 4), 8 sources each, int32 weights", flush=True)
    graphs, src, _ = wl.uneven(512, seed=4, weighted=True)
    check_batch(orc, graphs, src, 8, True)


def test_batch_split_tasks_duplicates_isolated(orc):
    """Graphs of >= 500k and >= 2M edges run as two tasks of 4 lanes and as
    one task per source; duplicate and isolated sources inside split tasks."""
    print("case: RMAT-17 (one task per source), RMAT-16 (two tasks), ER 2000", flush=True)
    big = dawn.HostGraph.rmat(17, 16, 41)
    mid = dawn.HostGraph.rmat(16, 16, 42)
    small = dawn.HostGraph.er(2000, 4.0, 43)
    assert big.m >= 2000000 and 500000 <= mid.m < 2000000
    iso = int(np.flatnonzero(big.degrees() == 0)[0])
    s_big = dawn.gen_sources(big.n, 8, 41)
    s_big[3] = s_big[1]
    s_big[6] = iso
    s_mid = dawn.gen_sources(mid.n, 8, 42)
    s_mid[5] = s_mid[4]
    s_small = dawn.gen_sources(small.n, 8, 43)
    st = check_batch(orc, [small, big, mid], np.concatenate([s_small, s_big, s_mid]), 8, False)
    assert st["batches"] == 1 + 8 + 2  # tasks


def test_batch_weighted_multigraph_and_directed(orc):
    """Weighted tasks (near-far relaxation) on a multigraph whose sources
    list duplicate edges with other weights (the last listed seeds), with
    zero-weight edges, and on a directed graph with weights spanning three
    orders of magnitude."""
    rng = np.random.default_rng(17)
    n = 3000
    u = rng.integers(0, n, 20000).astype(np.int32)
    v = rng.integers(0, n, 20000).astype(np.int32)
    w = rng.integers(0, 1000, 20000).astype(np.int32)  # zero weights included
    at = u < 40
    u = np.concatenate([u, u[at], u[at]])
    v = np.concatenate([v, v[at], v[at]])
    w = np.concatenate([w, rng.integers(1, 1000, int(at.sum())).astype(np.int32),
                        rng.integers(1, 1000, int(at.sum())).astype(np.int32)])
    multi = dawn.HostGraph.from_edges(n, u, v, w, symmetrize=True, dedup=False)
    du = rng.integers(0, 5000, 40000).astype(np.int32)
    dv = rng.integers(0, 5000, 40000).astype(np.int32)
    dw = (10 ** rng.integers(0, 3, 40000) * rng.integers(1, 10, 40000)).astype(np.int32)
    directed = dawn.HostGraph.from_edges(5000, du, dv, dw, symmetrize=False, dedup=True)
    src = np.concatenate([np.arange(0, 40, 5, dtype=np.int32), dawn.gen_sources(5000, 8, 17)])
    check_batch(orc, [multi, directed], src, 8, True)
