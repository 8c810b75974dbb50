"""Full-size parity at the headline configurations, through the C ABI.

- RMAT scale 24, ef 16, seed 99, 4096 sources (the benchmark workload, with
  its tune cache and learned network): the sums of all 4096 sources under
  the learned and the fixed schedule are equal; 16 sampled sources' rows
  and their sums from the full pass equal the oracle's; 8 more rows pass
  the shortest-path certificate; the reached-vertex and reached-edge
  totals (the GTEPS numerator) equal those computed from connected
  components (the rows of an undirected graph reach exactly the source's
  component).
- RMAT scale 20, ef 16, seed 3: all 4096 sums equal the oracle's.
- 2048^2 grid + 1% long-range edges, int32 weights, seed 2: all 16
  sources, single-source and as one 16-lane batched MSSP, equal the
  oracle's rows; closeness sums likewise.
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT, dawn

pytestmark = pytest.mark.gpu
INF = dawn.INF
TUNE_PATH = os.path.join(ROOT, "gal-dawn_amd", "tune_cache.txt")
MLP_PATH = os.path.join(ROOT, "gal-dawn_amd", "sched_mlp.json")


def log(msg):
    print(msg, flush=True)


def install_learned():
    w = json.load(open(MLP_PATH))
    dawn.sched_set_mlp(np.array(w["params"], np.float32), np.array(w["mean"], np.float32),
                       np.array(w["inv_std"], np.float32))


@pytest.fixture(scope="module")
def c5():
    log("case: RMAT scale 24 ef 16 seed 99, 4096 sources (tune cache + learned network)")
    dawn.tune_load(TUNE_PATH)
    install_learned()
    h = dawn.HostGraph.rmat(24, 16, 99)
    g = dawn.Graph(h)
    src = dawn.gen_sources(h.n, 4096, 99)
    log(f"  n={h.n} m={h.m} knobs={g.knobs()}")
    yield h, g, src
    g.close()
    dawn.sched_set_mlp(None)


@pytest.fixture(scope="module")
def c5_rows(c5, orc):
    """Rows of 24 sources spread over the four 1024-source batches of the
    full pass, and the oracle's rows of the first 16 of them."""
    h, g, src = c5
    pos = np.linspace(0, len(src) - 1, 24).astype(np.int64)
    chk = src[pos]
    log("  rows of 24 sampled sources (one batch), oracle rows of 16")
    dist, sums = g.mssp(chk, policy="learned")
    odist, osums, skipped, _ = orc.mssp(h.row_ptr, h.col, None, chk[:16])
    return pos, chk, dist, sums, odist, osums


@pytest.fixture(scope="module")
def c5_pass(c5):
    h, g, src = c5
    out = {}
    for pol in ("learned", "fixed"):
        log(f"  full pass, {pol} schedule (sums only)")
        _, sums, st = g.mssp(src, policy=pol, want_dist=False, stats=True)
        out[pol] = (sums, st)
        log(f"    steps {st['steps']} modes {st['mode_steps']} edges_reached {st['edges_reached']}")
    return out


def test_c5_sampled_rows_equal_oracle(c5, c5_rows, c5_pass):
    h, g, src = c5
    pos, chk, dist, sums, odist, osums = c5_rows
    assert np.array_equal(dist[:16], odist)
    assert np.array_equal(sums[:16], osums)
    # The same sources' sums inside the full 4096-source pass (their own
    # batches, two batches in flight) equal the oracle's.
    full = c5_pass["learned"][0]
    assert np.array_equal(full[pos[:16]], osums)
    assert np.array_equal(full[pos], sums)


def test_c5_more_rows_certified(c5, c5_rows, orc):
    h, g, src = c5
    pos, chk, dist, sums, _, _ = c5_rows
    for i in range(16, 24):
        s = int(chk[i])
        if h.degrees()[s] == 0:
            assert dist[i][s] == 0 and (dist[i] == INF).sum() == h.n - 1
            continue
        assert orc.certify(h.row_ptr, h.col, None, s, dist[i]), i


def test_c5_learned_and_fixed_sums_equal(c5, c5_pass):
    h, g, src = c5
    (sl, stl), (sf, stf) = c5_pass["learned"], c5_pass["fixed"]
    assert np.array_equal(sl, sf)
    assert stl["edges_reached"] == stf["edges_reached"]
    assert stl["vertices_reached"] == stf["vertices_reached"]
    # Isolated sources take no lane: batches of 1024 over the others.
    live = int((h.degrees()[src % h.n] > 0).sum())
    assert stl["skipped_sources"] == len(src) - live
    assert stl["batches"] == stf["batches"] == -(-live // 1024)


def test_c5_reached_totals_from_components(c5, c5_rows, c5_pass, orc):
    """vertices_reached / edges_reached of the full pass = sum over sources
    of |component(s)| and of the out-degrees in it: the giant component
    from a sampled row, small components by the oracle."""
    h, g, src = c5
    _, chk, dist, _, _, _ = c5_rows
    deg = h.degrees().astype(np.int64)
    big = int(np.argmax((dist != INF).sum(axis=1)))
    giant = dist[big] != INF
    giant_v, giant_e = int(giant.sum()), int(deg[giant].sum())
    log(f"  giant component: {giant_v} vertices, {giant_e} directed edges")
    exp_v = exp_e = 0
    other = {}
    for s in src:
        s = int(s) % h.n
        if deg[s] == 0:
            exp_v += 1
        elif giant[s]:
            exp_v += giant_v
            exp_e += giant_e
        else:
            if s not in other:
                row = orc.bfs(h.row_ptr, h.col, s, dense=False)
                r = row != INF
                other[s] = (int(r.sum()), int(deg[r].sum()))
            exp_v += other[s][0]
            exp_e += other[s][1]
    log(f"  {len(other)} sources outside the giant component")
    st = c5_pass["learned"][1]
    assert st["vertices_reached"] == exp_v
    assert st["edges_reached"] == exp_e


def test_c3_all_sums_equal_oracle(orc):
    log("case: RMAT scale 20 ef 16 seed 3, 4096 sources, all sums vs the oracle")
    h = dawn.HostGraph.rmat(20, 16, 3)
    g = dawn.Graph(h)
    srcs = dawn.gen_sources(h.n, 4096, 3)
    _, sums, st = g.mssp(srcs, want_dist=False, stats=True)
    _, osums, skipped, secs = orc.mssp(h.row_ptr, h.col, None, srcs, want_dist=False)
    log(f"  oracle {secs:.1f}s; engine {st['device_ms']:.1f} ms")
    assert np.array_equal(sums, osums)
    assert st["skipped_sources"] == int(skipped.sum())
    _, csums, _ = g.closeness(srcs, policy="learned")
    assert np.array_equal(csums, osums)
    g.close()


@pytest.fixture(scope="module")
def c2(orc):
    log("case: grid 2048x2048 + 1% long-range, w in [1,1000], seed 2, all 16 sources")
    h = dawn.HostGraph.grid(2048, 0.01, 2, weighted=True)
    g = dawn.Graph(h)
    srcs = dawn.gen_sources(h.n, 16, 2)
    odist, osums, _, secs = orc.mssp(h.row_ptr, h.col, h.weight, srcs)
    log(f"  oracle rows in {secs:.1f}s")
    yield h, g, srcs, odist, osums
    g.close()


def test_c2_single_source_all_16(c2):
    h, g, srcs, odist, _ = c2
    for i, s in enumerate(srcs):
        assert np.array_equal(g.sssp(int(s)), odist[i]), i


def test_c2_batched_16_sources(c2, orc):
    h, g, srcs, odist, osums = c2
    dist, sums, st = g.mssp(srcs, stats=True)
    assert np.array_equal(dist, odist)
    assert np.array_equal(sums, osums)
    assert st["batches"] == 1
    # Unique distances: Dijkstra on one row, the certificate on another.
    assert np.array_equal(dist[3], orc.dijkstra(h.row_ptr, h.col, h.weight, int(srcs[3])))
    assert orc.certify(h.row_ptr, h.col, h.weight, int(srcs[7]), dist[7])
    _, csums, _ = g.closeness(srcs)
    assert np.array_equal(csums, osums)
