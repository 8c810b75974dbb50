"""HIP engine vs the CPU oracle, bit-exact, through the C ABI.

Sizes: the small cases run every mode/policy; the headline workloads
(ER n=65,536 seed 1; 2048^2 grid + 1% long-range seed 2; RMAT scale 20
seed 3) run at full size with the oracle on sampled sources plus
size-independent checks (BFS level equations, symmetry of undirected
distances, sums == row sums).
"""
import numpy as np
import pytest

from conftest import dawn

INF = dawn.INF
pytestmark = pytest.mark.gpu
POLICIES = ["fixed", "push", "pull", "bitmap", "learned"]


def log(msg):
    print(msg, flush=True)


def bfs_level_equations_hold(h, source, d):
    """d is the BFS distance from source iff d[s]=0 and, for v != s,
    d[v] = 1 + min over in-neighbours (INF if none is reached)."""
    n = h.n
    rp, col = h.row_ptr, h.col
    rows = np.repeat(np.arange(n), np.diff(rp))
    # For undirected graphs in-neighbours == out-neighbours.
    du = d[rows].astype(np.int64)
    best = np.full(n, np.iinfo(np.int64).max)
    np.minimum.at(best, col, du)
    exp = np.where(best >= INF, INF, best + 1)
    exp[source] = 0
    return np.array_equal(exp, d.astype(np.int64))


@pytest.fixture(scope="module")
def er_c1():
    log("case: ER n=65536 avg deg 8 seed 1")
    h = dawn.HostGraph.er(65536, 8.0, 1)
    return h, dawn.Graph(h)


def test_bfs_c1_all_policies(er_c1, orc):
    h, g = er_c1
    srcs = [0] + list(dawn.gen_sources(h.n, 64, 1))
    for i, s in enumerate(srcs):
        if h.degrees()[s] == 0:
            continue
        ref = orc.bfs(h.row_ptr, h.col, s)
        pols = POLICIES if i < 4 else ["fixed"]
        for p in pols:
            d = g.bfs(s, policy=p)
            assert np.array_equal(d, ref), (s, p)


def test_bfs_fixed_schedule_matches_gapbs(er_c1, orc):
    h, g = er_c1
    for s in [0, 17, 4242]:
        modes = orc.dobfs_modes(h.row_ptr, h.col, s)
        d, st = g.bfs(s, policy="fixed", stats=True)
        assert st["steps"] == len(modes)
        assert st["mode_steps"]["pull"] == int((modes == 1).sum())
        # A pull step writes its level's bitmap only, so the top-down step
        # right after one reads the bitmap (BITMAP); the others the queue.
        after_pull = int(((modes[1:] == 0) & (modes[:-1] == 1)).sum())
        assert st["mode_steps"]["bitmap"] == after_pull
        assert st["mode_steps"]["push"] == int((modes == 0).sum()) - after_pull


def test_mssp_c1_batched_equals_oracle(er_c1, orc):
    h, g = er_c1
    srcs = np.concatenate([[0], dawn.gen_sources(h.n, 64, 1)]).astype(np.int32)
    for p in POLICIES:
        dist, sums = g.mssp(srcs, policy=p)
        odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p


def test_sssp_small_grid_all_policies(orc):
    log("case: grid 128x128 + 1% long-range, w in [1,1000], seed 2")
    h = dawn.HostGraph.grid(128, 0.01, 2, weighted=True)
    g = dawn.Graph(h)
    for s in dawn.gen_sources(h.n, 4, 2):
        ref = orc.sssp(h.row_ptr, h.col, h.weight, s, dense=False)
        assert np.array_equal(ref, orc.dijkstra(h.row_ptr, h.col, h.weight, s))
        for p in ["fixed", "push", "pull", "bitmap"]:
            assert np.array_equal(g.sssp(s, policy=p), ref), (s, p)


def test_sssp_c2_full_size(orc):
    log("case: grid 2048x2048 + 1% long-range, w in [1,1000], seed 2 (2 of 16 sources)")
    h = dawn.HostGraph.grid(2048, 0.01, 2, weighted=True)
    g = dawn.Graph(h)
    for s in dawn.gen_sources(h.n, 16, 2)[:2]:
        d = g.sssp(s)
        ref = orc.sssp(h.row_ptr, h.col, h.weight, s, dense=False)
        assert np.array_equal(d, ref)


def test_weighted_mssp_and_closeness(orc):
    h = dawn.HostGraph.er(3000, 6.0, 7, weighted=True)
    g = dawn.Graph(h)
    srcs = np.array([0, 5, 9, 2999], np.int32)
    dist, sums = g.mssp(srcs)
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, h.weight, srcs)
    assert np.array_equal(dist, odist)
    assert np.array_equal(sums, osums)


def test_mssp_rmat_small_all_rows(orc):
    log("case: RMAT scale 14 ef 16 seed 3, 200 sources")
    h = dawn.HostGraph.rmat(14, 16, 3)
    g = dawn.Graph(h)
    srcs = dawn.gen_sources(h.n, 200, 3)
    odist, osums, skipped, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    for p in POLICIES:
        dist, sums, st = g.mssp(srcs, policy=p, stats=True)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p
        assert st["skipped_sources"] == int(skipped.sum())


def test_closeness_exact(orc):
    h = dawn.HostGraph.rmat(12, 8, 5)
    g = dawn.Graph(h)
    srcs = dawn.gen_sources(h.n, 70, 5)
    cc, sums, ref_int = g.closeness(srcs)
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    for i in range(len(srcs)):
        c, s, r = orc.closeness(odist[i])
        assert sums[i] == s
        assert cc[i] == c  # bit-exact double
        assert ref_int[i] == r


def test_directed_graph_pull_uses_csc(orc):
    rng = np.random.default_rng(4)
    n, m = 5000, 30000
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    h = dawn.HostGraph.from_edges(n, u, v, None, symmetrize=False)
    hw = dawn.HostGraph.from_edges(n, u, v, rng.integers(1, 100, m).astype(np.int32), symmetrize=False)
    g, gw = dawn.Graph(h), dawn.Graph(hw)
    srcs = [s for s in range(0, n, 500) if h.degrees()[s] > 0]
    for s in srcs:
        ref = orc.bfs(h.row_ptr, h.col, s)
        for p in POLICIES:
            assert np.array_equal(g.bfs(s, policy=p), ref), (s, p)
        refw = orc.sssp(hw.row_ptr, hw.col, hw.weight, s)
        for p in ["fixed", "push", "pull", "bitmap"]:
            assert np.array_equal(gw.sssp(s, policy=p), refw), (s, p)
    dist, _ = g.mssp(np.array(srcs, np.int32), policy="pull")
    odist, _, _, _ = orc.mssp(h.row_ptr, h.col, None, np.array(srcs, np.int32))
    assert np.array_equal(dist, odist)


def test_isolated_source_status(orc):
    h = dawn.HostGraph.from_edges(10, np.array([0, 1], np.int32), np.array([1, 2], np.int32))
    g = dawn.Graph(h)
    with pytest.raises(dawn.DawnError) as e:
        g.bfs(7)
    assert e.value.status == dawn.ERR_ISOLATED_SOURCE
    dist, sums, st = g.mssp(np.array([7, 0], np.int32), stats=True)
    assert st["skipped_sources"] == 1
    assert dist[0][7] == 0 and (np.delete(dist[0], 7) == INF).all()
    assert list(dist[1][:3]) == [0, 1, 2]


def test_mssp_c3_full_size_properties(orc):
    log("case: RMAT scale 20 ef 16 seed 3, 4096 sources (64 checked by oracle)")
    h = dawn.HostGraph.rmat(20, 16, 3)
    g = dawn.Graph(h)
    srcs = dawn.gen_sources(h.n, 4096, 3)
    _, sums_all = g.mssp(srcs, want_dist=False)
    sub = srcs[:128]
    dist, sums = g.mssp(sub)
    assert np.array_equal(sums, sums_all[:128])
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, sub[:64])
    assert np.array_equal(dist[:64], odist)
    for i in range(64, 128, 8):
        assert bfs_level_equations_hold(h, int(sub[i]), dist[i])
    # symmetry of undirected hop distances between sources
    for i in range(0, 128, 16):
        for j in range(0, 128, 16):
            assert dist[i][sub[j]] == dist[j][sub[i]]


def test_tuned_knobs_keep_results_exact(orc, tmp_path):
    """Every knob setting the tuner can pick must leave results unchanged."""
    log("case: tuner on RMAT scale 13 ef 16 seed 4 and weighted grid 96")
    h = dawn.HostGraph.rmat(13, 16, 4)
    g = dawn.Graph(h)
    srcs = dawn.gen_sources(h.n, 96, 4)
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    dawn.tune(g)
    dist, sums = g.mssp(srcs)
    assert np.array_equal(dist, odist) and np.array_equal(sums, osums)
    # A graph created afterwards in the same class starts from the cache.
    g2 = dawn.Graph(h)
    assert np.array_equal(g2.mssp(srcs)[0], odist)
    hw = dawn.HostGraph.grid(96, 0.01, 4, weighted=True)
    gw = dawn.Graph(hw)
    dawn.tune(gw)
    od = orc.sssp(hw.row_ptr, hw.col, hw.weight, 5)
    assert np.array_equal(gw.sssp(5), od)
    p = str(tmp_path / "tune.txt")
    dawn.tune_save(p)
    assert len(open(p).read().splitlines()) >= 2


@pytest.mark.parametrize("words", [1, 2, 4, 8, 16])
def test_lane_words_exact(orc, words):
    """64, 128, 256, 512 and 1024 sources per batch, with a partial last
    batch (at 1024: six empty lane words) and duplicate sources: rows and
    sums equal the oracle's."""
    log(f"case: RMAT scale 13 ef 8 seed 6, 700 sources, lane_words={words}")
    h = dawn.HostGraph.rmat(13, 8, 6)
    g = dawn.Graph(h)
    g.set_knobs(lane_words=words)
    assert g.knobs()["lane_words"] == words
    srcs = dawn.gen_sources(h.n, 700, 6)
    srcs[7] = srcs[3]
    odist, osums, skipped, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    for p in ("fixed", "pull", "push", "bitmap"):
        dist, sums, st = g.mssp(srcs, policy=p, stats=True)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p
        assert st["skipped_sources"] == int(skipped.sum())
    _, lv = g.mssp_profile(srcs)
    # Isolated sources take no lane; the width is the knob, or less when
    # fewer lanes are needed.
    live = int((h.degrees()[srcs] > 0).sum())
    w = 1
    while w < words and 64 * w < live:
        w *= 2
    assert len(lv) == -(-live // (64 * w)) and (lv > 0).all()


@pytest.mark.parametrize("words,variant", [(1, 1), (4, 0), (4, 1), (4, 2), (4, 3), (16, 0), (16, 2), (16, 3)])
def test_pull_variants_exact(orc, words, variant):
    """Every light-pull variant (variant 1 leaves the fold of every vertex to
    the finalize pass; the others fold light vertices in the pull kernel and
    heavy ones in a separate pass), with heavy chunks present: rows and sums
    equal the oracle's under the pull-only and fixed schedules."""
    log(f"case: RMAT scale 13 ef 16 seed 21, 600 sources, lane_words={words}, pull_variant={variant}")
    h = dawn.HostGraph.rmat(13, 16, 21)
    g = dawn.Graph(h)
    g.set_knobs(lane_words=words, pull_variant=variant, heavy_degree=16, pull_chunk=64)
    srcs = dawn.gen_sources(h.n, 600, 21)
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    for p in ("pull", "fixed", "learned"):
        dist, sums = g.mssp(srcs, policy=p)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p


@pytest.mark.parametrize("words,count", [(1, 50), (8, 300), (16, 1100)])
def test_long_traversals_flush_windows(orc, words, count):
    """Traversals deeper than one captured chunk (a 40x40 grid: up to 78
    levels) flush the level history in several windows; every row entry
    equals the oracle's, and the grid's unit distances are Manhattan."""
    log(f"case: grid 40x40 no shortcuts, {count} sources, lane_words={words}")
    side = 40
    h = dawn.HostGraph.grid(side, 0.0, 3)
    g = dawn.Graph(h)
    g.set_knobs(lane_words=words)
    srcs = dawn.gen_sources(h.n, count, 3)
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    ys, xs = np.divmod(np.arange(h.n), side)
    for i in range(0, count, max(1, count // 7)):
        sy, sx = divmod(int(srcs[i]), side)
        assert np.array_equal(odist[i], np.abs(ys - sy) + np.abs(xs - sx))
    for p in ("fixed", "pull", "push"):
        dist, sums = g.mssp(srcs, policy=p)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p
    _, csums, _ = g.closeness(srcs, policy="learned")  # sums only, every window
    assert np.array_equal(csums, osums)


@pytest.mark.parametrize("n", [10, 100])
def test_tiny_graph_wide_batch(orc, n):
    """A path far smaller than one batch (1100 sources taken mod n, so every
    vertex is the source of many lanes; padded vertex range): rows are the
    closed form |i - s| under every schedule."""
    log(f"case: path of {n} vertices, 1100 sources, lane_words=16")
    u = np.arange(n - 1, dtype=np.int32)
    h = dawn.HostGraph.from_edges(n, u, u + 1)
    g = dawn.Graph(h)
    g.set_knobs(lane_words=16)
    srcs = dawn.gen_sources(1 << 20, 1100, 5)
    want = np.abs(np.arange(n)[None, :] - (srcs % n)[:, None]).astype(np.int32)
    for p in ("fixed", "pull", "push", "bitmap"):
        dist, sums = g.mssp(srcs, policy=p)
        assert np.array_equal(dist, want), p
        assert np.array_equal(sums, want.sum(axis=1)), p


@pytest.mark.parametrize("count", [1, 9, 17, 33, 70])
def test_weighted_batched_mssp_exact(orc, count):
    """Batched min-plus (8, 16, 32, 64 lanes; a partial second batch at 70):
    rows and sums equal the oracle's, with a duplicate source."""
    log(f"case: grid 40 + 2% seed 9 weighted, {count} sources")
    h = dawn.HostGraph.grid(40, 0.02, 9, weighted=True)
    g = dawn.Graph(h)
    srcs = dawn.gen_sources(h.n, count, 9)
    if count > 5:
        srcs[5] = srcs[2]
    odist, osums, skipped, _ = orc.mssp(h.row_ptr, h.col, h.weight, srcs)
    deg = h.degrees().astype(np.int64)
    reached = odist != INF
    for p in ("fixed", "push", "pull"):
        dist, sums, st = g.mssp(srcs, policy=p, stats=True)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p
        assert st["batches"] == -(-count // 64)
        # Reached vertices and the out-degrees of reached vertices (GTEPS).
        assert st["vertices_reached"] == int(reached.sum())
        assert st["edges_reached"] == int((reached * deg[None, :]).sum())
        if p == "pull":
            assert st["mode_steps"]["push"] == 0 and st["mode_steps"]["pull"] > 0
        if p == "push":
            assert st["mode_steps"]["pull"] == 0


def test_weighted_batched_multigraph_and_isolated(orc):
    """Duplicate edges with different weights keep the reference's seeding
    (a source's later listed duplicate wins, sssp.cpp:40-46); isolated
    sources give INF rows with 0 at the source and a zero sum."""
    log("case: weighted multigraph n=300 with isolated vertices")
    rng = np.random.default_rng(11)
    n = 300
    u = rng.integers(0, 250, 900).astype(np.int32)  # 250..299 isolated
    v = rng.integers(0, 250, 900).astype(np.int32)
    w = rng.integers(1, 50, 900).astype(np.int32)
    u = np.concatenate([u, u[:200]])  # duplicates with other weights
    v = np.concatenate([v, v[:200]])
    w = np.concatenate([w, rng.integers(1, 50, 200).astype(np.int32)])
    h = dawn.HostGraph.from_edges(n, u, v, w, symmetrize=True, dedup=False)
    g = dawn.Graph(h)
    deg = np.diff(h.row_ptr)
    srcs = np.array([0, 1, 2, 260, 3, 299, int(np.argmax(deg))], np.int32)
    odist, osums, skipped, _ = orc.mssp(h.row_ptr, h.col, h.weight, srcs)
    for p in ("fixed", "push", "pull"):
        dist, sums, st = g.mssp(srcs, policy=p, stats=True)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p
        assert st["skipped_sources"] == int(skipped.sum()) == 2


@pytest.mark.parametrize("count", [5, 40, 64])
def test_weighted_batched_skewed_directed(orc, count):
    """Batched min-plus on a skewed directed graph: light frontier vertices
    (out-degree < 16, one lane each walking their active source lanes) and
    heavy ones (a wave each) in the same push steps; pull steps over the
    in-edges (CSC), vertices of in-degree >= heavy_degree by in-edge chunks
    (a wave each, several chunks per vertex at heavy_degree 16 / 64-edge
    chunks); rows equal Dijkstra's under push, pull and the rule."""
    log(f"case: weighted directed RMAT scale 12 ef 8 seed 17, {count} sources")
    h = dawn.HostGraph.rmat(12, 8, 17, weighted=True)
    rows = np.repeat(np.arange(h.n, dtype=np.int32), np.diff(h.row_ptr))
    keep = (rows + h.col) % 3 != 0  # drop a third of the arcs: asymmetric
    rp, cc = _directed_csr(h.n, rows[keep], h.col[keep])
    order = np.lexsort((h.col[keep], rows[keep]))
    ww = h.weight[keep][order]
    hd = dawn.HostGraph.from_csr(rp, cc, ww, directed=True)
    g = dawn.Graph(hd)
    deg = np.diff(rp)
    assert (deg >= 16).any() and (deg < 16).any()
    srcs = dawn.gen_sources(hd.n, count, 17)
    srcs[0] = int(np.argmax(deg))
    want = np.stack([orc.dijkstra(rp, cc, ww, int(s)) for s in srcs])
    indeg = np.bincount(cc, minlength=hd.n)
    assert indeg.max() > 2 * 64  # several in-edge chunks per heavy vertex below
    for knobs in ({}, {"heavy_degree": 16, "pull_chunk": 64}):
        if knobs:
            g.set_knobs(**knobs)
        for p in ("fixed", "push", "pull"):
            dist, sums = g.mssp(srcs, policy=p)
            for i in range(count):
                assert np.array_equal(dist[i], want[i]), (knobs, p, i)


@pytest.mark.parametrize("directed", [False, True])
def test_hub_order_view_exact(orc, directed):
    """Hub-first renumbering of the batched engine (graphs >= 65,536
    vertices): rows in the original order, identical with and without it."""
    log(f"case: RMAT scale 16 ef 8 seed 12, directed={directed}, hub order on/off")
    h = dawn.HostGraph.rmat(16, 8, 12)
    if directed:
        # Keep one direction of each edge (lower id -> higher id).
        rows = np.repeat(np.arange(h.n, dtype=np.int32), np.diff(h.row_ptr))
        keep = rows < h.col
        h = dawn.HostGraph.from_csr(*_directed_csr(h.n, rows[keep], h.col[keep]), directed=True)
    g = dawn.Graph(h)
    srcs = dawn.gen_sources(h.n, 130, 12)
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    for hub in (1, 0):
        g.set_knobs(hub_order=hub)
        dist, sums = g.mssp(srcs, policy="fixed")
        assert np.array_equal(dist, odist), hub
        assert np.array_equal(sums, osums), hub


@pytest.mark.parametrize("directed", [False, True])
def test_wide_batches_hub_view_exact(orc, directed):
    """1024 sources per batch (records split over lane pairs), two batches in
    flight on two streams, hub-first view, every step mode: rows and sums
    equal the oracle's."""
    log(f"case: RMAT scale 16 ef 8 seed 14, directed={directed}, 1100 sources, lane_words=16")
    h = dawn.HostGraph.rmat(16, 8, 14)
    if directed:
        rows = np.repeat(np.arange(h.n, dtype=np.int32), np.diff(h.row_ptr))
        keep = rows > h.col
        h = dawn.HostGraph.from_csr(*_directed_csr(h.n, rows[keep], h.col[keep]), directed=True)
    g = dawn.Graph(h)
    g.set_knobs(lane_words=16, hub_order=1)
    # 1100 sources with out-edges (two batches: isolated sources take no
    # lane) and 9 isolated ones interleaved; one duplicate.
    deg = h.degrees()
    pool = dawn.gen_sources(h.n, 6000, 14)
    srcs = pool[deg[pool] > 0][:1100].copy()
    srcs[900] = srcs[5]
    srcs = np.insert(srcs, np.arange(0, 1100, 123), np.flatnonzero(deg == 0)[:9]).astype(np.int32)
    odist, osums, skipped, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    assert int(skipped.sum()) == 9
    for p in ("fixed", "pull", "push", "bitmap"):
        dist, sums, st = g.mssp(srcs, policy=p, stats=True)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p
        assert st["batches"] == 2 and st["skipped_sources"] == 9


@pytest.mark.parametrize("count", [1024 + 40, 1024 + 100, 2048 + 200, 3072 + 300, 1024 + 700])
def test_narrow_last_batch_exact(orc, count):
    """A partly filled last batch runs at the fewest lane words holding its
    lanes (40 lanes: 1 word, 100: 2, 200: 4, 300: 8; 700 needs all 16), on
    the stream of its turn (after 1 or 3 full batches the second stream,
    after 2 the first): host rows, device rows, sums-only closeness and the
    profiling pass's batches equal the oracle's."""
    log(f"case: RMAT scale 12 ef 8 seed 31, {count} live sources, lane_words=16")
    h = dawn.HostGraph.rmat(12, 8, 31)
    g = dawn.Graph(h)
    g.set_knobs(lane_words=16, hub_order=1)
    deg = h.degrees()
    rng = np.random.default_rng(count)
    srcs = rng.choice(np.flatnonzero(deg > 0), count).astype(np.int32)  # duplicates too
    srcs = np.insert(srcs, [5, count // 2], np.flatnonzero(deg == 0)[:2]).astype(np.int32)
    odist, osums, skipped, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    assert int(skipped.sum()) == 2
    for p in ("fixed", "learned", "pull", "push"):
        dist, sums, st = g.mssp(srcs, policy=p, stats=True)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p
        assert st["batches"] == -(-count // 1024) and st["skipped_sources"] == 2
    dev = dawn.DeviceArray((len(srcs), h.n), np.int32).fill_bytes(0x5A)
    _, sums = g.mssp(srcs, policy="learned", out=dev)
    assert np.array_equal(dev.download(), odist) and np.array_equal(sums, osums)
    dev.close()
    _, csums, _ = g.closeness(srcs, policy="learned")
    assert np.array_equal(csums, osums)
    _, lv = g.mssp_profile(srcs, policy="learned")
    assert len(lv) == -(-count // 1024) and (lv > 0).all()
    # The per-step mode timings record the same batches: one row per step
    # of the fixed schedule's traversal.
    _, _, st = g.mssp(srcs, policy="fixed", want_dist=False, stats=True)
    modes, _, times = g.sched_trace(srcs, policy="fixed")
    assert len(modes) == st["steps"] and (times > 0).all()


def _directed_csr(n, u, v):
    order = np.lexsort((v, u))
    u, v = u[order], v[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, u + 1, 1)
    return np.cumsum(rp).astype(np.int32), v.astype(np.int32)


@pytest.mark.parametrize("words", [1, 16])
def test_zero_in_degree_sources_exact(orc, words):
    """Vertices of in-degree 0 start done (no step reads them): sources of
    in-degree 0, a batch whose lanes all share one such source, and isolated
    vertices keep rows equal to the oracle's under every policy."""
    log(f"case: directed layered graph n=3000 with in-degree-0 sources, lane_words={words}")
    rng = np.random.default_rng(21)
    n = 3000
    roots = np.arange(0, 200, dtype=np.int32)          # in-degree 0, out-edges only
    mid = np.arange(200, 2600, dtype=np.int32)         # reached from roots and each other
    u = np.concatenate([rng.choice(roots, 4000), rng.choice(mid, 9000)]).astype(np.int32)
    v = np.concatenate([rng.choice(mid, 4000), rng.choice(mid, 9000)]).astype(np.int32)
    h = dawn.HostGraph.from_csr(*_directed_csr(n, u, v), directed=True)  # 2600.. isolated
    g = dawn.Graph(h)
    g.set_knobs(lane_words=words)
    lanes = 64 * words
    batches = [
        rng.choice(roots, lanes).astype(np.int32),
        np.full(lanes, 7, np.int32),
        np.concatenate([rng.choice(roots, lanes // 2), rng.choice(mid, lanes // 2)]).astype(np.int32),
    ]
    for srcs in batches:
        odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
        for p in POLICIES:
            dist, sums = g.mssp(srcs, policy=p)
            assert np.array_equal(dist, odist), p
            assert np.array_equal(sums, osums), p


def test_closeness_sums_only_wide_batches(orc):
    """Closeness computes only the sums (no distance rows): two 1024-source
    batches in flight, then an MSSP with rows on the same graph, then a
    weighted graph's closeness; all equal the oracle."""
    log("case: RMAT scale 15 ef 8 seed 23, 1100 sources, closeness then mssp, lane_words=16")
    h = dawn.HostGraph.rmat(15, 8, 23)
    g = dawn.Graph(h)
    g.set_knobs(lane_words=16)
    pool = dawn.gen_sources(h.n, 6000, 23)
    srcs = pool[h.degrees()[pool] > 0][:1100].copy()  # two batches of lanes
    srcs[7] = int(np.flatnonzero(h.degrees() == 0)[0])  # an isolated source takes none
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    cc, sums, ref_int, st = g.closeness(srcs, policy="learned", stats=True)
    assert np.array_equal(sums, osums)
    assert st["batches"] == 2 and st["skipped_sources"] == 1
    for i in (0, 1, 555, 1099):
        c, s, r = orc.closeness(odist[i])
        assert cc[i] == c and ref_int[i] == r
    dist, sums2 = g.mssp(srcs)
    assert np.array_equal(dist, odist) and np.array_equal(sums2, osums)
    hw = dawn.HostGraph.er(2000, 5.0, 24, weighted=True)
    gw = dawn.Graph(hw)
    ws = np.array([0, 7, 1999], np.int32)
    wd, wsum, _, _ = orc.mssp(hw.row_ptr, hw.col, hw.weight, ws)
    _, wsums, _ = gw.closeness(ws)
    assert np.array_equal(wsums, wsum)


def test_all_vertex_closeness_exact(orc):
    """Every vertex a source (isolated ones included), sums only: closeness,
    int64 sums and the reference's integer-division value equal the
    oracle's for every vertex."""
    log("case: ER n=3000 deg 3 seed 25, closeness of all 3000 vertices, lane_words=16")
    h = dawn.HostGraph.er(3000, 3.0, 25)
    g = dawn.Graph(h)
    g.set_knobs(lane_words=16)
    everyone = np.arange(h.n, dtype=np.int32)
    cc, sums, ref_int = g.closeness(everyone, policy="learned")
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, everyone)
    assert np.array_equal(sums, osums)
    assert (h.degrees() == 0).any()
    for v in range(h.n):
        assert (cc[v], sums[v], ref_int[v]) == orc.closeness(odist[v]), v


@pytest.mark.parametrize("heavy,chunk", [(48, 1024), (48, 1000), (32, 512)])
def test_heavy_chunks_multi_round_exact(orc, heavy, chunk):
    """Heavy in-edge chunks longer and shorter than one round of prefetched
    probes (512 edges at 1024 sources per batch), ragged chunk tails and
    vertices split over several chunks: rows and sums equal the oracle's."""
    log(f"case: RMAT scale 16 ef 16 seed 31, 1024 sources, heavy_degree={heavy}, pull_chunk={chunk}")
    h = dawn.HostGraph.rmat(16, 16, 31)
    g = dawn.Graph(h)
    g.set_knobs(lane_words=16, heavy_degree=heavy, pull_chunk=chunk)
    srcs = dawn.gen_sources(h.n, 1024, 31)
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, None, srcs)
    for p in ("pull", "learned"):
        dist, sums = g.mssp(srcs, policy=p)
        assert np.array_equal(dist, odist), p
        assert np.array_equal(sums, osums), p


@pytest.mark.parametrize("weighted", [False, True])
def test_apsp_block_rows_equal_oracle(orc, weighted):
    """dawn_apsp_block (APSP_GPU::run, src/algorithm/gpu/apsp.cu:9-111):
    the rows of sources [first, first+count) equal the oracle's, for the
    whole vertex range in one call and for blocks that start and end inside
    a batch; isolated vertices get INF rows with 0 at the source."""
    log(f"case: APSP blocks, weighted={weighted}")
    if weighted:
        h = dawn.HostGraph.grid(36, 0.02, 41, weighted=True)  # 1296 vertices
    else:
        h = dawn.HostGraph.rmat(11, 8, 41)  # 2048 vertices, isolated ones included
    g = dawn.Graph(h)
    everyone = np.arange(h.n, dtype=np.int32)
    odist, osums, _, _ = orc.mssp(h.row_ptr, h.col, h.weight, everyone)
    dist, sums = g.apsp_block(0, h.n)
    assert np.array_equal(dist, odist)
    assert np.array_equal(sums, osums)
    for first, count in ((5, 70), (1000, 77), (h.n - 3, 3)):
        dist, sums = g.apsp_block(first, count, policy="learned")
        assert np.array_equal(dist, odist[first:first + count]), (first, count)
        assert np.array_equal(sums, osums[first:first + count])
    with pytest.raises(dawn.DawnError):
        g.apsp_block(h.n - 1, 2)


def test_minplus_sums_beyond_int32_are_unreachable(orc):
    """Paths whose weight reaches INT32_MAX are not representable: the
    single-source, batched and step-level min-plus engines all treat them as
    absent (no wrap to negative distances), as the oracle does.  Path
    0-1-2-3-4 with weights 1.5e9, 6e8, 1, 1.5e9 plus a side branch 0-5 of
    2.1e9."""
    u = np.array([0, 1, 2, 3, 0], np.int32)
    v = np.array([1, 2, 3, 4, 5], np.int32)
    w = np.array([1500000000, 600000000, 1, 1500000000, 2100000000], np.int32)
    h = dawn.HostGraph.from_edges(6, u, v, w)
    g = dawn.Graph(h)
    want = {s: orc.dijkstra(h.row_ptr, h.col, h.weight, s, overflow_ok=True) for s in range(6)}
    assert list(want[0]) == [0, 1500000000, 2100000000, 2100000001, INF, 2100000000]
    for s in range(6):
        assert orc.certify(h.row_ptr, h.col, h.weight, s, want[s])
        for p in ("fixed", "push", "pull", "bitmap"):
            assert np.array_equal(g.sssp(s, policy=p), want[s]), (s, p)
    srcs = np.arange(6, dtype=np.int32)
    dist, sums = g.mssp(srcs)
    assert np.array_equal(dist, np.stack([want[s] for s in range(6)]))
    assert list(sums) == [int(want[s][want[s] != INF].astype(np.int64).sum()) for s in range(6)]
    _, csums, _ = g.closeness(srcs)
    assert np.array_equal(csums, sums)


def test_zero_weight_edges_all_engines(orc):
    """Zero-weight edges (allowed: weights are non-negative): single-source
    min-plus under every mode and the batched engine under push, pull and
    the rule agree with Dijkstra."""
    rng = np.random.default_rng(23)
    n = 4000
    u = rng.integers(0, n, 24000).astype(np.int32)
    v = rng.integers(0, n, 24000).astype(np.int32)
    w = rng.integers(0, 5, 24000).astype(np.int32)  # many zeros
    h = dawn.HostGraph.from_edges(n, u, v, w, symmetrize=True)
    assert (h.weight == 0).sum() > 1000
    g = dawn.Graph(h)
    srcs = [int(s) for s in dawn.gen_sources(n, 12, 23) if h.degrees()[s] > 0]
    want = np.stack([orc.dijkstra(h.row_ptr, h.col, h.weight, s) for s in srcs])
    for i, s in enumerate(srcs[:4]):
        for p in ["fixed", "push", "pull", "bitmap"]:
            assert np.array_equal(g.sssp(s, policy=p), want[i]), (s, p)
    for p in ["fixed", "push", "pull"]:
        dist, _ = g.mssp(np.array(srcs, np.int32), policy=p)
        assert np.array_equal(dist, want), p
