"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against something it does not restate:
closed forms, brute force (Floyd-Warshall), scipy's csgraph routines, the
textbook Dijkstra, the reference's own CPU code built from its sources
(oracle/_ref) and hand-derived GAPBS schedules.  A test at the end plants a
mistake in a copy of the oracle's output and checks that the pins catch it.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

from conftest import dawn

INF = dawn.INF


def csr_of(n, edges, weights=None, symmetric=True):
    u = np.array([e[0] for e in edges], np.int32)
    v = np.array([e[1] for e in edges], np.int32)
    w = None if weights is None else np.array(weights, np.int32)
    return dawn.HostGraph.from_edges(n, u, v, w, symmetrize=symmetric, dedup=True)


def floyd(n, rp, col, w=None):
    d = np.full((n, n), np.inf)
    np.fill_diagonal(d, 0)
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            d[r, col[k]] = min(d[r, col[k]], 1 if w is None else w[k])
    for k in range(n):
        d = np.minimum(d, d[:, k:k + 1] + d[k:k + 1, :])
    out = np.where(np.isinf(d), INF, d).astype(np.int64)
    return out


def scipy_dist(h, source, weighted):
    n = h.n
    data = h.weight.astype(np.float64) if weighted else np.ones(h.m)
    A = sp.csr_matrix((data, h.col.copy(), h.row_ptr.copy()), shape=(n, n))
    d = csg.shortest_path(A, directed=True, indices=source, unweighted=not weighted,
                          method="D")
    return np.where(np.isinf(d), INF, d).astype(np.int64)


# ---- closed forms ------------------------------------------------------------

def test_bfs_path_closed_form(orc):
    n = 50
    h = csr_of(n, [(i, i + 1) for i in range(n - 1)])
    for s in (0, 17, n - 1):
        for dense in (True, False):
            d = orc.bfs(h.row_ptr, h.col, s, dense=dense)
            assert np.array_equal(d, np.abs(np.arange(n) - s))


def test_bfs_cycle_star_complete(orc):
    n = 31
    h = csr_of(n, [(i, (i + 1) % n) for i in range(n)])
    d = orc.bfs(h.row_ptr, h.col, 3)
    k = np.abs(np.arange(n) - 3)
    assert np.array_equal(d, np.minimum(k, n - k))
    star = csr_of(20, [(0, i) for i in range(1, 20)])
    d = orc.bfs(star.row_ptr, star.col, 5)
    exp = np.full(20, 2)
    exp[0], exp[5] = 1, 0
    assert np.array_equal(d, exp)
    comp = csr_of(12, [(i, j) for i in range(12) for j in range(i + 1, 12)])
    d = orc.bfs(comp.row_ptr, comp.col, 7)
    assert np.array_equal(d, np.where(np.arange(12) == 7, 0, 1))


def test_bfs_two_components_unreachable(orc):
    h = csr_of(10, [(0, 1), (1, 2), (5, 6), (6, 7)])
    d = orc.bfs(h.row_ptr, h.col, 0)
    assert list(d[:3]) == [0, 1, 2]
    assert all(d[3:] == INF)
    dref = orc.bfs(h.row_ptr, h.col, 0, encoding="ref")
    assert all(dref[3:] == 0) and dref[0] == 0  # reference encoding: 0 = unvisited


def test_sssp_unit_grid_is_manhattan(orc):
    side = 9
    edges = []
    for r in range(side):
        for c in range(side):
            v = r * side + c
            if c + 1 < side:
                edges.append((v, v + 1))
            if r + 1 < side:
                edges.append((v, v + side))
    h = csr_of(side * side, edges, weights=[1] * len(edges))
    d = orc.sssp(h.row_ptr, h.col, h.weight, 0)
    rr, cc = np.divmod(np.arange(side * side), side)
    assert np.array_equal(d, rr + cc)


def test_isolated_source_rejected(orc):
    h = csr_of(5, [(0, 1)])
    with pytest.raises(ValueError):
        orc.bfs(h.row_ptr, h.col, 3)


# ---- brute force and library routines ---------------------------------------

@pytest.mark.parametrize("seed", range(6))
def test_bfs_and_sssp_vs_floyd_warshall(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(8, 40))
    m = int(rng.integers(n, 3 * n))
    directed = bool(seed % 2)
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    w = rng.integers(1, 50, m).astype(np.int32)
    h = dawn.HostGraph.from_edges(n, u, v, w, symmetrize=not directed)
    hu = dawn.HostGraph.from_edges(n, u, v, None, symmetrize=not directed)
    fw_w = floyd(n, h.row_ptr, h.col, h.weight)
    fw_u = floyd(n, hu.row_ptr, hu.col)
    deg = np.diff(h.row_ptr)
    for s in range(n):
        if deg[s] == 0:
            continue
        for dense in (True, False):
            assert np.array_equal(orc.bfs(hu.row_ptr, hu.col, s, dense=dense), fw_u[s])
            assert np.array_equal(orc.sssp(h.row_ptr, h.col, h.weight, s, dense=dense), fw_w[s])
        assert np.array_equal(orc.dijkstra(h.row_ptr, h.col, h.weight, s), fw_w[s])


def test_vs_scipy_on_er_and_grid(orc):
    h = dawn.HostGraph.er(3000, 6.0, 11, weighted=True)
    for s in (0, 77, 2999):
        if h.degrees()[s] == 0:
            continue
        assert np.array_equal(orc.sssp(h.row_ptr, h.col, h.weight, s), scipy_dist(h, s, True))
        assert np.array_equal(orc.bfs(h.row_ptr, h.col, s), scipy_dist(h, s, False))
    g = dawn.HostGraph.grid(40, 0.01, 2, weighted=True)
    assert np.array_equal(orc.sssp(g.row_ptr, g.col, g.weight, 9, dense=False), scipy_dist(g, 9, True))


def test_mssp_matches_single_source_and_skips_isolated(orc):
    h = dawn.HostGraph.from_edges(40, np.arange(0, 30, dtype=np.int32),
                                  np.arange(1, 31, dtype=np.int32))
    src = np.array([0, 35, 10, 30 + 40 * 3, -1], np.int32)  # 35 isolated, mod n, negative
    dist, sums, skipped, _ = orc.mssp(h.row_ptr, h.col, None, src)
    assert list(skipped) == [0, 1, 0, 0, 1]
    assert np.array_equal(dist[0], orc.bfs(h.row_ptr, h.col, 0))
    assert dist[1][35] == 0 and (np.delete(dist[1], 35) == INF).all()
    assert np.array_equal(dist[3], orc.bfs(h.row_ptr, h.col, 30))
    for i in range(len(src)):
        r = dist[i]
        assert sums[i] == r[r != INF].sum()


def test_closeness_closed_form(orc):
    n = 21
    h = csr_of(n, [(i, i + 1) for i in range(n - 1)])
    d = orc.bfs(h.row_ptr, h.col, 0)
    c, s, ref_int = orc.closeness(d)
    assert s == n * (n - 1) // 2
    assert c == (n - 1) / s
    assert ref_int == float((n - 1) // s)  # integer division => 0 here


# ---- GAPBS direction-optimizing schedule (test/gabps/bfs.cc:143-176) ---------

def test_dobfs_hand_derived_schedules(orc):
    star = csr_of(100, [(0, i) for i in range(1, 100)])  # m = 198
    # Centre: scout 99 > 198/15 -> bottom-up; awake 99 >= 1 keeps it; then 0.
    assert list(orc.dobfs_modes(star.row_ptr, star.col, 0)) == [1, 1]
    # Leaf: scout 1 <= 13 -> top-down (e2c 197); centre's scout 99 > 13 ->
    # bottom-up twice (awake 98 >= 1, then 0 < 98 and 0 <= 100/18).
    assert list(orc.dobfs_modes(star.row_ptr, star.col, 7)) == [0, 1, 1]
    path = csr_of(100, [(i, i + 1) for i in range(99)])  # m = 198
    modes = orc.dobfs_modes(path.row_ptr, path.col, 0)
    # Top-down while scout (2 per interior vertex) <= e2c/15: e2c = 198-1-2k.
    # First k with 2 > (197-2k)//15 is k = 84 -> step 85 goes bottom-up.
    assert len(modes) == 100
    assert (modes[:85] == 0).all() and modes[85] == 1


# ---- the reference's own CPU code --------------------------------------------

ref_only = pytest.mark.skipif(not __import__("oracle").ref_available(),
                              reason="oracle/_ref not built (reference sources absent)")


@ref_only
@pytest.mark.parametrize("seed", range(3))
def test_oracle_matches_reference_build_bfs(orc, seed):
    h = dawn.HostGraph.rmat(10, 8, 100 + seed)
    deg = h.degrees()
    srcs = [s for s in np.random.default_rng(seed).integers(0, h.n, 12) if deg[s] > 0]
    for s in srcs:
        for sparse in (False, True):
            r = orc.ref_bfs(h.row_ptr, h.col, s, sparse=sparse)
            assert np.array_equal(orc.bfs(h.row_ptr, h.col, s, encoding="ref"), r)


@ref_only
def test_oracle_matches_reference_build_sssp_and_closeness(orc):
    h = dawn.HostGraph.grid(24, 0.02, 5, weighted=True)
    for s in (0, 100, 575):
        r = orc.ref_sssp(h.row_ptr, h.col, h.weight, s)  # float, exact below 2^24
        o = orc.sssp(h.row_ptr, h.col, h.weight, s)
        assert np.array_equal(o.astype(np.float32), r)
    hu = dawn.HostGraph.er(500, 3.0, 9)
    for s in range(0, 500, 50):
        if hu.degrees()[s] == 0:
            continue
        d = orc.bfs(hu.row_ptr, hu.col, s)
        _, _, ref_int = orc.closeness(d)
        assert ref_int == orc.ref_closeness(hu.row_ptr, hu.col, None, s)


@ref_only
def test_oracle_multigraph_seeding_matches_reference_build():
    """A multigraph whose sources list duplicate out-edges with different
    weights: the reference seeds dist[v] with each listed weight in row
    order (the last listed wins; rows are sorted by (column, weight), so the
    larger), never relaxes the source, and relaxes from there
    (src/algorithm/cpu/sssp.cpp:40-46).  The oracle gives the reference
    build's distances (its unreached sentinel, about 3.4e38, <-> INT32_MAX)."""
    import oracle as orc_mod
    rng = np.random.default_rng(21)
    n = 240
    u = rng.integers(0, 200, 700).astype(np.int32)  # 200..239 unreachable
    v = rng.integers(0, 200, 700).astype(np.int32)
    w = rng.integers(1, 60, 700).astype(np.int32)
    src = np.array([0, 3, 7], np.int32)
    # Every edge at the sources repeated three times with other weights.
    at = np.isin(u, src)
    k = int(at.sum())
    u = np.concatenate([u] + [u[at]] * 3)
    v = np.concatenate([v] + [v[at]] * 3)
    w = np.concatenate([w] + [rng.integers(1, 60, k).astype(np.int32) for _ in range(3)])
    h = dawn.HostGraph.from_edges(n, u, v, w, symmetrize=True, dedup=False)
    dup = differs = 0
    for s in src:
        seg = slice(h.row_ptr[s], h.row_ptr[s + 1])
        cols = h.col[seg]
        dup += len(cols) - len(np.unique(cols))
        r = orc_mod.ref_sssp(h.row_ptr, h.col, h.weight, int(s))
        o = orc_mod.sssp(h.row_ptr, h.col, h.weight, int(s))
        reached = o != INF
        assert np.array_equal(reached, r < 3e38)  # the reference's large sentinel
        assert np.array_equal(o[reached].astype(np.float32), r[reached])
        differs += int((o != orc_mod.dijkstra(h.row_ptr, h.col, h.weight, int(s))).sum())
    assert dup > 10  # the case is exercised
    # ... and tells the reference's seeding from true shortest paths.
    assert differs > 0

def test_mlp_oracle_gradient_finite_differences(orc):
    rng = np.random.default_rng(0)
    p = dawn.mlp_init(3).astype(np.float64) + rng.normal(0, 0.05, dawn.SCHED_PARAMS)
    x = rng.normal(0, 1, (16, 8))
    y = rng.normal(0, 1, (16, 3))
    loss, g = orc.mlp_loss_grad(p, x, y)
    eps = 1e-6
    idx = rng.choice(dawn.SCHED_PARAMS, 60, replace=False)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        num = (orc.mlp_loss_grad(pp, x, y)[0] - orc.mlp_loss_grad(pm, x, y)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])
    out = orc.mlp_forward(p, x)
    assert abs(((out - y) ** 2).mean() - loss) < 1e-12


# ---- planted mistakes are caught ---------------------------------------------

def test_planted_oracle_mistakes_fail_pins(orc):
    h = dawn.HostGraph.er(400, 5.0, 4, weighted=True)
    s = int(np.argmax(h.degrees()))
    good = orc.sssp(h.row_ptr, h.col, h.weight, s)
    ref = scipy_dist(h, s, True)
    assert np.array_equal(good, ref)
    reached = np.flatnonzero((good != INF) & (good > 0))
    bad = good.copy()
    bad[reached[0]] += 1  # an off-by-one distance
    assert not np.array_equal(bad, ref)
    bad2 = good.copy()
    bad2[reached[1]] = INF  # a dropped vertex
    assert not np.array_equal(bad2, ref)


# ---- shortest-path certificate -------------------------------------------------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_certify_accepts_distances_and_rejects_planted_mistakes(orc, seed):
    """The certificate accepts scipy's distances (unit and positive int
    weights, undirected and directed) and rejects every planted mistake: one
    entry off by one either way, an unreached vertex marked reached, a
    reached one marked unreached, a wrong source entry, and a consistent
    shift of a whole subtree (a valid-looking potential that is not tight)."""
    rng = np.random.default_rng(seed)
    n = 400
    u = rng.integers(0, n, 1500).astype(np.int32)
    v = rng.integers(0, n, 1500).astype(np.int32)
    w = rng.integers(1, 60, 1500).astype(np.int32)
    for sym in (True, False):
        for weighted in (False, True):
            h = dawn.HostGraph.from_edges(n, u, v, w if weighted else None, symmetrize=sym)
            s = int(np.argmax(np.diff(h.row_ptr)))
            d = scipy_dist(h, s, weighted).astype(np.int32)
            ww = h.weight if weighted else None
            assert orc.certify(h.row_ptr, h.col, ww, s, d)
            reached = np.flatnonzero((d != INF) & (np.arange(n) != s))
            unreached = np.flatnonzero(d == INF)
            x = int(reached[len(reached) // 2])
            for delta in (-1, 1):
                bad = d.copy()
                bad[x] += delta
                assert not orc.certify(h.row_ptr, h.col, ww, s, bad), (sym, weighted, delta)
            bad = d.copy()
            bad[x] = INF
            assert not orc.certify(h.row_ptr, h.col, ww, s, bad)
            bad = d.copy()
            bad[s] = 1
            assert not orc.certify(h.row_ptr, h.col, ww, s, bad)
            if len(unreached):
                bad = d.copy()
                bad[unreached[0]] = int(d[reached].max()) + 1
                assert not orc.certify(h.row_ptr, h.col, ww, s, bad)
            # Every reached vertex other than the source shifted up by the
            # same amount: edge inequalities still hold, tightness fails at
            # the source's neighbours.
            bad = d.copy()
            bad[reached] += 1
            assert not orc.certify(h.row_ptr, h.col, ww, s, bad)


def test_certify_path_sums_beyond_int32_are_absent(orc):
    """Path 0-1-2-3 with weights 1.5e9: vertex 2 (3e9) and 3 are unreachable
    in int32, so the certificate takes [0, 1.5e9, INF, INF] and rejects a
    wrapped negative value."""
    h = dawn.HostGraph.from_edges(4, np.array([0, 1, 2], np.int32), np.array([1, 2, 3], np.int32),
                                  np.array([1500000000] * 3, np.int32))
    good = np.array([0, 1500000000, INF, INF], np.int32)
    assert orc.certify(h.row_ptr, h.col, h.weight, 0, good)
    wrapped = good.copy()
    wrapped[2] = np.int32(-1294967296)
    assert not orc.certify(h.row_ptr, h.col, h.weight, 0, wrapped)
    for row in (orc.dijkstra(h.row_ptr, h.col, h.weight, 0, overflow_ok=True),
                orc.sssp(h.row_ptr, h.col, h.weight, 0, overflow_ok=True),
                orc.sssp(h.row_ptr, h.col, h.weight, 0, dense=False, overflow_ok=True),
                orc.mssp(h.row_ptr, h.col, h.weight, [0], overflow_ok=True)[0][0]):
        assert np.array_equal(row, good)
    with pytest.raises(ValueError):
        orc.sssp(h.row_ptr, h.col, h.weight, 0)


def test_weighted_closeness_restatement_matches_reference_build(orc):
    """The reference's weighted closeness value (float accumulate of the row,
    FLT_MAX for unreached vertices, cc.cpp:132-133) restated from an int32
    row equals DAWN::CC_CPU::kernel_Weighted of the reference build: on a
    connected weighted grid, and on graphs with one and with several
    unreachable vertices (about (n-1)/FLT_MAX, then 0)."""
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built")
    cases = [dawn.HostGraph.grid(20, 0.05, 81, weighted=True)]
    # One unreachable vertex (isolated 39), then three (37..39).
    for iso in (1, 3):
        n = 40
        u = np.arange(n - iso - 1, dtype=np.int32)
        cases.append(dawn.HostGraph.from_edges(n, u, u + 1, (u % 7 + 1).astype(np.int32)))
    for h in cases:
        for s in (0, 5):
            row = orc.sssp(h.row_ptr, h.col, h.weight, s)
            want = orc.ref_closeness(h.row_ptr, h.col, h.weight, s)
            assert orc.weighted_closeness_ref(row) == want, (h.n, s)
    assert orc.weighted_closeness_ref(orc.sssp(cases[2].row_ptr, cases[2].col, cases[2].weight, 0)) == 0
