"""dawn_bfs_noalloc, the allocation-free variant of dawn_bfs, through the C
ABI: equal to dawn_bfs (distances, steps, modes, reached counts) on calls
in a row on graphs of different sizes, the later ones smaller, with host
and device outputs; graphs above the scratch region's 4,194,304 vertices
are rejected with DAWN_ERR_SCRATCH_LIMIT."""
import numpy as np
import pytest

from conftest import dawn

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def check_same(g, s, policy, orc_row):
    want, wst = g.bfs(s, policy=policy, stats=True)
    got, gst = g.bfs_noalloc(s, policy=policy, stats=True)
    assert np.array_equal(want, orc_row), (s, policy)
    assert np.array_equal(got, want), (s, policy)
    assert gst["vertices_reached"] == wst["vertices_reached"]
    assert gst["edges_reached"] == wst["edges_reached"]
    if policy != "learned":  # the learned policy may fall back differently
        assert gst["steps"] == wst["steps"]
        assert gst["mode_steps"] == wst["mode_steps"]


def test_noalloc_equals_bfs_calls_in_a_row(orc):
    log("case: dawn_bfs_noalloc on RMAT scale 16 ef 16 seed 51, then ER n=3000 seed 52, then RMAT again")
    big = dawn.HostGraph.rmat(16, 16, 51)
    small = dawn.HostGraph.er(3000, 5.0, 52)
    gb, gs = dawn.Graph(big), dawn.Graph(small)
    for h, g in ((big, gb), (small, gs), (big, gb)):
        deg = h.degrees()
        for s in (int(np.argmax(deg)), int(np.flatnonzero(deg > 0)[-1])):
            ref = orc.bfs(h.row_ptr, h.col, s)
            for policy in ("fixed", "push", "pull", "learned"):
                check_same(g, s, policy, ref)
    # Device output buffer.
    s = int(np.argmax(small.degrees()))
    out = dawn.DeviceArray(small.n, np.int32)
    gs.bfs_noalloc(s, out=out)
    assert np.array_equal(out.download(), orc.bfs(small.row_ptr, small.col, s))
    out.close()


def test_noalloc_largest_graph_and_limit(orc):
    log("case: dawn_bfs_noalloc on a 2048x2048 grid (the maximum size), then 2049x2049 (rejected)")
    h = dawn.HostGraph.grid(2048, 0.01, 53)
    assert h.n == 1 << 22
    g = dawn.Graph(h)
    s = 12345
    check_same(g, s, "fixed", orc.bfs(h.row_ptr, h.col, s))
    g.close()
    hx = dawn.HostGraph.grid(2049, 0.0, 54)
    gx = dawn.Graph(hx)
    with pytest.raises(dawn.DawnError) as e:
        gx.bfs_noalloc(0)
    assert e.value.status == dawn.ERR_SCRATCH_LIMIT
    # The ordinary entry point still takes it.
    d = gx.bfs(0)
    assert d[hx.n - 1] == 2 * 2048
    # And a small graph afterwards still works on the same region.
    hs = dawn.HostGraph.er(500, 4.0, 55)
    gs = dawn.Graph(hs)
    s = int(np.argmax(hs.degrees()))
    check_same(gs, s, "fixed", orc.bfs(hs.row_ptr, hs.col, s))
