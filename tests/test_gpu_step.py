"""Step-level entry point (dawn_step) driven from Python on device buffers:
a BFS and a min-plus SSSP built from single steps must equal the oracle,
with push and pull steps mixed."""
import numpy as np
import pytest

from conftest import dawn

pytestmark = pytest.mark.gpu
INF = dawn.INF


def log(msg):
    print(msg, flush=True)


def run_steps(g, n, source, modes):
    words = (n + 31) // 32
    d0 = np.full(n, INF, np.int32)
    d0[source] = 0
    dist = dawn.DeviceArray(n, np.int32).upload(d0)
    a0 = np.zeros(words, np.uint32)
    a0[source >> 5] = np.uint32(1) << np.uint32(source & 31)
    alpha = dawn.DeviceArray(words, np.uint32).upload(a0)
    beta = dawn.DeviceArray(words, np.uint32)
    changed = dawn.DeviceArray(1, np.int64)
    level, steps = 1, 0
    while True:
        beta.fill_bytes(0)
        g.step(alpha, beta, dist, changed, level=level, mode=modes[steps % len(modes)])
        g.synchronize()
        # alpha is cleared by the step, as the reference's SOVMP clears it.
        assert not alpha.download().any()
        steps += 1
        if int(changed.download()[0]) == 0:
            break
        alpha, beta = beta, alpha
        level += 1
        assert steps <= n + 1
    return dist.download(), steps


@pytest.mark.parametrize("modes", [("push",), ("pull",), ("push", "pull")])
def test_step_bfs_equals_oracle(orc, modes):
    log(f"case: RMAT scale 12 ef 8 seed 7, BFS by dawn_step, modes {modes}")
    h = dawn.HostGraph.rmat(12, 8, 7)
    g = dawn.Graph(h)
    deg = np.diff(h.row_ptr)
    for s in [v for v in (0, 5, 333, 1000, 2047) if deg[v] > 0][:3]:
        d, steps = run_steps(g, h.n, s, modes)
        ref = orc.bfs(h.row_ptr, h.col, s)
        assert np.array_equal(d, ref), s
        reached = ref[ref != INF]
        assert steps == int(reached.max()) + 1  # one step per level, then one empty


@pytest.mark.parametrize("modes", [("push",), ("pull",), ("pull", "push")])
def test_step_minplus_equals_oracle(orc, modes):
    log(f"case: grid 48 + 1% seed 8 weighted, min-plus by dawn_step, modes {modes}")
    h = dawn.HostGraph.grid(48, 0.01, 8, weighted=True)
    g = dawn.Graph(h)
    for s in (0, 777):
        d, _ = run_steps(g, h.n, s, modes)
        assert np.array_equal(d, orc.dijkstra(h.row_ptr, h.col, h.weight, s)), s


def test_step_rejects_host_pointers():
    h = dawn.HostGraph.er(256, 4.0, 3)
    g = dawn.Graph(h)

    class Host:
        def __init__(self, a):
            self.a = a

        def data_ptr(self):
            return self.a.ctypes.data

    host = Host(np.zeros(8, np.uint32))
    dev = dawn.DeviceArray(8, np.uint32)
    dist = dawn.DeviceArray(256, np.int32)
    ch = dawn.DeviceArray(1, np.int64)
    with pytest.raises(dawn.DawnError):
        g.step(host, dev, dist, ch, level=1)


def test_step_minplus_large_weights_do_not_wrap(orc):
    """Min-plus steps with weights near 2^31: a candidate sum at or above
    INT32_MAX is ignored (the vertex stays unreached), never wrapped."""
    u = np.array([0, 1, 2, 3], np.int32)
    v = np.array([1, 2, 3, 4], np.int32)
    w = np.array([1500000000, 600000000, 1, 1500000000], np.int32)
    h = dawn.HostGraph.from_edges(5, u, v, w)
    g = dawn.Graph(h)
    want = orc.dijkstra(h.row_ptr, h.col, h.weight, 0, overflow_ok=True)
    assert list(want) == [0, 1500000000, 2100000000, 2100000001, INF]
    for modes in (("push",), ("pull",)):
        d, _ = run_steps(g, h.n, 0, modes)
        assert np.array_equal(d, want), modes
