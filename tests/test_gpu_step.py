"""Step-level entry point (dawn_step) driven from Python on device tensors:
a BFS and a min-plus SSSP built from single steps must equal the oracle,
with push and pull steps mixed."""
import numpy as np
import pytest

from conftest import dawn

pytestmark = pytest.mark.gpu
INF = dawn.INF


def log(msg):
    print(msg, flush=True)


def run_steps(g, n, source, weighted, modes):
    import torch
    dev = "cuda"
    words = (n + 31) // 32
    dist = torch.full((n,), INF, dtype=torch.int32, device=dev)
    dist[source] = 0
    alpha = torch.zeros(words, dtype=torch.int32, device=dev)
    beta = torch.zeros(words, dtype=torch.int32, device=dev)
    bit = np.array([1 << (source & 31)], np.uint32).view(np.int32)[0]
    alpha[source >> 5] = int(bit)
    changed = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    level, steps = 1, 0
    while True:
        beta.zero_()
        g.step(alpha, beta, dist, changed, level=level, mode=modes[steps % len(modes)], stream=stream)
        # alpha is cleared by the step, as the reference's SOVMP clears it.
        assert int(alpha.abs().sum().item()) == 0
        steps += 1
        if int(changed.item()) == 0:
            break
        alpha, beta = beta, alpha
        level += 1
        assert steps <= n + 1
    return dist.cpu().numpy(), steps


@pytest.mark.parametrize("modes", [("push",), ("pull",), ("push", "pull")])
def test_step_bfs_equals_oracle(orc, modes):
    log(f"case: RMAT scale 12 ef 8 seed 7, BFS by dawn_step, modes {modes}")
    h = dawn.HostGraph.rmat(12, 8, 7)
    g = dawn.Graph(h)
    deg = np.diff(h.row_ptr)
    for s in [v for v in (0, 5, 333, 1000, 2047) if deg[v] > 0][:3]:
        d, steps = run_steps(g, h.n, s, False, modes)
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
        d, _ = run_steps(g, h.n, s, True, modes)
        assert np.array_equal(d, orc.dijkstra(h.row_ptr, h.col, h.weight, s)), s


def test_step_rejects_host_pointers():
    import torch
    h = dawn.HostGraph.er(256, 4.0, 3)
    g = dawn.Graph(h)
    host = torch.zeros(8, dtype=torch.int32)
    dev = torch.zeros(8, dtype=torch.int32, device="cuda")
    dist = torch.zeros(256, dtype=torch.int32, device="cuda")
    ch = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(dawn.DawnError):
        g.step(host, dev, dist, ch, level=1)
