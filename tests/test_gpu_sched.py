"""Scheduler MLP and policies on the GPU.

- forward pass and loss gradient of the HIP MLP kernels vs the CPU
  double-precision oracle (tolerances in the asserts);
- training lowers the loss;
- the decide kernel's on-device inference picks the same mode as the host
  argmin over the HIP forward pass on the same features;
- every policy yields identical distances;
- a network trained on recorded step timings of small graphs is compared
  with the fixed alpha/beta rule on held-out recorded steps (same state,
  same timings) and must not lose to it.
"""
import numpy as np
import pytest

from conftest import dawn

pytestmark = pytest.mark.gpu


def rand_net(seed):
    rng = np.random.default_rng(seed)
    return (dawn.mlp_init(seed) + rng.normal(0, 0.05, dawn.SCHED_PARAMS)).astype(np.float32)


def test_mlp_forward_vs_oracle(orc):
    rng = np.random.default_rng(1)
    p = rand_net(1)
    x = rng.normal(0, 1, (513, 8)).astype(np.float32)
    got = dawn.mlp_forward(p, x)
    exp = orc.mlp_forward(p.astype(np.float64), x.astype(np.float64))
    np.testing.assert_allclose(got, exp, rtol=1e-4, atol=1e-5)


def test_mlp_grad_vs_oracle(orc):
    rng = np.random.default_rng(2)
    p = rand_net(2)
    x = rng.normal(0, 1, (300, 8)).astype(np.float32)
    y = rng.normal(0, 1, (300, 3)).astype(np.float32)
    loss, g = dawn.mlp_loss_grad(p, x, y)
    eloss, eg = orc.mlp_loss_grad(p.astype(np.float64), x.astype(np.float64), y.astype(np.float64))
    assert loss == pytest.approx(eloss, rel=1e-4)
    np.testing.assert_allclose(g, eg, rtol=1e-3, atol=2e-5)


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(3)
    x = rng.normal(0, 1, (2000, 8)).astype(np.float32)
    w = rng.normal(0, 1, (8, 3))
    y = (np.tanh(x @ w) + 0.1 * x[:, :3]).astype(np.float32)
    p0 = dawn.mlp_init(4)
    l0, _ = dawn.mlp_loss_grad(p0, x, y)
    p1, l1 = dawn.mlp_train(p0, x, y, epochs=60, batch=128, lr=3e-3, seed=5)
    assert l1 < 0.3 * l0


@pytest.fixture(scope="module")
def small_model():
    """Records step timings on small training graphs and trains a network."""
    feats, times, fixed = [], [], []
    for kind, size, seed in [("rmat", 14, 10), ("rmat", 15, 11), ("rmat", 16, 12), ("er", 1 << 15, 13),
                             ("grid", 128, 14)]:
        if kind == "rmat":
            h = dawn.HostGraph.rmat(size, 16, seed)
        elif kind == "er":
            h = dawn.HostGraph.er(size, 8.0, seed)
        else:
            h = dawn.HostGraph.grid(size, 0.01, seed)
        g = dawn.Graph(h)
        f, t, fx = g.sched_record(dawn.gen_sources(h.n, 128, seed), repeats=3)
        feats.append(f)
        times.append(t)
        fixed.append(fx)
        g.close()
    X = np.concatenate(feats)
    T = np.concatenate(times)
    Y = np.log2(np.maximum(T, 1e-4) * 1e3).astype(np.float32)
    mean, std = X.mean(0), X.std(0)
    std[std < 1e-6] = 1
    Xn = ((X - mean) / std).astype(np.float32)
    p, loss = dawn.mlp_train(dawn.mlp_init(7), Xn, Y, epochs=300, batch=128, lr=3e-3, seed=7)
    return p, mean.astype(np.float32), (1 / std).astype(np.float32), loss


def test_device_inference_matches_host_forward(small_model, orc):
    p, mean, inv_std, _ = small_model
    h = dawn.HostGraph.rmat(15, 16, 99)
    g = dawn.Graph(h)
    src = dawn.gen_sources(h.n, 128, 99)
    dawn.sched_set_mlp(p, mean, inv_std)
    try:
        modes, feat, _ = g.sched_trace(src, policy="learned")
        host = dawn.mlp_forward(p, (feat - mean) * inv_std).argmin(axis=1)
        assert np.array_equal(modes, host)
        d_l, s_l = g.mssp(src, policy="learned")
        d_f, s_f = g.mssp(src, policy="fixed")
        odist, _, _, _ = orc.mssp(h.row_ptr, h.col, None, src)
        assert np.array_equal(d_l, odist) and np.array_equal(d_f, odist)
    finally:
        dawn.sched_set_mlp(None)


def test_learned_vs_fixed_on_held_out_steps(small_model):
    p, mean, inv_std, _ = small_model
    h = dawn.HostGraph.rmat(16, 16, 98)  # held out: seed 98 is not a training seed
    g = dawn.Graph(h)
    feat, times, fixed = g.sched_record(dawn.gen_sources(h.n, 256, 98), repeats=3)
    learned = dawn.mlp_forward(p, (feat - mean) * inv_std).argmin(axis=1)
    idx = np.arange(len(times))
    best = times.min(axis=1).sum()
    t_fixed = times[idx, fixed].sum()
    t_learned = times[idx, learned].sum()
    print(f"held-out steps: oracle-best {best:.3f} ms, fixed {t_fixed:.3f} ms, "
          f"learned {t_learned:.3f} ms, gain {t_fixed / t_learned:.3f}x", flush=True)
    assert t_learned <= 1.05 * t_fixed
    assert t_learned <= 1.25 * best


@pytest.fixture(scope="module")
def single_model():
    """Single-source steps recorded on small training graphs (every mode
    timed from identical state) and a network trained on them."""
    feats, times, fixed = [], [], []
    for kind, size, seed in [("rmat", 16, 10), ("rmat", 17, 11), ("er", 1 << 17, 12),
                             ("grid", 384, 13), ("grid", 256, 14)]:
        if kind == "rmat":
            h = dawn.HostGraph.rmat(size, 16, seed)
        elif kind == "er":
            h = dawn.HostGraph.er(size, 8.0, seed)
        else:
            h = dawn.HostGraph.grid(size, 0.01, seed)
        g = dawn.Graph(h)
        pool = dawn.gen_sources(h.n, 200, seed)
        f, t, fx = g.sched_record(pool[h.degrees()[pool] > 0][:6], repeats=2, single=True,
                                  max_rows=8192)
        feats.append(f)
        times.append(t)
        fixed.append(fx)
        g.close()
    X = np.concatenate(feats)
    T = np.concatenate(times)
    Y = np.log2(np.maximum(T, 1e-4) * 1e3).astype(np.float32)
    mean, std = X.mean(0), X.std(0)
    std[std < 1e-6] = 1
    Xn = ((X - mean) / std).astype(np.float32)
    p, loss = dawn.mlp_train(dawn.mlp_init(9), Xn, Y, epochs=300, batch=128, lr=3e-3, seed=9)
    return p, mean.astype(np.float32), (1 / std).astype(np.float32)


def test_single_source_record_consistent(orc):
    """Every recorded row belongs to a step of a non-isolated source, the
    fixed rule's recorded choices are those dawn_bfs makes, and the trace
    under a forced policy reports that mode at every step."""
    h = dawn.HostGraph.rmat(15, 16, 31)
    g = dawn.Graph(h)
    deg = h.degrees()
    s = int(np.argmax(deg))
    iso = int(np.flatnonzero(deg == 0)[0])
    feat, times, fixed = g.sched_record([s, iso], repeats=1, single=True)
    _, st = g.bfs(s, policy="fixed", stats=True)
    assert len(times) == st["steps"]
    assert int((fixed == 1).sum()) == st["mode_steps"]["pull"]
    assert int((fixed == 0).sum()) == st["mode_steps"]["push"]
    assert (times > 0).all()
    modes, _, _ = g.sched_trace([s], policy="push", single=True)
    assert (modes == 0).all() and len(modes) == len(times)


def test_single_source_learned_beats_fixed_on_held_out_steps(single_model, orc):
    """On held-out graphs the single-source network's choices cost less than
    the fixed alpha/beta rule's on the same recorded steps, and stay within
    25% of the per-step oracle-best; dawn_bfs under the installed network
    gives the oracle's distances."""
    p, mean, inv_std = single_model
    tot = {"fixed": 0.0, "learned": 0.0, "best": 0.0}
    for kind, size, seed in [("rmat", 17, 98), ("grid", 320, 97)]:
        h = dawn.HostGraph.rmat(size, 16, seed) if kind == "rmat" else dawn.HostGraph.grid(size, 0.01, seed)
        g = dawn.Graph(h)
        pool = dawn.gen_sources(h.n, 200, seed)
        src = pool[h.degrees()[pool] > 0][:6]
        feat, times, fixed = g.sched_record(src, repeats=2, single=True, max_rows=8192)
        learned = dawn.mlp_forward(p, (feat - mean) * inv_std).argmin(axis=1)
        idx = np.arange(len(times))
        tot["fixed"] += times[idx, fixed].sum()
        tot["learned"] += times[idx, learned].sum()
        tot["best"] += times.min(axis=1).sum()
        dawn.sched_set_mlp(p, mean, inv_std, single=True)
        try:
            for s in src[:2]:
                assert np.array_equal(g.bfs(int(s), policy="learned"), orc.bfs(h.row_ptr, h.col, int(s)))
        finally:
            dawn.sched_set_mlp(None, single=True)
        g.close()
    print(f"held-out single-source steps: oracle-best {tot['best']:.3f} ms, fixed {tot['fixed']:.3f} ms, "
          f"learned {tot['learned']:.3f} ms, gain {tot['fixed'] / tot['learned']:.3f}x", flush=True)
    assert tot["learned"] < tot["fixed"]
    assert tot["learned"] <= 1.25 * tot["best"]
