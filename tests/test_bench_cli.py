"""bench.py's command-line contract that needs no GPU."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_full_and_tail_split_follows_the_engine_order():
    """The profiling pass's split: sources of out-degree 0 take no lane, the
    others by decreasing out-degree (stable), full batches of 64*W lanes
    first, the partly filled last batch after them."""
    sys.path.insert(0, ROOT)
    import numpy as np

    import bench
    import __graft_entry__

    dawn = __graft_entry__.load_package()
    # A path 0..99 (ends of degree 1, the rest 2) and isolated 100..119.
    u = np.arange(99, dtype=np.int32)
    h = dawn.HostGraph.from_edges(120, u, u + 1)
    src = np.concatenate([[0, 110], np.arange(1, 70), [99, 115]]).astype(np.int32)
    full, tail = bench.full_and_tail(h, src, W=1)
    assert full.tolist() == list(range(1, 65))           # degree 2, caller's order
    assert tail.tolist() == list(range(65, 70)) + [0, 99]  # then the degree-1 ends
    full, tail = bench.full_and_tail(h, src[:10], W=1)  # fewer than 64 lanes: all in the last
    assert full.tolist() == [] and tail.tolist() == list(range(1, 9)) + [0]
