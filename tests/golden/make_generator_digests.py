"""Writes generator_digests.json: sha256 prefixes of the seeded generators'
CSR arrays (host build), so any change to the generators' output is caught by
tests/test_host.py::test_golden_generator_digests, and the device build
(tests/test_gpu_build.py) is held to the same arrays at the benchmark sizes
(RMAT scale 20 seed 3, the weighted 2048^2 grid seed 2, RMAT scale 24 seed
99).  Run from the repo root (the scale-24 graph takes about a minute)."""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

dawn = __graft_entry__.load_package()


def digest(h):
    m = hashlib.sha256()
    m.update(h.row_ptr.tobytes())
    m.update(h.col.tobytes())
    if h.weighted:
        m.update(h.weight.tobytes())
    return m.hexdigest()[:16]


got = {
    "er_n4096_d8_s1": digest(dawn.HostGraph.er(4096, 8.0, 1)),
    "rmat_s12_ef16_s3": digest(dawn.HostGraph.rmat(12, 16, 3)),
    "grid_64_1pct_s2_w": digest(dawn.HostGraph.grid(64, 0.01, 2, weighted=True)),
    "sources_n65536_64_s1": hashlib.sha256(dawn.gen_sources(65536, 64, 1).tobytes()).hexdigest()[:16],
    "rmat_s20_ef16_s3": digest(dawn.HostGraph.rmat(20, 16, 3)),
    "grid_2048_1pct_s2_w": digest(dawn.HostGraph.grid(2048, 0.01, 2, weighted=True)),
    "rmat_s24_ef16_s99": digest(dawn.HostGraph.rmat(24, 16, 99)),
}
with open(os.path.join(ROOT, "tests", "golden", "generator_digests.json"), "w") as f:
    json.dump(got, f, indent=1)
print(got)
