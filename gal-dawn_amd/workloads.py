"""The seeded synthetic workloads the engine is measured and tested on.

er65k         ER n=65,536, avg degree 8, seed 1; source 0 + 64 uniform sources
grid2048      2048x2048 4-neighbour grid + 1% random long-range edges, int32
              weights uniform in [1,1000], seed 2; 16 uniform sources
rmat20        RMAT(.57,.19,.19,.05) scale 20, ef 16, symmetrized, dedup,
              seed 3; 4096 uniform sources (bit-packed batches of 64*W lanes)
uneven512     512 graphs, seed 4: ER (avg degree 4-32), RMAT (scale 10-17),
              grids with 1% long-range edges; vertex counts log-uniform in
              [1k, 130k]; 8 sources each; unweighted or int32-weighted
rmat24        RMAT scale 24, ef 16, symmetrized, dedup, seed 99; 4096 sources
"""
from __future__ import annotations

import math

import numpy as np


def _dawn():
    import sys
    return sys.modules["gal_dawn_amd"]


def er65k():
    d = _dawn()
    h = d.HostGraph.er(65536, 8.0, 1)
    src = np.concatenate([[0], d.gen_sources(h.n, 64, 1)]).astype(np.int32)
    return h, src


def grid2048(side: int = 2048):
    d = _dawn()
    h = d.HostGraph.grid(side, 0.01, 2, weighted=True, wmin=1, wmax=1000)
    return h, d.gen_sources(h.n, 16, 2)


def rmat20(scale: int = 20, sources: int = 4096):
    d = _dawn()
    h = d.HostGraph.rmat(scale, 16, 3)
    return h, d.gen_sources(h.n, sources, 3)


def uneven(count: int = 512, seed: int = 4, weighted: bool = False, spg: int = 8,
           nmin: int = 1000, nmax: int = 130000):
    """Returns (graphs, sources[count*spg], families)."""
    d = _dawn()
    graphs, srcs, fams = [], [], []
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        n_target = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        fam = int(rng.integers(3))
        gseed = seed * 100003 + i
        if fam == 0:
            deg = float(rng.integers(4, 33))
            h = d.HostGraph.er(int(round(n_target)), deg, gseed, weighted=weighted)
            fams.append(f"er{deg:.0f}")
        elif fam == 1:
            scale = int(min(17, max(10, round(math.log2(n_target)))))
            h = d.HostGraph.rmat(scale, 16, gseed, weighted=weighted)
            fams.append(f"rmat{scale}")
        else:
            side = int(min(362, max(64, round(math.sqrt(n_target)))))
            h = d.HostGraph.grid(side, 0.01, gseed, weighted=weighted)
            fams.append(f"grid{side}")
        graphs.append(h)
        srcs.append(d.gen_sources(h.n, spg, gseed))
    return graphs, np.concatenate(srcs).astype(np.int32), fams


def rmat24(scale: int = 24, sources: int = 4096, seed: int = 99):
    d = _dawn()
    h = d.HostGraph.rmat(scale, 16, seed)
    return h, d.gen_sources(h.n, sources, seed)
