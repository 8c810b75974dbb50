"""The drop-in host functions against the reference's own code: the four
MatrixMarket readers (DAWN::IO::readGraph*, src/io.cpp:2-192), the COO
helpers (DAWN::Matrix::transpose* / coo2Csr*, src/matrix.cpp:3-82) and the
source-less outfile forms (src/io.cpp:252-281).  tests/dropin_dump.cpp is
compiled against include/dawn_amd.hxx + libdawn_amd.so and against the
reference's headers + oracle/_ref/libdawn_ref.so (the reference's sources
compiled where they lie); both dumps must be identical.  CPU only; skipped
where the reference sources or oracle/_ref are absent."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

REF = "/root/reference"
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libdawn_ref.so")
pytestmark = pytest.mark.skipif(not (os.path.isdir(os.path.join(REF, "include")) and
                                     os.path.exists(REF_LIB)),
                                reason="reference sources / oracle/_ref not present")


@pytest.fixture(scope="module")
def progs(tmp_path_factory):
    d = tmp_path_factory.mktemp("dropin")
    src = os.path.join(ROOT, "tests", "dropin_dump.cpp")
    ours = str(d / "ours")
    ref = str(d / "ref")
    pkg = os.path.join(ROOT, "gal-dawn_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", ours,
                    "-L", pkg, "-ldawn_amd", f"-Wl,-rpath,{pkg}", "-Wl,-rpath,/opt/rocm/lib"],
                   check=True)
    subprocess.run(["g++", "-O1", "-std=c++14", "-w", "-DUSE_REF", "-I", os.path.join(REF, "include"), src,
                    "-o", ref, "-L", os.path.dirname(REF_LIB), "-ldawn_ref", "-fopenmp",
                    f"-Wl,-rpath,{os.path.dirname(REF_LIB)}"], check=True)
    return d, ours, ref


def run_both(progs, *args):
    d, ours, ref = progs
    outs = []
    for exe, tag in ((ours, "ours"), (ref, "ref")):
        out = str(d / f"{tag}.txt")
        subprocess.run([exe, args[0], args[1], out, *args[2:]], check=True, capture_output=True,
                       timeout=120)
        outs.append(out)
    return outs


def write_mtx(path, n, edges, field, symmetry):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {symmetry}\n% comment\n")
        f.write(f"{n} {n} {len(edges)}\n")
        for e in edges:
            f.write(" ".join(str(x) for x in e) + "\n")


@pytest.mark.parametrize("field,symmetry", [("pattern", "symmetric"), ("pattern", "general"),
                                            ("integer", "general"), ("real", "symmetric")])
def test_readers_match_reference(progs, field, symmetry):
    d = progs[0]
    rng = np.random.default_rng(71)
    n = 60
    edges = []
    for _ in range(400):
        r, c = int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))
        if field == "pattern":
            edges.append((r, c))
        else:
            edges.append((r, c, int(rng.integers(1, 9))))
    edges += edges[:50]  # duplicates
    edges += [(5, 5)] if field == "pattern" else [(5, 5, 3)]  # self-loop
    path = str(d / f"{field}_{symmetry}.mtx")
    write_mtx(path, n, edges, field, symmetry)
    readers = ["readGraph", "readGraph_Directed"]
    if field != "pattern":
        readers += ["readGraph_Weighted", "readGraph_Directed_Weighted"]
    for rd in readers:
        a, b = run_both(progs, rd, path, str(n), str(len(edges)))
        assert open(a).read() == open(b).read(), rd


def test_coo_helpers_match_reference(progs):
    d = progs[0]
    rng = np.random.default_rng(72)
    n, nnz = 40, 300
    path = str(d / "coo.txt")
    with open(path, "w") as f:
        f.write(f"{n} {nnz}\n")
        for _ in range(nnz):
            f.write(f"{rng.integers(0, n)} {rng.integers(0, n)} {rng.integers(0, 5)}\n")
    a, b = run_both(progs, "coo", path)
    assert open(a).read() == open(b).read()


def test_sourceless_outfile_matches_reference(progs):
    d = progs[0]
    rng = np.random.default_rng(73)
    vals = rng.integers(-2, 40, 200)
    path = str(d / "dist.txt")
    with open(path, "w") as f:
        f.write(f"{len(vals)}\n" + " ".join(str(int(x)) for x in vals) + "\n")
    a, b = run_both(progs, "outfile", path)
    for ext in (".int", ".float"):
        assert open(a + ext).read() == open(b + ext).read(), ext
