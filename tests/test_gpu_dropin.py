"""The drop-in CLI mains (gal-dawn_amd/bin/*_gpu, built on include/dawn_amd.hxx)
against the CPU oracle and, where built, the reference's own CPU code: the
"src dst dist" output files must be identical."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, dawn

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "gal-dawn_amd", "bin")


def run(app, *args):
    r = subprocess.run([os.path.join(BIN, app), *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout, r.stderr)
    return r.stdout


def expected_lines(src, dist):
    return "".join(f"{src} {j} {d}\n" for j, d in enumerate(dist)
                   if j != src and d != dawn.INF and d > 0)


@pytest.fixture(scope="module")
def mtx(tmp_path_factory):
    d = tmp_path_factory.mktemp("mtx")
    h = dawn.HostGraph.rmat(12, 8, 5)
    p = str(d / "g.mtx")
    h.write_mtx(p)
    hw = dawn.HostGraph.grid(40, 0.02, 6, weighted=True)
    pw = str(d / "w.mtx")
    hw.write_mtx(pw)
    return d, p, pw


def test_bfs_gpu_main_output_file(mtx, orc):
    d, p, _ = mtx
    g = dawn.HostGraph.read_mtx(p)
    src = int(np.argmax(g.degrees()))
    out = str(d / "bfs.txt")
    text = run("bfs_gpu", "-i", p, "-o", out, "-p", "true", "-s", str(src))
    assert "Nodes:" in text and "Time:" in text
    got = open(out).read()
    assert got == expected_lines(src, orc.bfs(g.row_ptr, g.col, src))
    if orc.ref_available():
        refd = orc.ref_bfs(g.row_ptr, g.col, src)  # reference encoding
        ref_text = "".join(f"{src} {j} {x}\n" for j, x in enumerate(refd) if j != src and x > 0)
        assert got == ref_text


def test_sssp_gpu_main_output_file(mtx, orc):
    d, _, pw = mtx
    g = dawn.HostGraph.read_mtx(pw, weighted=True)
    out = str(d / "sssp.txt")
    run("sssp_gpu", "-i", pw, "-o", out, "-p", "true", "-s", "17")
    dist = orc.sssp(g.row_ptr, g.col, g.weight, 17)
    lines = open(out).read().splitlines()
    got = {int(l.split()[1]): float(l.split()[2]) for l in lines}
    exp = {j: float(x) for j, x in enumerate(dist) if j != 17 and x != 0}
    exp = {j: (3.40282e38 if dist[j] == dawn.INF else v) for j, v in exp.items()}
    assert got.keys() == exp.keys()
    for j in got:
        assert got[j] == pytest.approx(exp[j], rel=1e-6)


def weighted_lines(path):
    out = {}
    for line in open(path).read().splitlines():
        a, b, c = line.split()
        out[int(b)] = float(c)
    return out


def test_mssp_gpu_main_source_list(mtx, orc, tmp_path):
    """The mssp_gpu main writes the row of graph.source when it is in the
    list (src/algorithm/gpu/bfs.cu:97-102); list entries are taken mod n."""
    _, p, _ = mtx
    g = dawn.HostGraph.read_mtx(p)
    lst = tmp_path / "src.txt"
    lst.write_text("% sources\n" + "\n".join(str(x) for x in [3, 4099, 2147483647, 77]) + "\n")
    src = 4099 % g.n
    assert g.degrees()[src] > 0
    out = str(tmp_path / "mssp.txt")
    text = run("mssp_gpu", "-i", p, "-l", str(lst), "-s", str(src), "-p", "true", "-o", out)
    assert "Time:" in text and "Start print" in text
    assert open(out).read() == expected_lines(src, orc.bfs(g.row_ptr, g.col, src))


def test_mssp_gpu_main_weighted_row(mtx, orc, tmp_path):
    _, _, pw = mtx
    g = dawn.HostGraph.read_mtx(pw, weighted=True)
    lst = tmp_path / "wsrc.txt"
    lst.write_text("5\n17\n900\n")
    out = str(tmp_path / "wmssp.txt")
    run("mssp_gpu", "-i", pw, "-l", str(lst), "-s", "17", "-p", "true", "-o", out, "-w", "true")
    dist = orc.dijkstra(g.row_ptr, g.col, g.weight, 17)
    got = weighted_lines(out)
    exp = {j: float(x) for j, x in enumerate(dist) if j != 17 and x != dawn.INF}
    assert got == pytest.approx(exp, rel=1e-6)


def test_apsp_gpu_main_rows(mtx, orc, tmp_path):
    """The apsp_gpu main runs every vertex as a source and writes the row of
    graph.source (src/algorithm/gpu/apsp.cu:9-60, bfs.cu:97-102)."""
    _, p, pw = mtx
    g = dawn.HostGraph.read_mtx(p)
    src = int(np.argmax(g.degrees()))
    out = str(tmp_path / "apsp.txt")
    text = run("apsp_gpu", "-i", p, "-o", out, "-p", "true", "-s", str(src))
    assert "Time:" in text
    assert open(out).read() == expected_lines(src, orc.bfs(g.row_ptr, g.col, src))
    gw = dawn.HostGraph.read_mtx(pw, weighted=True)
    outw = str(tmp_path / "wapsp.txt")
    run("apsp_gpu", "-i", pw, "-o", outw, "-p", "true", "-s", "33", "-w", "true")
    dist = orc.dijkstra(gw.row_ptr, gw.col, gw.weight, 33)
    exp = {j: float(x) for j, x in enumerate(dist) if j != 33 and x != dawn.INF}
    assert weighted_lines(outw) == pytest.approx(exp, rel=1e-6)


def test_cc_gpu_main_value(mtx, orc):
    _, p, _ = mtx
    g = dawn.HostGraph.read_mtx(p)
    src = int(np.argmax(g.degrees()))
    text = run("cc_gpu", "-i", p, "-s", str(src))
    val = float([l for l in text.splitlines() if l.startswith("Closeness")][0].split(":")[1])
    _, _, ref_int = orc.closeness(orc.bfs(g.row_ptr, g.col, src))
    assert val == pytest.approx(ref_int)


def test_isolated_source_exits_zero(tmp_path):
    p = tmp_path / "iso.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n4 4 1\n2 1\n")
    text = run("bfs_gpu", "-i", str(p), "-s", "3")
    assert "isolated" in text


def test_cc_gpu_main_weighted_value(mtx, orc):
    """CC_GPU::run_Weighted returns the reference's float value
    (src/algorithm/cpu/cc.cpp:132-133, restated by
    oracle.weighted_closeness_ref and pinned against the reference build)."""
    _, _, pw = mtx
    g = dawn.HostGraph.read_mtx(pw, weighted=True)
    text = run("cc_gpu", "-i", pw, "-s", "17", "-w", "true")
    val = float([l for l in text.splitlines() if l.startswith("Closeness")][0].split(":")[1])
    want = float(orc.weighted_closeness_ref(orc.sssp(g.row_ptr, g.col, g.weight, 17)))
    assert abs(val - want) <= 5e-6  # the main prints five decimals
