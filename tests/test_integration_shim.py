"""The C++ shim of INTEGRATION.md section 1, compiled as written and linked to
libgcl.so.  It is built against a stand-in with the member layout of the
reference's CSRGraph (src/CSRGraph.h:9-15: vertexNum, edgeNum, dst, srcPtr,
degree, color), fed golden graphs, and its canonical_sequential matrix must
give the reference binary's `output` bytes.  Its Nauty_Parallel must keep the
reference's contract (nauty_parallel.h:9-13): the n*n matrix on success, an
empty vector on failure (here: no GPU).  Runs on the CPU only."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import gcl
from conftest import golden_graph, golden_output

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# stand-in for the caller's graph type: the reference's member names and types
CSRGRAPH_H = r"""
#pragma once
class CSRGraph {
public:
    unsigned int vertexNum = 0;
    unsigned int edgeNum = 0;
    unsigned int* dst = nullptr;
    unsigned int* srcPtr = nullptr;
    int* degree = nullptr;
    int* color = nullptr;
};
"""

# reads "n nnz" then row offsets, column ids and colours, runs both shim
# functions and writes: the parallel result's size, then the sequential matrix
MAIN_CPP = r"""
#include <cstdio>
#include <fstream>
SHIM
int main(int argc, char** argv) {
    std::ifstream in(argv[1]);
    unsigned n = 0, nnz = 0;
    in >> n >> nnz;
    CSRGraph g;
    g.vertexNum = n;
    g.edgeNum = nnz / 2;
    g.srcPtr = new unsigned int[n + 1];
    g.dst = new unsigned int[nnz ? nnz : 1];
    g.color = new int[n ? n : 1];
    for (unsigned i = 0; i <= n; ++i) in >> g.srcPtr[i];
    for (unsigned i = 0; i < nnz; ++i) in >> g.dst[i];
    for (unsigned i = 0; i < n; ++i) in >> g.color[i];
    const unsigned max_aut = (unsigned)std::atoi(argv[3]);
    unsigned int nodes = 0;
    std::vector<unsigned char> par = Nauty_Parallel(g, (int)max_aut, &nodes, "");
    std::vector<unsigned char> seq = canonical_sequential(g, max_aut);
    FILE* f = std::fopen(argv[2], "wb");
    unsigned long long ps = par.size();
    std::fwrite(&ps, sizeof ps, 1, f);
    std::fwrite(par.data(), 1, par.size(), f);
    std::fwrite(seq.data(), 1, seq.size(), f);
    std::fclose(f);
    return 0;
}
"""


def _shim_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text.split("## 1.", 1)[1].split("\n## ", 1)[0]
    blocks = re.findall(r"```cpp\n(.*?)```", sec, re.S)
    assert len(blocks) == 1, "section 1 should hold exactly one C++ block"
    return blocks[0]


@pytest.fixture(scope="module")
def shim_exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    lib_dir = os.path.dirname(gcl.LIB_PATH)
    if not os.path.exists(gcl.LIB_PATH):
        pytest.fail(f"{gcl.LIB_PATH} is not built")
    d = tmp_path_factory.mktemp("shim")
    (d / "CSRGraph.h").write_text(CSRGRAPH_H)
    src = MAIN_CPP.replace("SHIM", _shim_source())
    (d / "main.cpp").write_text(src)
    exe = d / "shim"
    cmd = [gxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-Wno-unused-function", f"-I{d}",
           f"-I{os.path.join(ROOT, 'include')}", str(d / "main.cpp"), "-o", str(exe),
           f"-L{lib_dir}", "-lgcl", f"-Wl,-rpath,{lib_dir}", "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def _run(exe, g, max_aut, tmp_path):
    inp, out = tmp_path / "g.txt", tmp_path / "m.bin"
    color = g.color if g.color is not None else np.zeros(g.n, np.int32)
    with open(inp, "w") as f:
        f.write(f"{g.n} {len(g.col)}\n")
        f.write(" ".join(map(str, np.asarray(g.row_ptr, np.int64))) + "\n")
        f.write(" ".join(map(str, np.asarray(g.col, np.int64))) + "\n")
        f.write(" ".join(map(str, np.asarray(color, np.int64))) + "\n")
    r = subprocess.run([exe, str(inp), str(out), str(max_aut)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    ps = int(np.frombuffer(raw[:8], np.uint64)[0])
    par = np.frombuffer(raw[8:8 + ps], np.uint8)
    seq = np.frombuffer(raw[8 + ps:], np.uint8)
    return par, seq


def _keys_of(mat, n):
    return np.flatnonzero(mat.reshape(n, n)).astype(np.uint32)


def test_shim_compiles_and_sequential_matches_reference_output(shim_exe, golden, tmp_path):
    done = 0
    for e in golden["singles"]:
        g = golden_graph(e["name"])
        if g.n > 400:
            continue
        par, seq = _run(shim_exe, g, e["max_aut"], tmp_path)
        assert seq.size == g.n * g.n, e["name"]
        text = gcl.format_adjacency(_keys_of(seq, g.n), g.n).encode()
        assert text == golden_output(e), e["name"]
        # no GPU here: the reference's failure contract is an empty vector
        assert par.size == 0 or np.array_equal(par, seq), e["name"]
        done += 1
    assert done >= 20
