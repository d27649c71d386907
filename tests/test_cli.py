"""The gpu_canon_label CLI (reference main.cpp surface) in --sequential mode:
output file bytes, verdict lines and error handling."""
import os
import subprocess

import pytest

import gcl
from conftest import GOLDEN, golden_output

CLI = gcl.CLI_PATH


def run(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


def gpath(name):
    return os.path.join(GOLDEN, "graphs", name + ".dimacs")


def test_cli_output_bytes_sequential(golden, tmp_path):
    for e in golden["singles"][::3]:
        p = run([gpath(e["name"]), "--sequential", f"--max-aut={e['max_aut']}"], tmp_path)
        assert p.returncode == 0, p.stderr
        assert (tmp_path / "output").read_bytes() == golden_output(e), e["name"]
        assert f"Total search tree nodes created (sequential): {e['nodes']}" in p.stdout


def test_cli_verdicts_sequential(golden, tmp_path):
    for pr in golden["pairs"]:
        p = run([gpath(pr["a"]), gpath(pr["b"]), "--sequential"], tmp_path)
        assert p.returncode == 0, p.stderr
        assert f"\nResult: {pr['verdict']}\n" in p.stdout
        assert "Graph 1: " in p.stdout and "Computing canonical form..." in p.stdout


def test_cli_help(tmp_path):
    p = run(["-h"], tmp_path)
    assert p.returncode == 0 and "--max-aut=N" in p.stdout and "Usage:" in p.stdout
    assert run(["--help"], tmp_path).returncode == 0


def test_cli_errors(tmp_path):
    p = run([], tmp_path)
    assert p.returncode == 1 and "[ERROR] Invalid arguments." in p.stderr and "Usage:" in p.stdout
    g = gpath("c6")
    assert run([g, g, g], tmp_path).returncode == 1
    p = run([g, "--frobnicate"], tmp_path)
    assert p.returncode == 1 and "[ERROR] Unknown option: --frobnicate" in p.stderr
    p = run([str(tmp_path / "missing.dimacs"), "--sequential"], tmp_path)
    assert p.returncode == 1 and "[ERROR] Failed to open DIMACS file:" in p.stderr
    for bad in ("--max-aut=abc", "--max-aut=0", "--max-aut="):
        p = run([g, bad, "--sequential"], tmp_path)
        assert p.returncode == 1 and "[ERROR]" in p.stderr, bad
    assert not (tmp_path / "output").exists()


def test_cli_rejects_multigraph_and_bad_ids(tmp_path):
    f = tmp_path / "multi.dimacs"
    f.write_text("p edge 3 3\ne 1 2\ne 2 1\ne 2 3\n")
    p = run([str(f), "--sequential"], tmp_path)
    assert p.returncode == 1 and "repeated edge" in p.stderr
    f2 = tmp_path / "range.dimacs"
    f2.write_text("p edge 3 1\ne 1 9\n")
    p = run([str(f2), "--sequential"], tmp_path)
    assert p.returncode == 1 and "[ERROR]" in p.stderr


def test_cli_dimacs_conventions(tmp_path):
    # comments, loops (dropped), colour lines, 'p' tag text, trailing blanks
    f = tmp_path / "conv.dimacs"
    f.write_text("c a comment\np edge 4 5\nn 2 7\ne 1 2\ne 2 3\ne 3 3\ne 3 4\ne 4 1\n\n")
    p = run([str(f), "--sequential"], tmp_path)
    assert p.returncode == 0, p.stderr
    g = gcl.read_dimacs(str(f))
    assert g.n == 4 and g.m == 4 and list(g.color) == [0, 7, 0, 0]
    r = gcl.canon_label(g, engine="cpu")
    assert (tmp_path / "output").read_text() == gcl.format_adjacency(r.keys, g.n)
