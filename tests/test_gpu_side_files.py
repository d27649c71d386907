"""Side files (gcl_set_side_files; counterpart of Nauty_Parallel's
output_prefix and its per-block CSVs, reference src/nauty_parallel.cu:882-981):
one row per worker whose graph and node totals equal the launch's results,
nodes per CU adding up to the same total, and the CLI naming them after the
graph file's stem when GCL_SIDE_FILES is set."""
import csv
import os
import subprocess
import sys

import pytest

import gcl

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def test_side_files_of_a_batch(tmp_path):
    graphs = [gcl.random_regular(2 * n, 3, n) for n in (70, 150, 300)] + \
             [gcl.permute(gcl.erdos_renyi_mean_degree(n, 5.0, n), n) for n in (20, 90, 400, 900)] + \
             [gcl.torus(12, 17)]
    prefix = str(tmp_path / "run")
    _log(f"side files batch of {len(graphs)}")
    gcl.set_side_files(prefix)
    try:
        r = gcl.batch_canon(gcl.Batch(graphs))
    finally:
        gcl.set_side_files(None)
    rows = _rows(prefix + "_counters.csv")
    assert len(rows) == r.workers
    assert sum(int(x["GRAPHS"]) for x in rows) == len(graphs)
    assert sum(int(x["NODES"]) for x in rows) == r.total_nodes
    assert sum(int(x["HELPED_CHILDREN"]) for x in rows) > 0  # the regular graphs publish root children
    # root children published by the regular graphs are adopted from helpers
    # or rejected by their code
    assert sum(int(x["ADOPTED_CHILDREN"]) + int(x["REJECTED_CHILDREN"]) for x in rows) > 0
    for x in rows:
        assert int(x["XCC"]) < 16 and int(x["SE"]) < 4 and int(x["CU"]) < 16
        pct = [float(x[k]) for k in ("SEARCH_pct", "WAIT_CHILDREN_pct", "HELP_pct", "IDLE_pct")]
        assert all(0.0 <= p <= 100.0 for p in pct) and sum(pct) <= 100.5
        # the phases of the own search lie within its share
        phases = [float(x[k]) for k in ("REFINE_pct", "CLASSIFY_pct", "ORBITS_pct", "BACKTRACK_pct")]
        assert all(p >= 0.0 for p in phases) and sum(phases) <= float(x["SEARCH_pct"]) + 0.5
    assert sum(float(x["REFINE_pct"]) for x in rows if int(x["GRAPHS"]) > 0) > 0.0
    cu = _rows(prefix + "_nodes_per_cu.csv")
    assert sum(int(x["nodes"]) for x in cu) == r.total_nodes
    # off again: a later launch writes nothing
    os.remove(prefix + "_counters.csv")
    gcl.batch_canon(gcl.Batch(graphs[:2]))
    assert not os.path.exists(prefix + "_counters.csv")


def test_cli_side_files_named_after_the_graph(tmp_path):
    g = gcl.random_regular(200, 3, 5)
    f = str(tmp_path / "rr3_200.dimacs")
    gcl.write_dimacs(f, g)
    cli = gcl.CLI_PATH
    env = dict(os.environ, GCL_SIDE_FILES="1")
    _log("cli side files")
    p = subprocess.run([cli, f, "--parallel"], cwd=str(tmp_path), env=env, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    rows = _rows(str(tmp_path / "rr3_200_counters.csv"))
    assert sum(int(x["GRAPHS"]) for x in rows) == 1
    assert sum(int(x["NODES"]) for x in rows) == 201
    assert os.path.exists(str(tmp_path / "rr3_200_nodes_per_cu.csv"))
