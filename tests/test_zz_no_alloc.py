"""The allocation-free entry point (gcl_canon_label_noalloc) through the C
ABI: its labels and canonical keys equal gcl_canon_label's, bit for bit, on
two calls in a row (the second with a smaller graph, so that the static
device array holds the first call's contents when the second starts), and
sizes beyond its limits are refused with GCL_E_CAPACITY."""
import sys

import numpy as np
import pytest

import gcl

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check_same(g):
    ref = gcl.canon_label(g, max_aut=10000, engine="gpu")
    got = gcl.canon_label_noalloc(g, max_aut=10000)
    assert np.array_equal(got.label, ref.label), g.n
    assert np.array_equal(got.keys, ref.keys), g.n
    assert got.nodes >= 1 and got.alg_bytes >= got.csr_bytes
    return got


def test_noalloc_matches_canon_label_on_two_calls():
    big = gcl.permute(gcl.random_regular(900, 3, 11), 5)  # publishes root children
    small = gcl.permute(gcl.erdos_renyi_mean_degree(200, 6.0, 12), 6)
    _log(f"noalloc first call n={big.n}")
    _check_same(big)
    _log(f"noalloc second call n={small.n}")
    _check_same(small)
    # coloured and highly symmetric graphs, then the first graph again
    for g in (gcl.torus(16, 20), gcl.cfi(40, True, 3), big):
        _log(f"noalloc n={g.n}")
        _check_same(g)
    colored = gcl.Graph(small.n, small.row_ptr, small.col, (np.arange(small.n) % 3).astype(np.int32))
    _log("noalloc coloured")
    _check_same(colored)


def test_noalloc_refuses_beyond_its_limits():
    for g in (gcl.random_regular(1100, 3, 1), gcl.erdos_renyi_mean_degree(1000, 40.0, 2)):
        with pytest.raises(gcl.GclError) as ei:
            gcl.canon_label_noalloc(g)
        assert ei.value.status == -6
