"""Launch tuning of the batched path (gcl_device_batch_create_tuned,
gcl_batch_autotune): the labels and node counts of a mixed batch are the same
under every tuning (speculation threshold, smallest size class), and the
autotuner returns one of its candidates, which then runs the batch to the
same results."""
import sys

import numpy as np
import pytest

import gcl

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(batch, tuning):
    db = gcl.DeviceBatch(batch, max_aut=64, tuning=tuning)
    try:
        db.run()
        r = db.results()
        return r.labels.copy(), r.nodes.copy()
    finally:
        db.close()


def test_gpu_tunings_give_identical_results():
    batch = gcl.mix_batch(1500, 17)
    base = gcl.batch_tuning_default()
    _log(f"default tuning {base}")
    ref_labels, ref_nodes = _run(batch, None)
    for t in ({"spec_min_n": 0, "min_class_n": 256}, {"spec_min_n": 64, "min_class_n": 64},
              {"spec_min_n": 256, "min_class_n": 128}):
        _log(f"tuning {t}")
        labels, nodes = _run(batch, t)
        assert np.array_equal(labels, ref_labels), t
        assert np.array_equal(nodes, ref_nodes), t


def test_gpu_autotune_returns_a_candidate():
    batch = gcl.mix_batch(2048, 23)
    _log("autotune")
    best, ms, tried = gcl.autotune(batch, max_aut=64, sample=512)
    _log(f"best {best} {ms:.2f} ms of {tried}")
    assert tried == 12 and ms > 0
    assert best["spec_min_n"] in (0, 64, 128, 256) and best["min_class_n"] in (64, 128, 256)
    a = _run(batch, best)
    b = _run(batch, None)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_gpu_tuning_rejects_unknown_class():
    with pytest.raises(gcl.GclError):
        gcl.DeviceBatch(gcl.mix_batch(8, 1), tuning={"spec_min_n": 128, "min_class_n": 100})
