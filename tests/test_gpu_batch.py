"""Batched boxes on the device: a captured HIP graph replays the eager
pipeline bit for bit (several replays, with eager work on the same stream
between them), and the autotuner changes timing only, not results."""
import numpy as np
import pytest

from monofonic_amd import F32, F64
from monofonic_amd.batch import Batch, autotune
from monofonic_amd.lpt import ICConfig, ICRun

BOXES = [(32, 101, 50.0), (48, 102, 80.0), (64, 103, 120.0)]


def _fields(runs):
    return [[r.field(c).download_r().copy() for c in range(6)] for r in runs]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F32, F64])
def test_graph_replay_matches_eager(ctx, prec):
    b = Batch(ctx, BOXES, lpt_order=2, prec=prec, tune=False)
    other = ICRun(ctx, ICConfig(n=32, box=70.0, zstart=24.0, lpt_order=3, prec=prec, seed=7))
    try:
        _replay_check(b, other)
    finally:
        b.close()
        other.close()


def _replay_check(b, other):
    b.execute()
    ref = _fields(b.runs)
    b.capture()
    for rep in range(3):
        for r in b.runs:
            for c in range(6):
                r.field(c).zero()
        other.execute()  # eager work between replays
        b.launch()
        got = _fields(b.runs)
        for i in range(len(BOXES)):
            for c in range(6):
                assert np.array_equal(got[i][c], ref[i][c]), (rep, i, c)
        print(f"graph replay {rep}: {len(BOXES)} boxes bit-identical to the eager run")


@pytest.mark.gpu
def test_autotune_keeps_results(ctx):
    cfg = ICConfig(n=64, box=100.0, zstart=24.0, lpt_order=2, prec=F32, seed=5)
    r = ICRun(ctx, cfg)
    r.execute()
    before = [r.field(c).download_r().copy() for c in range(6)]
    ms = autotune(ctx, 64, F32, 2)
    assert ms > 0
    r.execute()
    for c in range(6):
        assert np.array_equal(r.field(c).download_r(), before[c])
    print(f"autotuned 64^3 2LPT FP32: {ms:.3f} ms per box")
    r.close()
