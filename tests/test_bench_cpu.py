"""CPU checks of bench.py's argument handling (no GPU): any --steps/--warmup pair is
accepted and rounded up to whole 10-step blocks, so the workload keeps n_rec = 10."""
import subprocess
import sys

import pytest

from conftest import ROOT


def _bench():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import bench
    return bench


@pytest.mark.parametrize("steps,warmup,expect", [
    ((20, 5), None, (20, 5, 1)),      # 20 timed steps after the minimum warm-up of 100
    ((7, 3), None, (10, 10, 1)),
    ((500, 50), None, (50, 2, 10)),
    ((1, 0), None, (10, 10, 1)),
    ((100, 50), None, (50, 2, 2)),
    ((30, 200), None, (10, 20, 3)),
    ((40, 0), None, (20, 5, 2)),
])
def test_plan_steps(steps, warmup, expect):
    b = _bench()
    ni, wo, to = b.plan_steps(*steps)
    assert (ni, wo, to) == expect
    assert ni % 10 == 0                       # whole n_recalc blocks: solve_split keeps n_rec = 10
    assert ni * to >= steps[0] and ni * to - steps[0] < 10
    assert ni * wo >= steps[1] and ni * wo >= b.MIN_WARMUP


def test_plan_steps_never_exits():
    b = _bench()
    for s in range(1, 120):
        for w in range(0, 60):
            ni, wo, to = b.plan_steps(s, w)
            assert ni in (10, 20, 30, 40, 50) and to >= 1 and wo >= 1
            assert ni * to >= s and ni * to - s < 10 and ni * wo >= b.MIN_WARMUP


def test_gpus_above_one_exits_with_message():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2", "--steps", "20", "--warmup", "5"],
                       cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
