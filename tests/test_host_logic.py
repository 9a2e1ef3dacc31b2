"""Host-side logic that needs no GPU: the launch-shape autotuner's
coordinate descent (on a stand-in context with a known cost surface), the
batched workload's seeded size distribution, and bench.py's refusal of more
than one GPU."""
import os
import subprocess
import sys

from raygun_amd import autotune, batched

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class _FakeContext:
    """Holds a tuning; `cost` is the frame time the timer reports for it."""

    def __init__(self, start):
        self.t = dict(start)
        self.applied = []

    def tuning(self):
        return dict(self.t)

    def set_tuning(self, **t):
        self.t = dict(t)
        self.applied.append(dict(t))


def test_autotune_finds_the_minimum_of_a_separable_cost():
    start = {"wavefrontPasses": 6, "traceBlocksPerCU": 6, "shadeBlocksPerCU": 8, "drainBlocksPerCU": 4, "hostLag": 2}
    ctx = _FakeContext(start)
    other = _FakeContext(start)  # a second context sharing the tuning

    def cost():
        t = ctx.t
        return abs(t["wavefrontPasses"] - 12) + abs(t["traceBlocksPerCU"] - 8) * 2.0 + abs(t["drainBlocksPerCU"] - 2) * 0.5 + 7.0

    cand = {"wavefrontPasses": [6, 8, 10, 12, 14], "traceBlocksPerCU": [4, 6, 8], "drainBlocksPerCU": [2, 4, 8]}
    best, ms, table = autotune.tune(ctx, None, candidates=cand, contexts=[ctx, other], timer=cost)
    assert best["wavefrontPasses"] == 12 and best["traceBlocksPerCU"] == 8 and best["drainBlocksPerCU"] == 2
    assert ms == 7.0
    assert ctx.t == best and other.t == best  # every context left on the best tuning
    assert len(table) == 1 + 4 + 2 + 2  # the start, then each other candidate once


def test_autotune_confirmation_discards_a_lucky_timing():
    """The shape with traceBlocksPerCU 4 costs 6.0 but its first timing
    reads 5.0 (noise); 8 costs 5.5 every time. The descent alone keeps 4,
    the confirmation rounds keep 8."""
    start = {"wavefrontPasses": 6, "traceBlocksPerCU": 6}
    cand = {"traceBlocksPerCU": [4, 6, 8]}
    for confirm, want, want_ms in ((0, 4, 5.0), (3, 8, 5.5)):
        ctx = _FakeContext(start)
        seen = {}

        def cost():
            b = ctx.t["traceBlocksPerCU"]
            seen[b] = seen.get(b, 0) + 1
            if b == 4:
                return 5.0 if seen[b] == 1 else 6.0
            return {6: 7.0, 8: 5.5}[b]

        best, ms, table = autotune.tune(ctx, None, candidates=cand, timer=cost, confirm=confirm, confirm_reps=2)
        assert best["traceBlocksPerCU"] == want and ms == want_ms and ctx.t == best
        assert len(table) == 3
        assert sum(seen.values()) == 3 + (3 * 2 if confirm else 0)


def test_batched_scene_sizes_are_seeded_and_clipped():
    a = batched.scene_sizes(32, 3)
    assert a == batched.scene_sizes(32, 3)
    assert a != batched.scene_sizes(32, 4)
    assert all(10_000 <= s <= 4_000_000 for s in a)
    assert len(set(a)) > 20  # log-normal spread, not a constant


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
