"""The launch-shape autotuner's host logic with a stand-in handle (no GPU):
the fastest candidate is set and cached, FP64 skips workgroups above 256
threads, and a cached shape is applied without timing."""
import time

from pfmds_amd import autotune, capi, configs


class FakeCtx:
    """Steps cost a time that depends on the launch shape."""

    def __init__(self, cost):
        self.cost, self.shape, self.set_calls, self.runs = cost, None, [], []

    def set_tuning(self, lanes, block):
        self.shape = (lanes, block)
        self.set_calls.append(self.shape)

    def run(self, integ, dt, s0, s1, energy):
        self.runs.append((s0, s1))
        time.sleep(self.cost[self.shape] * (s1 - s0 + 1))


def test_tune_sets_and_caches_the_fastest_shape(tmp_path):
    cost = {c: 2e-4 for c in autotune.CANDIDATES}
    cost[(16, 512)] = 2e-5
    ctx = FakeCtx(cost)
    s = configs.fcc_cu_rgl(nx=3)
    path = str(tmp_path / "t.json")
    lanes, block, src, nxt = autotune.autotune(ctx, s, "nve", 1.0, capi.MIXED, first_step=11, path=path, steps=5)
    assert (lanes, block, src) == (16, 512, "tuned") and ctx.shape == (16, 512)
    # consecutive, non-overlapping step ranges starting at first_step
    assert ctx.runs[0][0] == 11 and all(a[1] + 1 == b[0] for a, b in zip(ctx.runs, ctx.runs[1:]))
    assert nxt == ctx.runs[-1][1] + 1
    ctx2 = FakeCtx(cost)
    lanes, block, src, nxt = autotune.autotune(ctx2, s, "nve", 1.0, capi.MIXED, first_step=5, path=path)
    assert (lanes, block, src, nxt) == (16, 512, "cache", 5) and ctx2.runs == []


def test_fp64_skips_large_workgroups(tmp_path):
    cost = {c: 1e-4 for c in autotune.CANDIDATES}
    ctx = FakeCtx(cost)
    _, _, table, _ = autotune.tune(ctx, "nve", 1.0, capi.FP64, steps=2, repeats=1)
    assert all(int(k.split("x")[1]) <= 256 for k in table)
    assert "8x256" in table


def test_signature_separates_workloads():
    a = configs.fcc_cu_rgl(nx=3)
    b = configs.fcc_cu_rgl(nx=4)
    assert autotune.signature(a, capi.MIXED) != autotune.signature(b, capi.MIXED)
    assert autotune.signature(a, capi.MIXED) != autotune.signature(a, capi.FP64)
    assert autotune.signature(a, capi.MIXED) == autotune.signature(configs.fcc_cu_rgl(nx=3), capi.MIXED)
