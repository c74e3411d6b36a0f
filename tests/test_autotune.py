"""The launch-shape autotuner's host logic with a stand-in handle (no GPU):
the fastest candidate is set and cached, FP64 skips workgroups above 256
threads, graphs and co-scheduled passes are kept only when faster, and a
cached shape is applied without timing."""
import re
import time

from pfmds_amd import autotune, capi, configs


class FakeCtx:
    """Steps cost a time that depends on the launch shape."""

    def __init__(self, cost, graph_gain=1.0, fusion_gain=1.0):
        self.cost, self.shape, self.set_calls, self.runs = cost, None, [], []
        self.graphs, self.graph_gain = False, graph_gain
        self.fusion, self.fusion_gain = True, fusion_gain

    def set_fusion(self, on):
        self.fusion = on

    def set_tuning(self, lanes, block):
        self.shape = (lanes, block)
        self.set_calls.append(self.shape)

    def set_graphs(self, on):
        self.graphs = on

    def run(self, integ, dt, s0, s1, energy):
        self.runs.append((s0, s1))
        t = self.cost[self.shape] * (s1 - s0 + 1)
        t = t / self.fusion_gain if self.fusion else t
        time.sleep(t / self.graph_gain if self.graphs else t)


def test_tune_sets_and_caches_the_fastest_shape(tmp_path):
    cost = {c: 2e-3 for c in autotune.CANDIDATES}
    cost[(16, 512)] = 2e-4
    ctx = FakeCtx(cost)
    s = configs.fcc_cu_rgl(nx=3)
    path = str(tmp_path / "t.json")
    lanes, block, graphs, src, nxt = autotune.autotune(ctx, s, "nve", 1.0, capi.MIXED, first_step=11, path=path,
                                                       steps=5)
    assert (lanes, block, src) == (16, 512, "tuned") and ctx.shape == (16, 512)
    # consecutive, non-overlapping step ranges starting at first_step
    assert ctx.runs[0][0] == 11 and all(a[1] + 1 == b[0] for a, b in zip(ctx.runs, ctx.runs[1:]))
    assert nxt == ctx.runs[-1][1] + 1
    ctx2 = FakeCtx(cost)
    lanes, block, graphs2, src, nxt = autotune.autotune(ctx2, s, "nve", 1.0, capi.MIXED, first_step=5, path=path)
    assert (lanes, block, graphs2, src, nxt) == (16, 512, graphs, "cache", 5) and ctx2.runs == []


def test_graphs_chosen_when_faster():
    cost = {c: 1e-3 for c in autotune.CANDIDATES}
    fast = FakeCtx(cost, graph_gain=3.0)
    assert autotune.tune(fast, "nve", 1.0, capi.MIXED, steps=3, repeats=1)[2] is True and fast.graphs
    slow = FakeCtx(cost, graph_gain=0.3)
    assert autotune.tune(slow, "nve", 1.0, capi.MIXED, steps=3, repeats=1)[2] is False and not slow.graphs


def test_fusion_kept_only_when_faster():
    cost = {c: 1e-3 for c in autotune.CANDIDATES}
    fast = FakeCtx(cost, fusion_gain=3.0)
    autotune.tune(fast, "nve", 1.0, capi.MIXED, steps=3, repeats=1)
    assert fast.fusion and fast.fused
    slow = FakeCtx(cost, fusion_gain=0.3)
    autotune.tune(slow, "nve", 1.0, capi.MIXED, steps=3, repeats=1)
    assert not slow.fusion and not slow.fused


def test_fp64_skips_large_workgroups(tmp_path):
    cost = {c: 1e-4 for c in autotune.CANDIDATES}
    ctx = FakeCtx(cost)
    _, _, _, table, _ = autotune.tune(ctx, "nve", 1.0, capi.FP64, steps=2, repeats=1)
    assert all(int(re.match(r"\d+x(\d+)", k).group(1)) <= 256 for k in table)
    assert "8x256" in table


def test_signature_separates_workloads():
    a = configs.fcc_cu_rgl(nx=3)
    b = configs.fcc_cu_rgl(nx=4)
    assert autotune.signature(a, capi.MIXED) != autotune.signature(b, capi.MIXED)
    assert autotune.signature(a, capi.MIXED) != autotune.signature(a, capi.FP64)
    assert autotune.signature(a, capi.MIXED) == autotune.signature(configs.fcc_cu_rgl(nx=3), capi.MIXED)


class PassCtx(FakeCtx):
    """Two interactions; interaction 2's own passes run faster at 16 lanes."""

    n_inter = 2

    def __init__(self, cost, **kw):
        super().__init__(cost, **kw)
        self.pass_set = {}

    def set_pass_lanes(self, inter, lanes):
        self.pass_set[inter] = lanes

    def run(self, integ, dt, s0, s1, energy):
        f = 0.5 if (not self.fusion and self.pass_set.get(2) == 16) else 1.0
        self.runs.append((s0, s1))
        t = self.cost[self.shape] * (s1 - s0 + 1) * f
        t = t / self.fusion_gain if self.fusion else t
        time.sleep(t / self.graph_gain if self.graphs else t)


def test_pass_lanes_with_separate_launches(tmp_path):
    cost = {c: 1e-3 for c in autotune.CANDIDATES}
    cost[(8, 256)] = 5e-4
    ctx = PassCtx(cost, fusion_gain=0.5)
    s = configs.binary_lj_gas(n=200, box=20.0, seed=5)
    path = str(tmp_path / "t.json")
    autotune.autotune(ctx, s, "nve", 1.0, capi.MIXED, path=path, steps=4)
    assert not ctx.fused and ctx.pass_lanes == {2: 16} and ctx.pass_set == {1: 0, 2: 16}
    ctx2 = PassCtx(cost)
    assert autotune.apply_cached(ctx2, autotune.signature(s, capi.MIXED), path)
    assert ctx2.pass_set == {2: 16} and ctx2.fusion is False and ctx2.runs == []


def test_pass_lanes_not_tried_with_shared_launches():
    cost = {c: 1e-3 for c in autotune.CANDIDATES}
    ctx = PassCtx(cost, fusion_gain=3.0)
    autotune.tune(ctx, "nve", 1.0, capi.MIXED, steps=3, repeats=1)
    assert ctx.fused and ctx.pass_lanes == {} and ctx.pass_set == {}
