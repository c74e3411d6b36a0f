"""Runtime group growth on the GPU (pfm_add_group_change) against the CPU
oracle and against itself.

Bar (FP64): positions, velocities and energies within 1e-10 relative of the
oracle after every chunk, active counts equal after every chunk, lists
equal as sets after rebuilds, for the deposition case with lj and with rjl.
Captured blocks are used away from the steps where a count changes and give
bit-identical results to direct launches; records carry the active counts
of their step.  The same bar holds for lj1g, for a tb sheet whose atoms join
one at a time, and for metal adatoms joining the rjl and morsec metal group
under graphene (tb + morsec with nearest-neighbour terms).  A growing
graphene group of morsec with a tb list stops at step 0 with the reference's
"not enough gr nearest neibs" status, on the GPU and in the oracle.  A
rebosc sheet growing the same way meets rebosc's own bars (forces within 1e-6
of the largest, positions and velocities within 1e-8).
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs, system
from pfmds_amd.system import Interaction
from tests import growth_case
from tests.harness import gpu_ctx, list_sets, oracle_ctx, rel_inf

pytestmark = pytest.mark.gpu

CHUNKS = [(0, 0), (1, 15), (16, 37), (38, 60)]


@pytest.mark.parametrize("kind,side", [("lj", 4), ("rjl", 4), ("lj", 8), ("lj1g", 4), ("lj1g", 8)])
def test_growth_matches_oracle(kind, side):
    # 4^3 atoms: lists searched per row; 8^3: three cells a side, searched per cell tile
    s = growth_case.deposition(n_side=side, kind=kind)
    print(f"[growth {kind}] n={s.n}", flush=True)
    g = gpu_ctx(s, capi.FP64)
    o = oracle_ctx(s)
    want = growth_case.expected_counts([c[1] for c in CHUNKS], s.n // 2, s.n, 5, 40, 7)
    for (s0, s1), n in zip(CHUNKS, want):
        g.run("nve", 1.0, s0, s1, True)
        o.run("nve", 1.0, s0, s1, True)
        assert g.group_count(2) == o.group_count(2) == n
        (pg, vg, fg), (po, vo, fo) = g.state(), o.state()
        assert rel_inf(pg, po) < 1e-10 and rel_inf(vg, vo) < 1e-10 and rel_inf(fg, fo) < 1e-10, s1
        eg, eo = g.energies(), o.energies()
        assert eg["pe"] == pytest.approx(eo["pe"], rel=1e-10) and eg["temp"] == pytest.approx(eo["temp"], rel=1e-10)
        if s1 % 10 == 0 or s1 == 15:
            ga, _ = list_sets(g, 1, 1)
            oa, _ = list_sets(o, 1, 1)
            assert ga == oa, s1
    g.close()
    o.close()


def test_growth_graphs_equal_direct_launches():
    # changes every 30 steps: blocks of 10 between them are captured
    s = growth_case.deposition(ts1=5, ts2=200, frec=30)
    out = []
    for graphs in (True, False):
        c = gpu_ctx(s, capi.MIXED)
        c.set_graphs(graphs)
        res = []
        for s0, s1 in ((0, 0), (1, 77), (78, 151)):
            c.run("nve", 1.0, s0, s1, True)
            res.append((c.state(), c.energies()["pe"], c.group_count(2)))
        launched = c.graph_count()[1]
        print(f"  graphs={graphs}: {launched} launches", flush=True)
        assert (launched > 3) if graphs else launched == 0
        out.append(res)
        c.close()
    for (sa, pa, na), (sb, pb, nb) in zip(*out):
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y)
        assert pa == pb and na == nb
    assert out[0][-1][2] == 32 + 1 + len(range(35, 152, 30))


def test_growth_counts_in_records():
    s = growth_case.deposition()
    c = gpu_ctx(s, capi.FP64)
    got = {}
    c.set_records((3,), (9,), 2, lambda r: got.__setitem__(r["step"], r))
    c.run("nve", 1.0, 0, 60, False)
    steps = sorted(got)
    want = growth_case.expected_counts(steps, 32, 64, 5, 40, 7)
    assert [int(got[t]["group_n"][2]) for t in steps] == want
    assert all(int(got[t]["group_n"][1]) == 32 and int(got[t]["group_n"][3]) == 64 for t in steps)
    # the temperature of a record divides by its step's count
    for t in steps:
        if "ke" in got[t]:
            n = int(got[t]["group_n"][2])
            assert got[t]["temp"] == pytest.approx(2 * got[t]["ke"] / capi_kb() / (3 * n), rel=1e-14)
    c.close()


def capi_kb():
    return 1.3806488 / 1.6021765654 * (1.0 / 10000.0)


@pytest.mark.parametrize("case", ["tb", "adatoms"])
def test_growth_tb_and_cos_match_oracle(case):
    s = growth_case.carbon_growth_tb() if case == "tb" else growth_case.metal_adatoms()
    g2 = 1 if case == "tb" else 2
    print(f"[growth {case}] n={s.n}", flush=True)
    g = gpu_ctx(s, capi.FP64)
    o = oracle_ctx(s)
    for s0, s1 in CHUNKS:
        g.run("nve", 0.5, s0, s1, True)
        o.run("nve", 0.5, s0, s1, True)
        assert g.group_count(g2) == o.group_count(g2)
        (pg, vg, fg), (po, vo, fo) = g.state(), o.state()
        assert rel_inf(pg, po) < 1e-10 and rel_inf(vg, vo) < 1e-10 and rel_inf(fg, fo) < 1e-10, s1
        assert g.energies()["pe"] == pytest.approx(o.energies()["pe"], rel=1e-10)
        if s1 % 10 == 0:
            for k, it in enumerate(s.interactions, start=1):
                for li in range(1, len(it.lists) + 1):
                    assert list_sets(g, k, li)[0] == list_sets(o, k, li)[0], (s1, k, li)
    size = len(s.group_indices(g2))
    assert o.group_count(g2) < size  # still growing at the end: inactive rows were exercised
    g.close()
    o.close()


def test_growing_graphene_group_with_tb_stops():
    s = growth_case.carbon_growth_with_cos()
    codes = []
    for ctx in (gpu_ctx(s, capi.FP64), oracle_ctx(s)):
        with pytest.raises(capi.PfmError) as ei:
            ctx.run("nve", 0.5, 0, 5, True)
        assert "gr nearest neibs" in ei.value.msg
        codes.append(ei.value.code)
        ctx.close()
    assert codes == [capi.E_GRAPHENE_NN, capi.E_GRAPHENE_NN]


def test_growth_rebosc_matches_oracle():
    s = growth_case.carbon_growth_tb()
    s.interactions = [Interaction("rebosc", list(configs.REBO_C), [(1, 1, 8, 2.3, 10)])]
    print(f"[growth rebosc] n={s.n}", flush=True)
    g = gpu_ctx(s, capi.FP64)
    o = oracle_ctx(s)
    for s0, s1 in CHUNKS:
        g.run("nve", 0.5, s0, s1, True)
        o.run("nve", 0.5, s0, s1, True)
        assert g.group_count(1) == o.group_count(1)
        (pg, vg, fg), (po, vo, fo) = g.state(), o.state()
        # rebosc's bars: central-difference forces within 1e-6 of the largest
        # force, trajectories within 1e-8
        assert rel_inf(pg, po) < 1e-8 and rel_inf(vg, vo) < 1e-8 and rel_inf(fg, fo) < 1e-6, s1
        assert g.energies()["pe"] == pytest.approx(o.energies()["pe"], rel=1e-10)
        if s1 % 10 == 0:
            assert list_sets(g, 1, 1)[0] == list_sets(o, 1, 1)[0], s1
    assert o.group_count(1) < len(s.group_indices(1))
    g.close()
    o.close()


def test_growth_in_a_batch_matches_single_runs():
    # three lj deposition systems of different sizes and seeds in one batch:
    # each replica grows its own group by the shared rule, as its own md() run
    reps = [growth_case.deposition(n_side=4, seed=101), growth_case.deposition(n_side=5, seed=102, spacing=4.0),
            growth_case.deposition(n_side=3, seed=103, spacing=4.4)]
    for s in reps:
        # a batch matches groups by their species lists: keep the static
        # all-atoms group apart from the growing group of the same species
        s.groups[2] = ["A", "D", "#"]
    b = system.merge_replicas(reps)
    print(f"[growth batch] {b.n_rep} replicas, {b.system.n} atoms", flush=True)
    g = system.build_batch_context(b, capi.FP64)
    singles = [gpu_ctx(s, capi.FP64) for s in reps]
    for s0, s1 in CHUNKS:
        g.run("nve", 1.0, s0, s1, True)
        for c in singles:
            c.run("nve", 1.0, s0, s1, True)
        pos, vel, frc = g.state()
        er = g.replica_energies()
        for r, c in enumerate(singles):
            sl = b.replica_slice(r)
            p1, v1, f1 = c.state()
            assert rel_inf(pos[sl], p1) < 1e-10 and rel_inf(vel[sl], v1) < 1e-10 and rel_inf(frc[sl], f1) < 1e-10, (r, s1)
            e = c.energies()
            assert er["pe"][r] == pytest.approx(e["pe"], rel=1e-10)
            assert er["temp"][r] == pytest.approx(e["temp"], rel=1e-10)
        # the batch matches groups by species: its growing group is group 1
        assert g.group_count(1) == sum(c.group_count(2) for c in singles)
    g.close()
    for c in singles:
        c.close()


def test_growth_in_a_batch_matches_oracle():
    """Each replica of a growing batch against the oracle running that
    replica alone: positions within 1e-10, velocities and forces within 1e-8
    relative, per-replica energies within 1e-9, the same group counts."""
    reps = [growth_case.deposition(n_side=4, seed=111), growth_case.deposition(n_side=3, seed=112, spacing=4.4)]
    for s in reps:
        s.groups[2] = ["A", "D", "#"]
    b = system.merge_replicas(reps)
    print(f"[growth batch vs oracle] {b.n_rep} replicas, {b.system.n} atoms", flush=True)
    g = system.build_batch_context(b, capi.FP64)
    oracles = [oracle_ctx(s) for s in reps]
    for s0, s1 in CHUNKS:
        g.run("nve", 1.0, s0, s1, True)
        for o in oracles:
            o.run("nve", 1.0, s0, s1, True)
        pos, vel, frc = g.state()
        er = g.replica_energies()
        for r, o in enumerate(oracles):
            sl = b.replica_slice(r)
            p1, v1, f1 = o.state()
            assert rel_inf(pos[sl], p1) < 1e-10 and rel_inf(vel[sl], v1) < 1e-8 and rel_inf(frc[sl], f1) < 1e-8, (r, s1)
            assert er["pe"][r] == pytest.approx(o.energies()["pe"], rel=1e-9)
        assert g.group_count(1) == sum(o.group_count(2) for o in oracles)
    g.close()
    for o in oracles:
        o.close()
