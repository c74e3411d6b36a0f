"""nvms blocks (pfm_set_speculation) against testing every step on the host.

Bar: the exit step and event, positions, velocities, forces, interaction
energies, kinetic energy, temperature and simulated time bit-identical to
the step-by-step path, whether the exit falls inside a block, on a block's
first step or never; over split calls, in both precisions, with and without
captured blocks, and with output records (records after the exit step are
dropped).  After an exit the next steps start from a fresh sort, so a
continuation agrees to rounding: over 20 steps, positions within 1e-11
relative in FP64 and 1e-8 in mixed precision.  Replica batches stop when
every replica meets the test: the same bar with per-replica energies.
"""
import time

import numpy as np
import pytest

from pfmds_amd import capi, configs, system
from tests.harness import gpu_ctx, rel_inf

pytestmark = pytest.mark.gpu


def run(s, prec, spec, ms_de, calls, graphs=True, records=None):
    c = gpu_ctx(s, prec)
    c.set_speculation(spec)
    c.set_graphs(graphs)
    got = []
    if records:
        c.set_records(records[0], records[1], 4, lambda r: got.append((r["step"], r.get("pe"), r.get("pos"))))
    out = None
    for a, b in calls:
        last, ev = c.run("nvms", 1.0, a, b, True, ms_de)
        out = (last, ev)
        if ev == capi.EV_CONVERGED:
            break
    st = c.state()
    e = c.energies()
    return c, out, st, e, got


def same(x, y):
    (ca, oa, sa, ea, ga), (cb, ob, sb, eb, gb) = x, y
    assert oa == ob
    for p, q in zip(sa, sb):
        assert np.array_equal(p, q)
    for k in ("pe", "ke", "temp", "time"):
        assert ea[k] == eb[k], k
    assert np.array_equal(ea["e_inter"], eb["e_inter"])
    assert [g[0] for g in ga] == [g[0] for g in gb]
    for g, h in zip(ga, gb):
        assert g[1] == h[1] and (g[2] is None) == (h[2] is None)
        if g[2] is not None:
            assert np.array_equal(g[2], h[2])


def pick_ms_de(s, prec, limit):
    """The smallest threshold of a ladder at which the step-by-step run exits before `limit`."""
    for ms_de in (1e-3, 1e-2, 5e-2, 0.2, 1.0):
        c, out, _, _, _ = run(s, prec, False, ms_de, [(0, limit)])
        c.close()
        if out[1] == capi.EV_CONVERGED and out[0] > 3:
            return ms_de, out[0]
    return 1e-30, None


@pytest.mark.parametrize("case,prec", [("fcc", capi.FP64), ("fcc", capi.MIXED), ("grcu", capi.MIXED)])
def test_nvms_blocks_bit_identical(case, prec):
    if case == "fcc":
        s = configs.fcc_cu_rgl(nx=6, seed=81, temperature=50.0)
    else:
        s = configs.graphene_on_cu(cx=10, cy=6, gx=10, gy=6, layers=4, seed=82, temperature=50.0)
    ms_de, tc = pick_ms_de(s, prec, 300)
    print(f"[nvms blocks {case}] n={s.n} ms_de={ms_de} exit={tc}", flush=True)
    for calls in ([(0, 300)], [(0, 0), (1, 37), (38, 300)]):
        for graphs in (True, False):
            a = run(s, prec, False, ms_de, calls, graphs)
            b = run(s, prec, True, ms_de, calls, graphs)
            same(a, b)
            # continuation after the exit: a fresh sort, rounding-level agreement
            if a[1][1] == capi.EV_CONVERGED:
                t = a[1][0]
                for x in (a, b):
                    x[0].run("nve", 1.0, t, t + 20, True)
                pa, pb = a[0].state(), b[0].state()
                tp, tv = (1e-11, 1e-9) if prec == capi.FP64 else (1e-8, 1e-6)
                assert rel_inf(pb[0], pa[0]) < tp and rel_inf(pb[1], pa[1]) < tv
            a[0].close()
            b[0].close()
    # no exit within the run
    same(run(s, prec, False, 1e-30, [(0, 45)]), run(s, prec, True, 1e-30, [(0, 45)]))


def test_nvms_blocks_with_records():
    s = configs.fcc_cu_rgl(nx=6, seed=83, temperature=50.0)
    ms_de, tc = pick_ms_de(s, capi.FP64, 300)
    a = run(s, capi.FP64, False, ms_de, [(0, 300)], records=((3,), (7,)))
    b = run(s, capi.FP64, True, ms_de, [(0, 300)], records=((3,), (7,)))
    same(a, b)
    if tc is not None:
        assert max(g[0] for g in b[4]) < tc
    a[0].close()
    b[0].close()


def test_nvms_blocks_medium_slab():
    """A 500-step nvms run of a medium graphene/Cu slab that never meets the
    exit test: the blocks reach the same last step, event, state and energies
    as testing every step, bit for bit (the timing is printed, not judged)."""
    s = configs.graphene_on_cu(cx=30, cy=17, gx=31, gy=18, layers=4, seed=84, temperature=50.0)
    res = []
    for spec in (False, True):
        c = gpu_ctx(s, capi.MIXED)
        c.set_speculation(spec)
        c.run("nvms", 0.5, 0, 20, True, 1e-30)
        t0 = time.perf_counter()
        out = c.run("nvms", 0.5, 21, 520, True, 1e-30)
        dt = (time.perf_counter() - t0) / 500
        print(f"  n={s.n} nvms blocks={spec}: {dt * 1e3:.4f} ms/step", flush=True)
        res.append((out, c.state(), c.energies()))
        c.close()
    (oa, sa, ea), (ob, sb, eb) = res
    assert oa == ob == (520, capi.EV_DONE)
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)
    for k in ("pe", "ke", "temp", "time"):
        assert ea[k] == eb[k], k
    assert np.array_equal(ea["e_inter"], eb["e_inter"])


def batch_replicas():
    # a dilute LJ lattice (small energy changes) and three Cu crystals near 50 K
    return [
        configs.lj_gas_lattice(400, 0.01, 11, 20.0),
        configs.fcc_cu_rgl(nx=4, seed=12, dt=0.5, temperature=50.0),
        configs.fcc_cu_rgl(nx=6, seed=13, dt=0.5, temperature=60.0),
        configs.fcc_cu_rgl(nx=5, ny=4, nz=7, seed=15, dt=0.5, temperature=40.0),
    ]


def run_batch(b, prec, spec, ms_de, calls, graphs=True):
    c = system.build_batch_context(b, prec)
    c.set_speculation(spec)
    c.set_graphs(graphs)
    out = None
    for a, e in calls:
        last, ev = c.run("nvms", 0.5, a, e, True, ms_de)
        out = (last, ev)
        if ev == capi.EV_CONVERGED:
            break
    res = (out, c.state(), c.replica_energies(), c.energies())
    c.close()
    return res


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED])
def test_nvms_blocks_batch(prec):
    b = system.merge_replicas(batch_replicas())
    ms_de, tc = 1e-30, None
    # the batch's largest per-step change falls from about 7e-2 to a plateau
    # near 5e-4 by step 150 (the first step changes by 9e-4)
    for m in (1e-4, 5.5e-4, 7e-4, 8.5e-4, 3e-3, 1e-2):
        out = run_batch(b, prec, False, m, [(0, 300)])[0]
        if out[1] == capi.EV_CONVERGED and out[0] > 3:
            ms_de, tc = m, out[0]
            break
    print(f"[nvms blocks batch] {b.n_rep} replicas n={b.system.n} ms_de={ms_de} exit={tc}", flush=True)
    assert tc is not None, "no exit within 300 steps: the batch exit is not exercised"
    for calls in ([(0, 300)], [(0, 0), (1, 37), (38, 300)]):
        for graphs in (True, False):
            (oa, sa, ra, ea), (ob, sb, rb, eb) = (run_batch(b, prec, sp, ms_de, calls, graphs) for sp in (False, True))
            assert oa == ob
            for x, y in zip(sa, sb):
                assert np.array_equal(x, y)
            for k in ("pe", "ke", "temp"):
                assert np.array_equal(ra[k], rb[k]), k
                assert ea[k] == eb[k], k
            assert np.array_equal(ra["e_inter"], rb["e_inter"])
