"""Every benchmark configuration on the GPU, at its full size.

  C0  the binary LJ gas, 4,000 atoms in the 60 A cube, against the CPU
      oracle: lists and FP64 forces at step 0, a 50-step nvt trajectory, and
      md()'s schedule (nvt, then nve, then nvms to its exit) driven through
      the same loop for both;
  C3  replicas: each replica of a batch against the oracle running that
      replica alone; the 512-replica, 3,374,724-atom batch through
      size-independent properties (every replica's force sum vanishes to
      2e-8 of its sum of |F| in mixed precision, two
      runs are bit-identical, the replicas' energies add up to the batch's);
  C4  the learned surrogate on the 94,560-atom graphene/Cu system: the
      surrogate's own forces sum to zero, the whole system's force sum
      vanishes to the mixed path's bar, two runs are bit-identical.
(C1 and C2 at full size: tests/test_gpu_fullsize.py.)

Bars: as the parity cases (lists as sets, forces 1e-10 at step 0, positions
1e-10 and velocities/forces 1e-8 after 50 steps); schedule: the same exit
step, log energies within 1e-9 relative, final positions within 1e-8
relative; batch replicas within 1e-10 of the oracle after 30 steps in FP64.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs, surrogate, system
from pfmds_amd.system import Integrator
from tests.harness import gpu_ctx, list_sets, oracle_ctx, rel_inf
from tests.test_gpu_batch import STEPS, replicas

pytestmark = pytest.mark.gpu


# ------------------------------------------------------------------------ C0

def c0():
    return configs.binary_lj_gas()  # 4,000 atoms, 60 A, seed 1000


def test_c0_step0_lists_and_forces():
    s = c0()
    assert s.n == 4000 and np.allclose(s.box, 60.0)
    g, o = gpu_ctx(s), oracle_ctx(s)
    for c in (g, o):
        c.run("nvt", 0.5, 0, 0, True)
    for k, it in enumerate(s.interactions, start=1):
        for li in range(1, len(it.lists) + 1):
            gs, gless = list_sets(g, k, li)
            os_, oless = list_sets(o, k, li)
            assert gs == os_, f"interaction {k} list {li}"
            if it.name == "lj1g":
                assert np.array_equal(gless, oless)
    assert rel_inf(g.state()[2], o.state()[2]) < 1e-10
    np.testing.assert_allclose(g.energies()["e_inter"], o.energies()["e_inter"], rtol=1e-10, atol=0)


def test_c0_trajectory_50_steps():
    s = c0()
    g, o = gpu_ctx(s), oracle_ctx(s)
    for c in (g, o):
        c.run("nvt", 0.5, 0, 0, True)
        c.run("nvt", 0.5, 1, 50, True)
    pg, vg, fg = g.state()
    po, vo, fo = o.state()
    assert rel_inf(pg, po) < 1e-10
    assert rel_inf(vg, vo) < 1e-8
    assert rel_inf(fg, fo) < 1e-8
    eg, eo = g.energies(), o.energies()
    np.testing.assert_allclose(eg["e_inter"], eo["e_inter"], rtol=1e-9)
    np.testing.assert_allclose(eg["e_nhc"], eo["e_nhc"], rtol=1e-8, atol=1e-12)


def test_c0_schedule_nvt_nve_nvms():
    """md()'s schedule over both contexts: 30 nvt steps, 30 nve steps, then
    nvms until abs(PE - PE_prev) <= ms_de."""
    s = c0()
    s.integrators = [Integrator("nvt", 0.5, 30, 10**9, 10), Integrator("nve", 0.5, 30, 10**9, 10),
                     Integrator("nvms", 0.5, 400, 10**9, 10)]
    s.md_step_limit = 460
    s.ms_de = 0.15  # the oracle exits at step 386
    out = []
    for c in (gpu_ctx(s), oracle_ctx(s)):
        recs, last = system.run_md(s, c, out_period=10)
        out.append((recs, last, c.state()))
    (rg, lg, sg), (ro, lo, so) = out
    print(f"[c0 schedule] exit step {lg} / {lo}, {len(rg)} log records", flush=True)
    assert lg == lo and lg < 460
    assert [r.step for r in rg] == [r.step for r in ro]
    for a, b in zip(rg, ro):
        assert a.integrator == b.integrator
        np.testing.assert_allclose(a.pe, b.pe, rtol=1e-9)
        np.testing.assert_allclose(a.ke, b.ke, rtol=1e-8, atol=1e-12)
    assert rel_inf(sg[0], so[0]) < 1e-8


# ------------------------------------------------------------------------ C3

def test_batch_replicas_match_oracle():
    reps = replicas()
    b = system.merge_replicas(reps)
    g = system.build_batch_context(b, capi.FP64)
    g.run("nve", 0.5, 0, STEPS, True)
    pos, vel, frc = g.state()
    er = g.replica_energies()
    for r, s in enumerate(reps):
        o = oracle_ctx(s)
        o.run("nve", 0.5, 0, STEPS, True)
        p1, v1, f1 = o.state()
        sl = b.replica_slice(r)
        print(f"[batch vs oracle] replica {r}: {s.n} atoms", flush=True)
        assert rel_inf(pos[sl], p1) < 1e-10
        assert rel_inf(vel[sl], v1) < 1e-8
        assert rel_inf(frc[sl], f1) < 1e-8
        e = o.energies()
        np.testing.assert_allclose(er["pe"][r], e["pe"], rtol=1e-9)
        np.testing.assert_allclose(er["ke"][r], e["ke"], rtol=1e-9)


def test_pair_search_compacted_staging_against_oracle():
    """Two sparse gases (under 14 atoms a cell: one wavefront per cell) beside
    a larger Cu crystal: lj1g A and lj A->B, and lj1g B with lj B->A, are
    searched in one pass (rows of A and of B, four lists), and the union of
    their candidate groups is under half of the atoms, so the pass stages
    them compacted.  Forces at step 0
    and the state after 30 steps against the oracle per replica (FP64)."""
    reps = [configs.lj_gas_lattice(600, 0.01, 31, 150.0, steps=STEPS),
            configs.lj_gas_lattice(500, 0.012, 32, 100.0, steps=STEPS),
            configs.fcc_cu_rgl(nx=7, seed=33, steps=STEPS, dt=0.5)]
    b = system.merge_replicas(reps)
    assert 2 * (reps[0].n + reps[1].n) < b.system.n
    g = system.build_batch_context(b, capi.FP64)
    g.run("nve", 0.5, 0, 0, True)
    f0 = g.state()[2]
    for r, s in enumerate(reps):
        o = oracle_ctx(s)
        o.run("nve", 0.5, 0, 0, True)
        print(f"[pair search] replica {r}: {s.n} atoms, step 0", flush=True)
        assert rel_inf(f0[b.replica_slice(r)], o.state()[2]) < 1e-10
    g.run("nve", 0.5, 1, STEPS, True)
    pos, vel, frc = g.state()
    for r, s in enumerate(reps):
        o = oracle_ctx(s)
        o.run("nve", 0.5, 0, STEPS, True)
        p1, v1, f1 = o.state()
        sl = b.replica_slice(r)
        assert rel_inf(pos[sl], p1) < 1e-10
        assert rel_inf(vel[sl], v1) < 1e-8
        assert rel_inf(frc[sl], f1) < 1e-8


def per_replica_force_sums(b, frc):
    out = []
    for r in range(b.n_rep):
        f = frc[b.replica_slice(r)]
        out.append(np.linalg.norm(f.sum(0)) / max(np.linalg.norm(f, axis=1).sum(), 1e-300))
    return np.array(out)


def test_c3_full_batch_properties():
    reps = configs.replica_mix()
    b = system.merge_replicas(reps)
    assert b.n_rep == 512 and b.system.n == 3374724
    res = []
    for _ in range(2):
        g = system.build_batch_context(b, capi.MIXED)
        g.run("nve", 0.5, 0, 20, True)
        res.append((g.state(), g.replica_energies(), g.energies()))
        g.close()
    (sa, ra, ea), (sb, rb, eb) = res
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)
    for k in ("pe", "ke"):
        assert np.array_equal(ra[k], rb[k])
    fs = per_replica_force_sums(b, sa[2])
    print(f"[c3 full] {b.n_rep} replicas, largest relative force sum {fs.max():.2e}", flush=True)
    # mixed precision: the RGL rows sum each block of four FP32 pair terms in
    # FP32, so the two sides of a pair do not cancel exactly (the LJ rows do)
    assert fs.max() < 2e-8
    np.testing.assert_allclose(ra["pe"].sum(), ea["pe"], rtol=1e-12)
    np.testing.assert_allclose(ra["ke"].sum(), ea["ke"], rtol=1e-12)


# ------------------------------------------------------------------------ C4

def test_c4_full_surrogate_properties():
    s = configs.graphene_on_cu()
    assert s.n == 94560
    _, sur_s, inter = surrogate.split_system(s)
    tr = surrogate.training_system(sur_s)
    n_c = int((s.mass < 13).sum())
    # a network with the system's own feature scaling (untrained: the
    # properties hold for any parameters)
    probe = system.build_context(tr, lambda bx, p, v, m: capi.Context.gpu(bx, p, v, m, capi.MIXED))
    probe.sur_set_params(1, surrogate.init_theta(np.zeros(40), np.ones(40), 0.0))
    probe.run("nve", 0.5, 0, 0, False)
    G = probe.sur_descriptors(1, n_c)
    theta = surrogate.init_theta(G.mean(0), 1.0 / (G.std(0) + 1e-3), -7.0, seed=5)
    probe.sur_set_params(1, theta)
    probe.run("nve", 0.5, 0, 0, True)
    f = probe.state()[2]
    rel = np.linalg.norm(f.sum(0)) / np.linalg.norm(f, axis=1).sum()
    print(f"[c4 full] surrogate-only force sum {rel:.2e}", flush=True)
    assert rel < 1e-10
    probe.close()
    res = []
    for _ in range(2):
        c = surrogate.surrogate_context(sur_s, theta)
        c.run("nvt", 0.5, 0, 25, True)
        res.append((c.state(), c.energies()))
        c.close()
    (sa, ea), (sb, eb) = res
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)
    assert ea["pe"] == eb["pe"]
    f = sa[2]
    rel = np.linalg.norm(f.sum(0)) / np.linalg.norm(f, axis=1).sum()
    print(f"[c4 full] whole-system force sum {rel:.2e}", flush=True)
    assert rel < 1e-6
