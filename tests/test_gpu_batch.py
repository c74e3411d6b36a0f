"""Replica batches (pfm_create_batch) against the same systems run one at a
time on the GPU, whose single-system path the parity tests pin to the CPU
oracle.

Bar: per replica, positions, velocities and forces after 30 nve steps within
1e-10 (FP64) and 1e-6 (mixed) relative of the single run (the replicas differ
from single runs only in the summation order of the step-0 zero_momentum);
per-replica energies within 1e-10 relative in FP64 and 1e-6 in mixed precision
(the FP32 pair terms of positions that differ in their last bits can round
apart by an FP32 ulp, and the mixed path sums them in FP32 blocks of four);
batches refuse Nose-Hoover chains.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs, system
from tests.harness import gpu_ctx, rel_inf

pytestmark = pytest.mark.gpu

STEPS = 30


def replicas():
    return [
        configs.lj_gas_lattice(400, 0.02, 11, 150.0, steps=STEPS),        # 27 A box: cell tiles
        configs.fcc_cu_rgl(nx=4, seed=12, steps=STEPS, dt=0.5),           # 14.5 A box: row search
        configs.fcc_cu_rgl(nx=6, seed=13, steps=STEPS, dt=0.5, temperature=900.0),  # cell tiles
        configs.lj_gas_lattice(300, 0.03, 14, 60.0, steps=STEPS),         # 21.5 A box: row search
        configs.fcc_cu_rgl(nx=5, ny=4, nz=7, seed=15, steps=STEPS, dt=0.5),
    ]


@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-10), (capi.MIXED, 1e-6)])
def test_batch_matches_single_runs(precision, tol):
    reps = replicas()
    b = system.merge_replicas(reps)
    print(f"[batch] {b.n_rep} replicas, {b.system.n} atoms, precision {precision}", flush=True)
    g = system.build_batch_context(b, precision)
    g.run("nve", 0.5, 0, STEPS, True)
    pos, vel, frc = g.state()
    er = g.replica_energies()
    for r, s in enumerate(reps):
        c = gpu_ctx(s, precision)
        c.run("nve", 0.5, 0, STEPS, True)
        p1, v1, f1 = c.state()
        sl = b.replica_slice(r)
        assert rel_inf(pos[sl], p1) < tol, f"replica {r} positions"
        assert rel_inf(vel[sl], v1) < tol, f"replica {r} velocities"
        assert rel_inf(frc[sl], f1) < tol, f"replica {r} forces"
        e = c.energies()
        etol = 1e-10 if precision == capi.FP64 else tol
        np.testing.assert_allclose(er["pe"][r], e["pe"], rtol=etol)
        np.testing.assert_allclose(er["ke"][r], e["ke"], rtol=etol)
        np.testing.assert_allclose(er["temp"][r], e["temp"], rtol=etol)
    tot = g.energies()
    np.testing.assert_allclose(tot["pe"], er["pe"].sum(), rtol=1e-10)


def test_batch_zero_momentum_per_replica():
    reps = replicas()[:3]
    for k, s in enumerate(reps):
        s.zero_momentum_period = 1
        s.vel = s.vel + np.array([0.001 * (k + 1), -0.002, 0.0005])  # every replica drifts differently
    b = system.merge_replicas(reps)
    g = system.build_batch_context(b, capi.FP64)
    g.run("nve", 0.5, 0, 3, True)
    _, vel, _ = g.state()
    for r in range(b.n_rep):
        sl = b.replica_slice(r)
        p = (b.system.mass[sl][:, None] * vel[sl]).sum(0)
        scale = (b.system.mass[sl][:, None] * np.abs(vel[sl])).sum(0)
        assert np.all(np.abs(p) < 1e-12 * scale), f"replica {r} momentum {p}"


def test_batch_refuses_chains():
    reps = replicas()[:2]
    b = system.merge_replicas(reps)
    g = system.build_batch_context(b, capi.FP64)
    g.add_nhc(b.system.all_moving, 300.0, 3, 1e5)
    with pytest.raises(capi.PfmError) as ei:
        g.run("nvt", 0.5, 0, 1, True)
    assert ei.value.code == capi.E_INVALID
