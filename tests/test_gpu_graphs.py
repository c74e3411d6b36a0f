"""Captured blocks of steps (HIP graphs in pfm_run) against the same steps
launched one by one.

Bar: positions, velocities, forces and energies bit-identical after every
chunk.  The chunks end on energy steps (eager) and span several replays of
each captured block, including the re-sort that swaps the atom buffers every
list period (two cached graphs alternate), NHC chains, lists with different
periods and replica batches.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs, system
from tests.harness import gpu_ctx

pytestmark = pytest.mark.gpu

CHUNKS = [(0, 0), (1, 47), (48, 95), (96, 130)]


def run_chunks(make, integ, dt, graphs, min_launches=0):
    c = make()
    c.set_graphs(graphs)
    out = []
    for s0, s1 in CHUNKS:
        c.run(integ, dt, s0, s1, True)
        e = c.energies()
        out.append((c.state(), e["pe"], e["ke"]))
    cap, launched = c.graph_count()
    print(f"  graphs={graphs}: {cap} captured, {launched} launched", flush=True)
    assert launched >= min_launches if graphs else launched == 0
    c.close()
    return out


def assert_same(a, b):
    for (sa, pa, ka), (sb, pb, kb) in zip(a, b):
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y)
        assert pa == pb and ka == kb


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_graphs_fcc_nve(precision):
    s = configs.fcc_cu_rgl(nx=8, seed=61)
    print(f"[graphs fcc] n={s.n} precision={precision}", flush=True)
    mk = lambda: gpu_ctx(s, precision)  # noqa: E731
    assert_same(run_chunks(mk, "nve", 1.0, True, 8), run_chunks(mk, "nve", 1.0, False))


def test_graphs_graphene_cu_nvt():
    s = configs.graphene_on_cu(cx=12, cy=7, gx=12, gy=7, layers=4, seed=62)
    print(f"[graphs gr/cu] n={s.n}", flush=True)
    mk = lambda: gpu_ctx(s, capi.MIXED)  # noqa: E731
    assert_same(run_chunks(mk, "nvt", 0.5, True, 8), run_chunks(mk, "nvt", 0.5, False))


def test_graphs_mixed_list_periods():
    # periods 5 and 10: a block of 10 steps holds one full re-sort and one
    # re-binning without permutation
    s = configs.lj_gas_lattice(1000, 0.02, 63, 300.0, dt=1.0, steps=130)
    s.interactions[0].lists = [(1, 1, 160, 8.5, 10)]
    s.interactions[1].lists = [(2, 2, 160, 8.5, 5)]
    s.interactions[2].lists = [(1, 2, 160, 8.5, 10), (2, 1, 160, 8.5, 10)]
    print(f"[graphs periods] n={s.n}", flush=True)
    mk = lambda: gpu_ctx(s, capi.MIXED)  # noqa: E731
    assert_same(run_chunks(mk, "nve", 1.0, True, 8), run_chunks(mk, "nve", 1.0, False))


def test_graphs_replica_batch():
    reps = configs.replica_mix(n_rep=6, seed=64)
    b = system.merge_replicas(reps)
    print(f"[graphs batch] replicas={b.n_rep} n={b.system.n}", flush=True)
    mk = lambda: system.build_batch_context(b, capi.MIXED)  # noqa: E731
    assert_same(run_chunks(mk, "nve", 0.5, True, 3), run_chunks(mk, "nve", 0.5, False))
