"""The energy reductions are compensated sums (Neumaier): an interaction's
energy equals math.fsum of its per-row energies (pfm_row_energies) to within
the compensated sum's error bound, on rows built to cancel almost exactly.

The system: isolated A-B pairs of a deep Lennard-Jones interaction (eps =
1000 eV, sigma = 2 A), half of them in the attractive well (-1000..-460 eV)
and half pressed into the repulsive wall at the distance whose energy is the
opposite of a partner's, so the rows sum to about 1e-10 of their absolute
sum.  A plain summation tree loses about 1e-16 log2(n) of the absolute sum
(1e-9 eV here), more than the whole result; the compensated sum keeps
2.2e-16 |sum| + 2 n eps^2 sum|x|.  Checked for a single system and per
replica of a batch.
"""
import math

import numpy as np
import pytest

from pfmds_amd import capi, system
from pfmds_amd.system import Integrator, Interaction, System
from tests.harness import gpu_ctx

pytestmark = pytest.mark.gpu

EPS, SIG = 1000.0, 2.0


def lj(r):
    x6 = (SIG / r) ** 6
    return 4 * EPS * (x6 * x6 - x6)


def cancelling_pairs(m=14, seed=3):
    """m^3 A-B pairs on a 10 A grid; even pairs in the well, odd ones at the
    repulsive distance of the opposite energy (bisection in float64)."""
    rng = np.random.default_rng(seed)
    n = m ** 3
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3) * 10.0 + 1.0
    r = np.empty(n)
    for k in range(0, n, 2):
        r[k] = rng.uniform(2.25, 2.8)
        if k + 1 < n:
            target = -lj(r[k])
            lo, hi = 1.7, SIG
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                lo, hi = (mid, hi) if lj(mid) > target else (lo, mid)
            r[k + 1] = 0.5 * (lo + hi)
    pos = np.concatenate([g, g + np.column_stack([r, np.zeros(n), np.zeros(n)])])
    box = np.full(3, 10.0 * m)
    mass = np.full(2 * n, 40.0)
    return System(box=box, pos=pos, vel=np.zeros((2 * n, 3)), mass=mass, species=["A"] * n + ["B"] * n,
                  groups=[["A"], ["B"], ["A", "B"], ["#"]],
                  interactions=[Interaction("lj", [EPS, SIG, 4.0, 5.0], [(1, 2, 4, 5.5, 1), (2, 1, 4, 5.5, 1)])],
                  integrators=[Integrator("nve", 0.5, 1, 10**9, 10**9)], all_moving=3, xyz_moving=3, z_moving=4,
                  all_atoms=3)


def bound(rows):
    """Neumaier's error bound for the tree of (s, c) pairs, with margin."""
    eps = 2.0 ** -53
    return 4 * eps * abs(math.fsum(rows)) + 4 * len(rows) * eps * eps * np.abs(rows).sum()


def naive(rows):
    s = 0.0
    for x in rows:
        s += x
    return s


def test_energy_is_a_compensated_sum():
    s = cancelling_pairs()
    g = gpu_ctx(s, capi.FP64)
    g.run("nve", 0.5, 0, 0, True)
    rows = g.row_energies(1)
    e = g.energies()["e_inter"][0]
    exact = math.fsum(rows)
    print(f"[compensated] rows {len(rows)}: sum|x| {np.abs(rows).sum():.3e}, fsum {exact:.3e}, "
          f"gpu {e:.3e}, sequential {naive(rows):.3e}", flush=True)
    assert np.abs(rows).sum() > 1e9 * abs(exact)  # the rows cancel
    assert abs(e - exact) <= bound(rows)
    # a plain sum of the same rows misses by far more than the bound
    assert abs(naive(rows) - exact) > 100 * bound(rows)


def test_replica_energies_are_compensated_sums():
    reps = [cancelling_pairs(m=10, seed=5), cancelling_pairs(m=12, seed=6)]
    b = system.merge_replicas(reps)
    g = system.build_batch_context(b, capi.FP64)
    g.run("nve", 0.5, 0, 0, True)
    rows = g.row_energies(1)
    per = g.replica_energies()["pe"]
    n0 = reps[0].n // 2
    for r, part in enumerate((rows[:n0], rows[n0:])):
        assert abs(per[r] - math.fsum(part)) <= bound(part)
    assert abs(g.energies()["e_inter"][0] - math.fsum(rows)) <= bound(rows)
