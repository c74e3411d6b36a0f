"""The HIP path at the benchmark sizes, through properties that do not depend
on the system size: fcc Cu RGL 32^3 cells (131,072 atoms) and graphene on
Cu(111) (94,560 atoms).

Bar: the perfect crystal has the CPU oracle's energy per atom (taken on a
4^3-cell crystal, same lattice) within 1e-10 relative and zero force; the sum
of all forces vanishes (Newton's third law of every pair, the full-list
gathers form each pair force twice with dr and -dr): |sum F| within 1e-12 of
sum |F_i| for RGL in FP64 (both sides of a pair run the same formula on
exactly opposite dr) and 1e-9 in mixed precision (each row sums every block
of four FP32 pair terms in FP32 before its FP64 sum, so the two sides of a
pair are rounded in different blocks; measured 6e-11), 1e-10 for graphene/Cu in FP64 and
1e-6 in mixed precision (the two sides of a Tersoff-Brenner bond or of a
graphene-metal pair come from different passes, whose FP32 roundings of large
cancelling terms do not cancel); two runs of the same
steps are bit-identical (deterministic binning and reductions); the mixed
path's force RMSE against FP64 stays within 1e-4 eV/A.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs
from tests.harness import gpu_ctx, oracle_ctx

pytestmark = pytest.mark.gpu


def test_perfect_fcc_energy_per_atom_matches_small_crystal():
    big = configs.fcc_cu_rgl(nx=32, sigma=0.0, temperature=0.0)
    small = configs.fcc_cu_rgl(nx=4, sigma=0.0, temperature=0.0)
    print(f"[fcc] {big.n} atoms vs oracle on {small.n}", flush=True)
    g = gpu_ctx(big)
    g.run("nve", 1.0, 0, 0, True)
    o = oracle_ctx(small)
    o.run("nve", 1.0, 0, 0, True)
    e_big = g.energies()["pe"] / big.n
    e_small = o.energies()["pe"] / small.n
    np.testing.assert_allclose(e_big, e_small, rtol=1e-10)
    f = g.state()[2]
    assert np.abs(f).max() < 1e-9


@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-12), (capi.MIXED, 1e-9)])
def test_momentum_conservation_fcc(precision, tol):
    s = configs.fcc_cu_rgl(nx=32)
    g = gpu_ctx(s, precision)
    g.run("nve", 1.0, 0, 0, True)
    f = g.state()[2]
    scale = np.linalg.norm(f, axis=1).sum()
    print(f"  sum {f.sum(0)} relative {np.linalg.norm(f.sum(0)) / scale:.2e}", flush=True)
    assert np.linalg.norm(f.sum(0)) < tol * scale, f.sum(0)


@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-10), (capi.MIXED, 1e-6)])
def test_momentum_conservation_graphene_cu(precision, tol):
    s = configs.graphene_on_cu()
    g = gpu_ctx(s, precision)
    g.run("nvt", 0.5, 0, 0, True)
    f = g.state()[2]
    scale = np.linalg.norm(f, axis=1).sum()
    print(f"  sum {f.sum(0)} relative {np.linalg.norm(f.sum(0)) / scale:.2e}", flush=True)
    assert np.linalg.norm(f.sum(0)) < tol * scale, f.sum(0)


@pytest.mark.parametrize("name", ["fcc", "graphene_cu"])
def test_runs_are_bit_identical(name):
    if name == "fcc":
        s, integ, dt = configs.fcc_cu_rgl(nx=32), "nve", 1.0
    else:
        s, integ, dt = configs.graphene_on_cu(), "nvt", 0.5
    out = []
    for _ in range(2):
        g = gpu_ctx(s, capi.MIXED)
        g.run(integ, dt, 0, 25, True)
        out.append(g.state())
        g.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_mixed_rmse_graphene_cu_full():
    s = configs.graphene_on_cu()
    g64, g32 = gpu_ctx(s, capi.FP64), gpu_ctx(s, capi.MIXED)
    for g in (g64, g32):
        g.run("nvt", 0.5, 0, 0, True)
    rmse = float(np.sqrt(np.mean((g32.state()[2] - g64.state()[2]) ** 2)))
    print(f"  rmse {rmse:.3e}", flush=True)
    assert rmse <= 1e-4
