"""rebosc on the GPU against the CPU oracle (the reference's energy and its
central differences over truncated lists).

Bars: lists identical as sorted sets; energies within 1e-10 relative; forces
within 1e-6 of the largest force (both sides are central differences with
dx = 1e-6 A, each within a few 1e-9 of -grad E); zero force on the perfect
sheet to 1e-6 eV/A; the force sum zero to 1e-7 of sum |F|; a 25-step nve
trajectory within 1e-8; graphene on Cu with rebosc feeding morsec's
nearest-neighbour list.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs, system
from tests.harness import gpu_ctx, list_sets, oracle_ctx, rel_inf

pytestmark = pytest.mark.gpu


def step0(ctx, s):
    ctx.run(s.integrators[0].name, s.integrators[0].dt, 0, 0, True)
    return ctx


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_rebo_perfect_and_buckled_sheets(precision):
    for s in (configs.graphene_rebo(gx=6, gy=4), configs.graphene_rebo(gx=7, gy=4, jitter=0.1, buckle=0.2, seed=73)):
        print(f"[rebo] n={s.n} precision={precision}", flush=True)
        g, o = step0(gpu_ctx(s, precision), s), step0(oracle_ctx(s), s)
        assert list_sets(g, 1, 1)[0] == list_sets(o, 1, 1)[0]
        eg, eo = g.energies()["pe"], o.energies()["pe"]
        assert abs(eg - eo) < 1e-10 * abs(eo)
        fg, fo = g.state()[2], o.state()[2]
        fmax = np.abs(fo).max()
        print(f"  E {eg:.12f} vs {eo:.12f}; |dF| {np.abs(fg - fo).max():.2e} of max {fmax:.2e}", flush=True)
        if fmax < 1e-6:
            assert np.abs(fg).max() < 1e-6
        else:
            assert np.abs(fg - fo).max() < 1e-6 * fmax
            assert np.linalg.norm(fg.sum(0)) < 1e-7 * np.linalg.norm(fg, axis=1).sum()


def test_rebo_trajectory():
    s = configs.graphene_rebo(gx=6, gy=4, jitter=0.05, buckle=0.1, seed=74, temperature=600.0)
    g, o = gpu_ctx(s), oracle_ctx(s)
    for c in (g, o):
        c.run("nve", 0.5, 0, 25, True)
    (pg, vg, fg), (po, vo, fo) = g.state(), o.state()
    print(f"[rebo traj] pos {rel_inf(pg, po):.2e} vel {rel_inf(vg, vo):.2e} frc {rel_inf(fg, fo):.2e}", flush=True)
    assert rel_inf(pg, po) < 1e-8 and rel_inf(vg, vo) < 1e-6 and rel_inf(fg, fo) < 1e-6


def test_graphene_cu_with_rebo_nearest_neighbours():
    s = configs.graphene_on_cu(cx=8, cy=5, gx=8, gy=5, layers=3, seed=75, jitter=0.1)
    tb = [it.name for it in s.interactions].index("tb")
    s.interactions[tb] = system.Interaction("rebosc", list(configs.REBO_C), [(1, 1, 8, 2.3, 10)])
    print(f"[rebo gr/cu] n={s.n}", flush=True)
    g, o = gpu_ctx(s), oracle_ctx(s)
    for c in (g, o):
        c.run("nvt", 0.5, 0, 0, True)
    assert abs(g.energies()["pe"] - o.energies()["pe"]) < 1e-10 * abs(o.energies()["pe"])
    fg, fo = g.state()[2], o.state()[2]
    assert np.abs(fg - fo).max() < 1e-6 * np.abs(fo).max()
    # morsec's nearest-neighbour list comes from the rebosc list
    assert list_sets(g, 3, 3)[0] == list_sets(o, 3, 3)[0]
