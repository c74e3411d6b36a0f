"""HIP path against the CPU oracle on the same seeded inputs (C-ABI calls).

Bar: neighbour lists identical as sorted sets (bit-exact membership test);
FP64 forces and per-interaction energies within 1e-10 relative (inf-norm of
the force difference over the inf-norm of the forces); short trajectories
within 1e-8; mixed precision force RMSE <= 1e-4 eV/A against FP64.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs
from tests.harness import gpu_ctx, list_sets, oracle_ctx, rel_inf

pytestmark = pytest.mark.gpu


def small_systems():
    return {
        "lj_gas": configs.binary_lj_gas(n=600, box=32.0, seed=11),
        "cu_rgl": configs.fcc_cu_rgl(nx=6, seed=12),
        "gr_cu": configs.graphene_on_cu(cx=12, cy=7, gx=12, gy=7, layers=4, seed=13),
    }


SYSTEMS = small_systems()


def _step0(s, precision=capi.FP64):
    g = gpu_ctx(s, precision)
    o = oracle_ctx(s)
    integ = s.integrators[0].name
    g.run(integ, s.integrators[0].dt, 0, 0, True, s.ms_de)
    o.run(integ, s.integrators[0].dt, 0, 0, True, s.ms_de)
    return g, o


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_lists_bit_exact(name):
    s = SYSTEMS[name]
    print(f"[lists] {name} n={s.n}", flush=True)
    g, o = _step0(s)
    for k, it in enumerate(s.interactions, start=1):
        for li in range(1, len(it.lists) + 1):
            gs, gless = list_sets(g, k, li)
            os_, oless = list_sets(o, k, li)
            assert gs == os_, f"{name} interaction {k} list {li} differs"
            if it.name == "lj1g":
                assert np.array_equal(gless, oless)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_forces_energies_fp64(name):
    s = SYSTEMS[name]
    print(f"[fp64] {name} n={s.n}", flush=True)
    g, o = _step0(s)
    _, _, fg = g.state()
    _, _, fo = o.state()
    assert rel_inf(fg, fo) < 1e-10
    eg, eo = g.energies(), o.energies()
    np.testing.assert_allclose(eg["e_inter"], eo["e_inter"], rtol=1e-10, atol=0)
    assert abs(eg["ke"] - eo["ke"]) <= 1e-12 * abs(eo["ke"])


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_mixed_force_rmse(name):
    s = SYSTEMS[name]
    print(f"[mixed] {name} n={s.n}", flush=True)
    g64 = gpu_ctx(s, capi.FP64)
    g32 = gpu_ctx(s, capi.MIXED)
    for c in (g64, g32):
        c.run(s.integrators[0].name, s.integrators[0].dt, 0, 0, True)
    f64, f32 = g64.state()[2], g32.state()[2]
    rmse = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    print(f"  force rmse {rmse:.3e} eV/A", flush=True)
    assert rmse <= 1e-4
    e64, e32 = g64.energies()["pe"], g32.energies()["pe"]
    assert abs(e64 - e32) <= 1e-5 * abs(e64)


@pytest.mark.parametrize("name,integ", [("lj_gas", "nvt"), ("cu_rgl", "nve"), ("gr_cu", "nvt"), ("lj_gas", "nvms")])
def test_short_trajectory(name, integ):
    s = SYSTEMS[name]
    print(f"[traj] {name} {integ}", flush=True)
    g = gpu_ctx(s)
    o = oracle_ctx(s)
    dt = 0.5
    for c in (g, o):
        c.run(integ, dt, 0, 0, True, 1e-12)
        c.run(integ, dt, 1, 25, True, 1e-12)
    pg, vg, fg = g.state()
    po, vo, fo = o.state()
    assert rel_inf(pg, po) < 1e-10
    assert rel_inf(vg, vo) < 1e-8
    assert rel_inf(fg, fo) < 1e-8
    eg, eo = g.energies(), o.energies()
    np.testing.assert_allclose(eg["e_inter"], eo["e_inter"], rtol=1e-9)
    np.testing.assert_allclose(eg["e_nhc"], eo["e_nhc"], rtol=1e-8, atol=1e-12)
