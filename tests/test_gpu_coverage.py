"""HIP path against the CPU oracle on the paths the basic parity cases leave
out: ljc (both `simplified` modes), morsec with `simplified` = T, lj1g with a
switching width R2-R1 = 2 A (where f_dfr_cut's missing 1/(R2-R1) changes the
force) and in a two-species group whose order is not ascending (lessnnum
stops at the first larger atom index), z-moving groups under nve and nvms
(integrate_verlet_z_*, molecular_static_1D), the reflecting wall
(invert_z_velocities), and the status codes the reference stops with.

Bar (as the basic parity cases): lists identical as sorted sets and lessnnum
identical; FP64 forces and per-interaction energies within 1e-10 relative at
step 0; 25-step trajectories with positions within 1e-10 and velocities and
forces within 1e-8; 300 free-flight steps against the wall within 1e-12;
status codes, atom indices and steps identical to the oracle's.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs
from pfmds_amd.system import Interaction
from tests.harness import gpu_ctx, list_sets, oracle_ctx, rel_inf
from tests.test_oracle_guards import err_parts, free_columns

pytestmark = pytest.mark.gpu


def graphene_metal(kind, simplified, seed=61):
    """A small graphene/Cu(111) slab with tb and ljc or morsec (settings order
    rjl, tb, then the graphene-metal term)."""
    s = configs.graphene_on_cu(cx=12, cy=7, gx=12, gy=7, layers=4, seed=seed, jitter=0.1)
    mor = s.interactions[2]
    if kind == "ljc":
        params = [0.005, 3.2, 2.0, 6.0, 7.0, 1.0 if simplified else 0.0]
    else:
        params = list(configs.MORSEC_CCU[:6]) + [1.0 if simplified else 0.0]
    s.interactions[2] = Interaction(kind, params, mor.lists)
    return s


def lj1g_wide_switch(order):
    """Binary gas with one lj1g over both species, R1 = 5.5, R2 = 7.5; the
    group lists B before A (order 'BA'), so its indexes do not ascend."""
    s = configs.binary_lj_gas(n=600, box=32.0, seed=62, r1=5.5, r2=7.5)
    s.groups = [["A"], ["B"], ["A", "B"] if order == "AB" else ["B", "A"], ["#"]]
    s.interactions = [Interaction("lj1g", [0.012, 3.5, 5.5, 7.5], [(3, 3, 160, 8.5, 20)])]
    return s


def z_moving_slab(seed=63):
    """Graphene/Cu where the graphene moves in xyz and the mobile Cu along z
    only (xyz_moving = C, z_moving = Cu)."""
    s = configs.graphene_on_cu(cx=12, cy=7, gx=12, gy=7, layers=4, seed=seed, temperature=200.0)
    s.xyz_moving, s.z_moving = 1, 5
    return s


CASES = {
    "ljc": lambda: graphene_metal("ljc", False),
    "ljc_simplified": lambda: graphene_metal("ljc", True),
    "morsec_simplified": lambda: graphene_metal("morsec", True),
    "lj1g_wide_BA": lambda: lj1g_wide_switch("BA"),
    "lj1g_wide_AB": lambda: lj1g_wide_switch("AB"),
    "z_moving": z_moving_slab,
}


def step0_pair(s, integ="nve"):
    g, o = gpu_ctx(s), oracle_ctx(s)
    for c in (g, o):
        c.run(integ, 0.5, 0, 0, True)
    return g, o


@pytest.mark.parametrize("name", list(CASES))
def test_step0_lists_forces_energies(name):
    s = CASES[name]()
    print(f"[coverage step0] {name} n={s.n}", flush=True)
    g, o = step0_pair(s)
    for k, it in enumerate(s.interactions, start=1):
        for li in range(1, len(it.lists) + 1):
            gs, gless = list_sets(g, k, li)
            os_, oless = list_sets(o, k, li)
            assert gs == os_, f"{name} interaction {k} list {li}"
            if it.name == "lj1g":
                assert np.array_equal(gless, oless), f"{name} lessnnum"
    fg, fo = g.state()[2], o.state()[2]
    assert rel_inf(fg, fo) < 1e-10
    np.testing.assert_allclose(g.energies()["e_inter"], o.energies()["e_inter"], rtol=1e-10, atol=0)
    if name.startswith("ljc") or name.startswith("morsec"):
        n_c = int((s.mass < 13).sum())
        np.testing.assert_allclose(g.normals(3, n_c), o.normals(3, n_c), atol=1e-12)


@pytest.mark.parametrize("name,integ", [("ljc", "nvt"), ("ljc_simplified", "nve"), ("morsec_simplified", "nvt"),
                                        ("lj1g_wide_BA", "nvt"), ("z_moving", "nve"), ("z_moving", "nvms"),
                                        ("z_moving", "nvt")])
def test_trajectory(name, integ):
    s = CASES[name]()
    print(f"[coverage traj] {name} {integ} n={s.n}", flush=True)
    g, o = gpu_ctx(s), oracle_ctx(s)
    for c in (g, o):
        c.run(integ, 0.5, 0, 0, True, 1e-12)
        c.run(integ, 0.5, 1, 25, True, 1e-12)
    pg, vg, fg = g.state()
    po, vo, fo = o.state()
    assert rel_inf(pg, po) < 1e-10
    assert rel_inf(vg, vo) < 1e-8
    assert rel_inf(fg, fo) < 1e-8
    np.testing.assert_allclose(g.energies()["e_inter"], o.energies()["e_inter"], rtol=1e-9)
    if name == "z_moving":
        # the z-moving Cu kept x, y; its z moved
        cu = np.array([sp == "Cu" for sp in s.species])
        assert np.array_equal(pg[cu, :2], s.pos[cu, :2])
        assert np.abs(pg[cu, 2] - s.pos[cu, 2]).max() > 1e-4


@pytest.mark.parametrize("precision,tol", [(capi.MIXED, 1e-4)])
def test_mixed_rmse_ljc(precision, tol):
    s = CASES["ljc"]()
    g64, g32 = gpu_ctx(s, capi.FP64), gpu_ctx(s, precision)
    for c in (g64, g32):
        c.run("nvt", 0.5, 0, 0, True)
    rmse = float(np.sqrt(np.mean((g32.state()[2] - g64.state()[2]) ** 2)))
    assert rmse <= tol


# ------------------------------------------------------------------ the wall

def test_reflecting_wall_free_flight():
    s = free_columns(n=64, vmax=0.6, seed=9)
    s.invert_z = True
    g, o = gpu_ctx(s), oracle_ctx(s)
    for c in (g, o):
        c.run("nve", 0.5, 0, 0, True)
        c.run("nve", 0.5, 1, 150, True)
        c.run("nve", 0.5, 151, 300, True)
    pg, vg, _ = g.state()
    po, vo, _ = o.state()
    assert rel_inf(pg, po) < 1e-12
    assert np.array_equal(np.sign(vg), np.sign(vo))
    assert rel_inf(vg, vo) < 1e-12


def test_reflecting_wall_gas_trajectory():
    """A hot LJ gas whose band holds atoms: the wall flips some v_z."""
    s = configs.binary_lj_gas(n=600, box=32.0, seed=64, temperature=600.0)
    s.invert_z = True
    g, o = gpu_ctx(s), oracle_ctx(s)
    for c in (g, o):
        c.run("nve", 1.0, 0, 0, True)
        c.run("nve", 1.0, 1, 30, True)
    pg, vg, fg = g.state()
    po, vo, fo = o.state()
    assert rel_inf(pg, po) < 1e-10 and rel_inf(vg, vo) < 1e-8 and rel_inf(fg, fo) < 1e-8


def test_nvms_exit_step_runs_the_wall():
    """md() applies check_positions and the wall at the top of the step whose
    exit test ends an nvms run: the final velocities carry it."""
    s = configs.binary_lj_gas(n=600, box=32.0, seed=65, temperature=100.0)
    s.invert_z = True
    res = []
    for c in (gpu_ctx(s), oracle_ctx(s)):
        c.run("nvms", 1.0, 0, 0, True, 1e-2)
        last, ev = c.run("nvms", 1.0, 1, 400, True, 1e-2)  # the oracle exits at step 297
        res.append((last, ev, c.state()))
    (lg, eg, sg), (lo, eo, so) = res
    assert eg == eo == capi.EV_CONVERGED and lg == lo
    # 297 steps of descent: rounding grows a little; a missing wall at the
    # exit step would flip the sign of v_z in the band (an O(1) difference)
    assert rel_inf(sg[0], so[0]) < 1e-8 and rel_inf(sg[1], so[1]) < 1e-6


# -------------------------------------------------------------- status codes

def both(s):
    return gpu_ctx(s), oracle_ctx(s)


def raised(fn, *args):
    with pytest.raises(capi.PfmError) as e:
        fn(*args)
    return e.value


@pytest.mark.parametrize("value,bad", [("L+0.9e-7", False), ("L+1.1e-7", True), ("-1.1e-7", True)])
def test_out_of_cell_at_start(value, bad):
    s = free_columns(n=9)
    L = s.box[0]
    s.pos[4, 0] = {"L+0.9e-7": L + 0.9e-7, "L+1.1e-7": L + 1.1e-7, "-1.1e-7": -1.1e-7}[value]
    g, o = both(s)
    if not bad:
        for c in (g, o):
            c.run("nve", 1.0, 0, 0, True)
        return
    eg, eo = raised(g.run, "nve", 1.0, 0, 0, True), raised(o.run, "nve", 1.0, 0, 0, True)
    assert eg.code == eo.code == capi.E_OUT_OF_CELL
    pg, po = err_parts(eg.msg), err_parts(eo.msg)
    assert pg[0] == po[0] == 5 and pg[4] == po[4] == 0
    np.testing.assert_allclose(pg[1:4], po[1:4], rtol=1e-9)


@pytest.mark.parametrize("split", [False, True])
def test_out_of_cell_after_drift(split):
    """A drift overshooting the box by more than one period is refused at the
    top of the next step (step 2), within one call or when the overshoot is
    the last drift of a call and the next call starts there."""
    s = free_columns(n=9)
    Lz, dt = s.box[2], 0.5
    s.pos[3, 2] = 0.5 * Lz
    s.vel[3, 2] += 1.7 * Lz / dt / (1.0 - s.mass[3] / s.mass.sum())
    out = []
    for c in both(s):
        c.run("nve", dt, 0, 0, True)
        if split:
            c.run("nve", dt, 1, 1, True)
            out.append(raised(c.run, "nve", dt, 2, 5, True))
        else:
            out.append(raised(c.run, "nve", dt, 1, 5, True))
    eg, eo = out
    assert eg.code == eo.code == capi.E_OUT_OF_CELL
    pg, po = err_parts(eg.msg), err_parts(eo.msg)
    assert pg[0] == po[0] == 4 and pg[4] == po[4] == 2
    np.testing.assert_allclose(pg[1:4], po[1:4], rtol=1e-9)


@pytest.mark.parametrize("what", ["empty_group", "negative_q1", "negative_temperature"])
def test_wrong_nhc_parameters(what):
    s = free_columns(n=9)
    group, T, q1 = 1, 300.0, 1e4
    if what == "empty_group":
        group = 2  # '#'
    elif what == "negative_q1":
        q1 = -1.0
    else:
        T = -5.0
    for c in both(s):
        e = raised(c.add_nhc, group, T, 3, q1)
        assert e.code == capi.E_NHC_PARAMS and "wrong nhc parameters" in e.msg


def test_unknown_interaction_kind():
    s = free_columns(n=9)
    for c in both(s):
        gn = np.array([1, 1], np.int32)
        nm = np.array([8], np.int32)
        rc = np.array([2.0])
        per = np.array([1], np.int32)
        p = np.zeros(4)
        fn = c._fn("add_interaction")
        rc_ = fn(c.h, 99, capi._dp(p), 4, 1, capi._ip(gn), capi._ip(nm), capi._dp(rc), capi._ip(per))
        assert rc_ == capi.E_UNKNOWN_KIND
        assert c.last_error() == "error: unknown interaction name"
