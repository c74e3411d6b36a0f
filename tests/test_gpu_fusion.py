"""Co-scheduled passes (pfm_set_fusion) against the same passes launched one
by one, and against the CPU oracle: graphene on a metal, and lj / lj1g /
rjl passes over disjoint atoms (a binary gas, a batch of replicas).

The fused launches run the same row arithmetic; only the order of the FP64
force sums of a graphene atom differs (its graphene-pass, tb and
normal-derivative contributions are added in another order).  Bar: FP64
forces within 1e-13 relative (inf-norm) of the separate launches and
energies within 1e-13; mixed precision within 1e-7 relative; the fused FP64
path against the oracle as the parity cases (forces 1e-10 at step 0,
positions 1e-10 after 25 nvt steps); morsec and ljc, `simplified` F and T,
lane counts 4, 8 and 16, captured blocks and direct launches.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs
from pfmds_amd.system import Interaction
from tests.harness import gpu_ctx, oracle_ctx, rel_inf

pytestmark = pytest.mark.gpu


def slab(kind="morsec", simplified=False, seed=71):
    s = configs.graphene_on_cu(cx=14, cy=8, gx=14, gy=8, layers=4, seed=seed, jitter=0.1)
    if kind == "ljc" or simplified:
        mor = s.interactions[2]
        if kind == "ljc":
            params = [0.005, 3.2, 2.0, 6.0, 7.0, 1.0 if simplified else 0.0]
        else:
            params = list(configs.MORSEC_CCU[:6]) + [1.0 if simplified else 0.0]
        s.interactions[2] = Interaction(kind, params, mor.lists)
    return s


def run(s, precision, fuse, lanes, graphs=True, steps=20):
    c = gpu_ctx(s, precision)
    c.set_fusion(fuse)
    c.set_tuning(lanes, 256)
    c.set_graphs(graphs)
    c.run("nvt", 0.5, 0, 0, True)
    f0 = c.state()[2]
    e0 = c.energies()["e_inter"]
    c.run("nvt", 0.5, 1, steps, True)
    return f0, e0, c.state(), c.energies()["e_inter"]


@pytest.mark.parametrize("kind,simplified", [("morsec", False), ("morsec", True), ("ljc", False)])
@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_fused_equals_separate_fp64(kind, simplified, lanes):
    s = slab(kind, simplified)
    print(f"[fusion fp64] {kind} simplified={simplified} lanes={lanes} n={s.n}", flush=True)
    a = run(s, capi.FP64, True, lanes)
    b = run(s, capi.FP64, False, lanes)
    assert rel_inf(a[0], b[0]) < 1e-13
    np.testing.assert_allclose(a[1], b[1], rtol=1e-13)
    assert rel_inf(a[2][0], b[2][0]) < 1e-12
    assert rel_inf(a[2][2], b[2][2]) < 1e-11
    np.testing.assert_allclose(a[3], b[3], rtol=1e-12)


@pytest.mark.parametrize("graphs", [True, False])
def test_fused_equals_separate_mixed(graphs):
    s = slab()
    a = run(s, capi.MIXED, True, 4, graphs)
    b = run(s, capi.MIXED, False, 4, graphs)
    assert rel_inf(a[0], b[0]) < 1e-7
    np.testing.assert_allclose(a[1], b[1], rtol=1e-7)
    assert rel_inf(a[2][0], b[2][0]) < 1e-9


def test_fused_against_oracle():
    s = slab(seed=72)
    g, o = gpu_ctx(s), oracle_ctx(s)
    g.set_fusion(True)
    for c in (g, o):
        c.run("nvt", 0.5, 0, 0, True)
    assert rel_inf(g.state()[2], o.state()[2]) < 1e-10
    np.testing.assert_allclose(g.energies()["e_inter"], o.energies()["e_inter"], rtol=1e-10)
    for c in (g, o):
        c.run("nvt", 0.5, 1, 25, True)
    pg, vg, fg = g.state()
    po, vo, fo = o.state()
    assert rel_inf(pg, po) < 1e-10 and rel_inf(vg, vo) < 1e-8 and rel_inf(fg, fo) < 1e-8


# ------------------------------------------ lj / lj1g / rjl passes sharing launches

def run_plain(ctx_fn, integ, fuse, lanes, steps=20, dt=0.5):
    c = ctx_fn()
    c.set_fusion(fuse)
    c.set_tuning(lanes, 256)
    c.run(integ, dt, 0, 0, True)
    f0, e0 = c.state()[2], c.energies()["e_inter"]
    c.run(integ, dt, 1, steps, True)
    return f0, e0, c.state(), c.energies()["e_inter"]


@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_gas_passes_sharing_launches_fp64(lanes):
    """Binary gas: lj1g A | lj1g B, then lj A->B | lj B->A."""
    s = configs.binary_lj_gas(n=600, box=32.0, seed=73)
    a = run_plain(lambda: gpu_ctx(s, capi.FP64), "nvt", True, lanes)
    b = run_plain(lambda: gpu_ctx(s, capi.FP64), "nvt", False, lanes)
    assert rel_inf(a[0], b[0]) < 1e-13
    np.testing.assert_allclose(a[1], b[1], rtol=1e-13)
    assert rel_inf(a[2][0], b[2][0]) < 1e-12
    np.testing.assert_allclose(a[3], b[3], rtol=1e-12)


@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-13), (capi.MIXED, 1e-7)])
def test_batch_passes_sharing_launches(precision, tol):
    """A batch of LJ gases and Cu crystals: lj1g A | lj1g B | rjl densities,
    then lj A->B | lj B->A | rjl forces; per-replica energies too."""
    from pfmds_amd import system
    from tests.test_gpu_batch import replicas
    b = system.merge_replicas(replicas())
    out = []
    for fuse in (True, False):
        c = system.build_batch_context(b, precision)
        c.set_fusion(fuse)
        c.set_tuning(8, 256)
        c.run("nve", 0.5, 0, 0, True)
        f0 = c.state()[2]
        c.run("nve", 0.5, 1, 20, True)
        out.append((f0, c.state(), c.replica_energies()))
    (fa, sa, ra), (fb, sb, rb) = out
    assert rel_inf(fa, fb) < tol
    assert rel_inf(sa[0], sb[0]) < tol
    np.testing.assert_allclose(ra["pe"], rb["pe"], rtol=max(tol, 1e-12))


def test_gas_sharing_launches_against_oracle():
    s = configs.binary_lj_gas(n=600, box=32.0, seed=74)
    g, o = gpu_ctx(s), oracle_ctx(s)
    g.set_fusion(True)
    for c in (g, o):
        c.run("nvt", 0.5, 0, 0, True)
        c.run("nvt", 0.5, 1, 25, True)
    pg, vg, fg = g.state()
    po, vo, fo = o.state()
    assert rel_inf(pg, po) < 1e-10 and rel_inf(vg, vo) < 1e-8 and rel_inf(fg, fo) < 1e-8
    np.testing.assert_allclose(g.energies()["e_inter"], o.energies()["e_inter"], rtol=1e-9)


def test_sharing_launches_with_partly_written_group_against_oracle():
    """lj1g over A, then lj1g over A+B: the second pass finds some of its
    atoms written by the first and others not, so the shared launches cannot
    write each atom's force first; the step zeroes forces and every pass adds
    (the gas above takes the other way: lj1g A | lj1g B write, the lj rows
    add)."""
    s = configs.binary_lj_gas(n=600, box=32.0, seed=77)
    a, b = s.interactions[0], s.interactions[1]
    s.interactions = [a, Interaction("lj1g", b.params, [(3, 3) + tuple(b.lists[0][2:])])]
    g, o = gpu_ctx(s), oracle_ctx(s)
    g.set_fusion(True)
    for c in (g, o):
        c.run("nvt", 0.5, 0, 0, True)
    assert rel_inf(g.state()[2], o.state()[2]) < 1e-10
    for c in (g, o):
        c.run("nvt", 0.5, 1, 25, True)
    pg, vg, fg = g.state()
    po, vo, fo = o.state()
    assert rel_inf(pg, po) < 1e-10 and rel_inf(vg, vo) < 1e-8 and rel_inf(fg, fo) < 1e-8
    np.testing.assert_allclose(g.energies()["e_inter"], o.energies()["e_inter"], rtol=1e-9)


# ------------------------------------------ lanes per row of one interaction

@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-13), (capi.MIXED, 1e-7)])
def test_pass_lanes_equal_shared_lanes(precision, tol):
    """Separate launches with interaction 1 (lj1g A) at 4 lanes and 3 (lj)
    at 16 against every pass at 8; captured blocks, so the block cache must
    tell the settings apart."""
    s = configs.binary_lj_gas(n=600, box=32.0, seed=75)
    out = []
    for per in ({}, {1: 4, 3: 16}):
        c = gpu_ctx(s, precision)
        c.set_fusion(False)
        c.set_tuning(8, 256)
        c.set_graphs(True)
        c.run("nvt", 0.5, 0, 0, True)
        c.run("nvt", 0.5, 1, 20, True)   # blocks captured at 8 lanes
        for k, v in per.items():
            c.set_pass_lanes(k, v)
        c.run("nvt", 0.5, 21, 60, True)
        out.append((c.state(), c.energies()["e_inter"]))
    (sa, ea), (sb, eb) = out
    assert rel_inf(sa[2], sb[2]) < tol and rel_inf(sa[0], sb[0]) < tol
    np.testing.assert_allclose(ea, eb, rtol=max(tol, 1e-12))


def test_pass_lanes_rejects_bad_values():
    s = configs.binary_lj_gas(n=200, box=20.0, seed=76)
    c = gpu_ctx(s)
    for inter, lanes in ((0, 8), (4, 8), (1, 3), (1, 64)):
        with pytest.raises(capi.PfmError):
            c.set_pass_lanes(inter, lanes)
    c.set_pass_lanes(1, 0)
