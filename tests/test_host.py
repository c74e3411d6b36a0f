"""Host-side logic without a GPU: the C-ABI library loads and exports every
entry point include/pfm.h declares, the product fails loudly when no GPU is
visible, the reference file formats round-trip, and md()'s schedule (integrator
switching, event steps, nvms exit) drives the library as the reference's loop
would."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from pfmds_amd import capi, configs, system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "pfm.h")).read()
    return sorted(set(re.findall(r"\b(pfm_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.GPU_LIB)
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s


def test_oracle_mirrors_the_abi():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "liborc.so"))
    for s in header_symbols():
        # GPU-only extras: profiling, launch shape, replica batches (pinned
        # against single-system GPU runs, not the oracle), output records
        # (pinned against GPU runs stopped at every recorded step) and the
        # learned surrogate (pinned against tests/sur_ref.py)
        if s in ("pfm_set_timing", "pfm_kernel_time", "pfm_set_tuning", "pfm_set_graphs", "pfm_graph_count", "pfm_set_fusion",
                 "pfm_set_pass_lanes",
                 "pfm_row_energies",
                 "pfm_create_batch", "pfm_replica_energies", "pfm_set_records", "pfm_set_speculation",
                 "pfm_set_chain", "pfm_copy_state", "pfm_force_rmse", "pfm_force_rmse_noalloc") \
                or s.startswith("pfm_sur_"):
            continue
        assert hasattr(lib, s.replace("pfm_", "orc_")), s


def _gpu_visible():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU error path")
def test_no_device_fails_loudly():
    s = configs.fcc_cu_rgl(nx=2)
    with pytest.raises(capi.PfmError) as e:
        capi.Context.gpu(s.box, s.pos, s.vel, s.mass)
    assert e.value.code == capi.E_NO_DEVICE


def test_settings_and_xyz_round_trip(tmp_path):
    s = configs.graphene_on_cu(cx=4, cy=3, gx=4, gy=3, layers=2)
    s.integrators[0].length = 7
    s.write_files(str(tmp_path))
    r = system.read_files(str(tmp_path))
    np.testing.assert_array_equal(r.pos, s.pos)
    np.testing.assert_array_equal(r.vel, s.vel)
    np.testing.assert_array_equal(r.mass, s.mass)
    assert r.species == s.species and r.box.tolist() == s.box.tolist()
    assert [g for g in r.groups] == [g + ["#"] * (max(len(x) for x in s.groups) - len(g)) for g in s.groups]
    for a, b in zip(r.interactions, s.interactions):
        assert a.name == b.name and a.lists == b.lists
        np.testing.assert_array_equal(a.params, b.params)
    assert [(i.name, i.dt, i.length) for i in r.integrators] == [(i.name, i.dt, i.length) for i in s.integrators]
    assert r.nhcs == s.nhcs and r.ms_de == s.ms_de
    for g in range(1, len(s.groups) + 1):
        np.testing.assert_array_equal(r.group_indices(g), s.group_indices(g))


class FakeCtx:
    """Records pfm_run calls; converges an nvms run after `conv` steps."""

    def __init__(self, conv=None):
        self.calls, self.conv = [], conv

    def run(self, integ, dt, s0, s1, energy, ms_de):
        self.calls.append((integ, dt, s0, s1, energy))
        if integ == "nvms" and self.conv is not None and s1 >= self.conv:
            return max(s0, self.conv), capi.EV_CONVERGED
        return s1, capi.EV_DONE

    def energies(self):
        return dict(e_inter=np.zeros(1), pe=0.0, ke=0.0, temp=0.0, e_nhc=np.zeros(0), time=0.0)


def test_schedule_switching_and_events():
    s = configs.fcc_cu_rgl(nx=2)
    s.integrators = [system.Integrator("nvt", 0.5, 10, 4, 3), system.Integrator("nve", 1.0, 0, 5, 5),
                     system.Integrator("nve", 1.0, 6, 100, 4), system.Integrator("nvms", 0.5, 100, 100, 100)]
    s.md_step_limit = 1000
    f = FakeCtx(conv=23)
    recs, last = system.run_md(s, f, out_period=1000)
    # steps 0..10 nvt (zero-length nve skipped), 11..16 nve, then nvms from 17 until converged at 23
    covered = [(c[0], c[2], c[3]) for c in f.calls]
    assert covered[0][:2] == ("nvt", 0) and all(c[0] == "nvt" for c in covered if c[2] <= 10)
    assert [c for c in covered if c[0] == "nve"][0][1] == 11
    assert [c for c in covered if c[0] == "nve"][-1][2] == 16
    assert [c for c in covered if c[0] == "nvms"][0][1] == 17
    assert last == 23
    # chunks end exactly on log (3 or 4 steps), snapshot and integrator boundaries
    ends = [c[3] for c in f.calls]
    assert ends[:5] == [0, 3, 4, 6, 8]
    steps = sorted(x for c in f.calls for x in range(c[2], min(c[3], last) + 1))
    assert steps == list(range(0, 24))  # every step entered once, up to the nvms exit
    assert f.calls[-1][2] <= last <= f.calls[-1][3]
    assert [r.step for r in recs][:4] == [0, 3, 6, 9]


def test_fortran_frontend_fails_loudly_without_gpu(tmp_path):
    exe = os.path.join(ROOT, "pfmds_amd", "bin", "run_md_simulation")
    if _gpu_visible():
        pytest.skip("checks the no-GPU error path")
    s = configs.fcc_cu_rgl(nx=2)
    s.write_files(str(tmp_path))
    p = subprocess.run([exe, "-i", "md_run_settings.txt", "-op", "10"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=60)
    assert p.returncode != 0
    assert "no HIP device" in p.stdout


def test_merge_replicas_groups_roles_interactions():
    """Batch layout: groups matched by species list, shared interactions
    deduplicated, differing role groups united, replica slices in order."""
    from pfmds_amd import system
    a = configs.lj_gas_lattice(100, 0.02, 1, 100.0, steps=10)
    b = configs.fcc_cu_rgl(nx=4, seed=2, steps=10, dt=0.5)
    c = configs.lj_gas_lattice(120, 0.03, 3, 200.0, steps=10)
    bt = system.merge_replicas([a, b, c])
    s = bt.system
    assert s.groups == [["A"], ["B"], ["Cu"], ["A", "B", "Cu"], ["#"]]
    assert [it.name for it in s.interactions] == ["lj1g", "lj1g", "lj", "rjl"]
    assert s.interactions[3].lists == [[3, 3, 96, 6.5, 10]]
    assert s.groups[s.all_moving - 1] == ["A", "B", "Cu"] and s.groups[s.z_moving - 1] == ["#"]
    assert list(bt.n_atoms) == [100, 256, 120] and s.n == 476
    assert bt.replica_slice(1) == slice(100, 356)
    np.testing.assert_array_equal(s.pos[bt.replica_slice(2)], c.pos)
    assert np.allclose(bt.boxes[1], b.box)
    # group A of the batch holds the A atoms of both gas replicas, in order
    ga = s.group_indices(1) - 1
    assert list(ga) == list(range(0, 50)) + list(range(356, 416))


def test_replica_mix_sizes():
    reps = configs.replica_mix(n_rep=16, seed=7)
    ns = np.array([r.n for r in reps])
    assert ns.min() >= 128 and ns.max() <= 40000
    for k, r in enumerate(reps):
        if k % 2 == 0:
            d = r.pos[:, None, :] - r.pos[None, :, :] if r.n <= 2000 else None
            if d is not None:
                d -= r.box * np.round(d / r.box)
                dist = np.sqrt((d ** 2).sum(-1)) + np.eye(r.n) * 1e9
                assert dist.min() > 3.0
        else:
            assert min(r.box) >= 4 * 3.615 - 1e-9


def test_batch_carries_group_changes():
    # merge_replicas matches groups by species and maps every replica's change
    # lines onto the batch's groups; lines that differ between replicas are refused
    import pytest

    from pfmds_amd import system
    from tests import growth_case

    reps = [growth_case.deposition(n_side=4, seed=1), growth_case.deposition(n_side=3, seed=2, spacing=4.4)]
    for s in reps:
        s.groups[2] = ["A", "D", "#"]
    b = system.merge_replicas(reps)
    g_to = b.system.groups.index(["A", "D"]) + 1
    g_from = b.system.groups.index(["A"]) + 1
    assert b.system.changes == [(g_from, g_to, 5, 40, 7)]
    assert b.system.all_moving == g_to and b.system.all_atoms == b.system.groups.index(["A", "D", "#"]) + 1
    reps[1].changes = [(1, 2, 6, 40, 7)]
    with pytest.raises(ValueError):
        system.merge_replicas(reps)
