"""The Fortran front end (pfmds_amd/bin/run_md_simulation), reading a
reference-format settings file and driving libpfmds through ISO_C_BINDING,
against the Python restatement of md()'s schedule driving the CPU oracle:
same log records (f24.6 as written), same nvms exit step, same final xyz,
and the outputs the front end writes from device records while the run goes
on: snapshots (f27.16), trajectory frames (f10.4) and the out block's centre
of mass (3f12.6) and largest speed (f12.6), each to the precision written."""
import os
import subprocess

import numpy as np
import pytest

from pfmds_amd import configs, system
from tests import growth_case
from tests.harness import oracle_ctx

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pfmds_amd", "bin", "run_md_simulation")


def parse_log(path, n_inter, n_nhc):
    recs = []
    for line in open(path):
        try:
            name = line[0:6].strip()
            step = int(line[6:15])
            vals = [float(line[15 + 24 * k:15 + 24 * (k + 1)]) for k in range(7)]
            off = 15 + 24 * 7
            e = [float(line[off + 20 * k:off + 20 * (k + 1)]) for k in range(n_inter + n_nhc)]
        except ValueError:
            continue
        recs.append((name, step, vals, e))
    return recs


@pytest.mark.parametrize("case", ["lj_gas", "gr_cu", "rebo", "growth"])
def test_frontend_matches_oracle(tmp_path, case):
    if case == "growth":
        # change_group lines in the settings; the trajectory group grows
        s = growth_case.deposition(n_side=6, kind="rjl")
        s.integrators = [system.Integrator("nve", 1.0, 30, 25, 5), system.Integrator("nvt", 0.5, 30, 1000, 5)]
        s.nhcs = [system.NHC(2, 150.0, 3, 1000.0)]
    elif case == "rebo":
        s = configs.graphene_rebo(gx=6, gy=4, jitter=0.05, buckle=0.1, seed=53)
        s.integrators = [system.Integrator("nve", 0.5, 40, 1000, 4)]
    elif case == "lj_gas":
        s = configs.binary_lj_gas(n=500, box=30.0, seed=51, schedule=(("nvt", 0.5, 30, 25, 5), ("nve", 0.5, 0, 5, 5),
                                                                       ("nve", 0.5, 30, 100, 5),
                                                                       ("nvms", 0.5, 400, 1000, 5)))
        s.ms_de = 1e-5
    else:
        s = configs.graphene_on_cu(cx=10, cy=6, gx=10, gy=6, layers=4, seed=52)
        s.integrators = [system.Integrator("nvms", 0.5, 40, 1000, 4), system.Integrator("nvt", 0.5, 40, 20, 4)]
        s.ms_de = 1e-12
    s.md_step_limit = sum(i.length for i in s.integrators)
    s.period_traj = 20 if case != "growth" else 7
    s.write_files(str(tmp_path))
    print(f"[frontend] {case} n={s.n}", flush=True)
    p = subprocess.run([EXE, "-i", "md_run_settings.txt", "-op", "10"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    got = parse_log(os.path.join(str(tmp_path), s.logfilename), len(s.interactions), len(s.nhcs))
    o = oracle_ctx(s)
    events = {}
    g_all = s.group_indices(s.all_atoms) - 1
    g_traj = s.group_indices(s.traj) - 1

    def on_event(t, kinds):
        p_, v_, _ = o.state()
        events[t] = (kinds, p_, v_, o.group_count(s.traj), o.group_count(s.all_atoms))

    want, last = system.run_md(s, o, out_period=10, on_event=on_event)
    assert [(r[0], r[1]) for r in got] == [(r.integrator, r.step) for r in want]
    for g, w in zip(got, want):
        ref = [w.time, w.conserved, w.e_nhc, w.total, w.pe, w.ke, w.temp]
        np.testing.assert_allclose(g[2], ref, atol=2e-6, rtol=1e-9)
        np.testing.assert_allclose(g[3], list(w.e_inter) + list(w.e_chains), atol=2e-9, rtol=1e-9)
    steps_line = [ln for ln in p.stdout.splitlines() if "steps number:" in ln]
    assert steps_line and int(steps_line[-1].split(":")[-1]) == last - 1
    # md_simulation.f90:246 after an nvms run
    de_line = [ln for ln in p.stdout.splitlines() if "potential energy difference:" in ln]
    if s.integrators[-1].name == "nvms":
        assert de_line and float(de_line[-1].split(":")[-1]) == pytest.approx(o.pe_change(), abs=1e-7)
    else:
        assert not de_line
    fin = os.path.join(str(tmp_path), "final_" + s.init_xyz_filename)
    _, pf, vf, _, _ = system.read_xyz(fin)
    po, vo, _ = o.state()
    g = s.group_indices(s.all_atoms) - 1
    np.testing.assert_allclose(pf, po[g], atol=1e-8)
    np.testing.assert_allclose(vf, vo[g], atol=1e-8)
    # snapshots and trajectory frames
    snaps = [t for t, ev in events.items() if "snapshot" in ev[0]]
    for t in snaps:
        _, ps, vs, _, _ = system.read_xyz(os.path.join(str(tmp_path), f"snapshot_{t:06d}.xyz"))
        ga = g_all[:events[t][4]]
        np.testing.assert_allclose(ps, events[t][1][ga], atol=1e-8)
        np.testing.assert_allclose(vs, events[t][2][ga], atol=1e-8)
    frames = read_traj(os.path.join(str(tmp_path), f"traj_{s.traj:02d}.xyz"))
    trajs = sorted(t for t, ev in events.items() if "traj" in ev[0])
    assert [f[0] for f in frames] == trajs
    for t, pt, vt in frames:
        # the trajectory group's active atoms at the step
        gt = g_traj[:events[t][3]]
        np.testing.assert_allclose(pt, events[t][1][gt], atol=1.0001e-4)
        np.testing.assert_allclose(vt, events[t][2][gt], atol=1.0001e-4)
    # the out block
    outs = parse_out(p.stdout)
    want_out = sorted(t for t, ev in events.items() if "out" in ev[0])
    assert [o_[0] for o_ in outs] == want_out
    for t, mc, vmax in outs:
        _, p_, v_, _, n_all = events[t]
        m = s.mass[g_all[:n_all]]
        np.testing.assert_allclose(mc, (m[:, None] * p_[g_all[:n_all]]).sum(0) / m.sum(), atol=1.0001e-6)
        np.testing.assert_allclose(vmax, np.sqrt((v_ * v_).sum(1).max()), atol=1.0001e-6)


def read_traj(path):
    """Frames of a write_particle_group_append file: (step, pos, vel)."""
    out = []
    with open(path) as f:
        while True:
            head = f.readline()
            if not head.strip():
                break
            n = int(head.split()[0])
            step = int(f.readline().split()[1])
            a = np.array([[float(x) for x in f.readline().split()[:7]] for _ in range(n)])
            out.append((step, a[:, 0:3], a[:, 3:6]))
    return out


def parse_out(text):
    """(step, centre of mass, largest speed) of every out block."""
    out, step, mc = [], None, None
    for ln in text.splitlines():
        if ln.startswith("step = "):
            step = int(ln.split("=")[1].split()[0])
        elif ln.startswith("c.o.m. position (A) = "):
            mc = np.array([float(x) for x in ln.split("=")[1].split()])
        elif ln.startswith("maximum velocity (A/fs) = "):
            out.append((step, mc, float(ln.split("=")[1])))
    return out
