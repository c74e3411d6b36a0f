"""Output records (pfm_set_records) against stopping the run at every
recorded step.

The reference computes the log and out values and writes snapshots and
trajectories inside its step loop (md_simulation.f90:188-229); the chunked
runs here do the same through pfm_run/pfm_energies/pfm_get_state, ending a
call on every recorded step.  A recorded run covers the same steps in one
call.  Bar: interaction energies, kinetic energy, temperature, chain
energies, simulated time, positions and velocities bit-identical; the out
block's sums (force sum, centre of mass and its velocity, largest speed)
against float64 sums of the chunked state, within 1e-12 relative (the
reference's own sums run in OpenMP order), the force sum within 1e-10 of
sum |f| (it is a rounding residue).  The records must keep the captured
blocks in use, survive a ring smaller than the number of frames, take a
record on the last step of a block (that block runs uncaptured), and stop
the run with the callback's exception.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs, system
from tests.harness import gpu_ctx

pytestmark = pytest.mark.gpu


def due(periods, s):
    return any(p > 0 and s % p == 0 for p in periods)


def chunked(make, integ, dt, s_end, e_per, x_per, ms_de=0.0):
    """Values at every recorded step from runs that stop there."""
    c = make()
    out = {}
    s = 0
    while s <= s_end:
        t = s
        while t < s_end and not (due(e_per, t) or due(x_per, t)):
            t += 1
        want_e = due(e_per, t)
        last, ev = c.run(integ, dt, s, t, want_e, ms_de)
        if ev == capi.EV_CONVERGED:
            break
        rec = {}
        if want_e:
            rec.update(c.energies())
        if due(x_per, t) or want_e:
            p, v, f = c.state()
            rec.update(pos=p, vel=v, frc=f)
        out[t] = rec
        s = t + 1
    c.close()
    return out


def recorded(make, integ, dt, s_end, e_per, x_per, slots, graphs=True, ms_de=0.0, split=(0,)):
    c = make()
    c.set_graphs(graphs)
    got = {}

    def cb(r):
        assert r["step"] not in got, "a step was recorded twice"
        assert not got or r["step"] > max(got), "records out of step order"
        got[r["step"]] = r

    c.set_records(e_per, x_per, slots, cb)
    bounds = list(split) + [s_end + 1]
    for a, b in zip(bounds[:-1], bounds[1:]):
        last, ev = c.run(integ, dt, a, b - 1, False, ms_de)
        if ev == capi.EV_CONVERGED:
            break
    launched = c.graph_count()[1]
    c.close()
    return got, launched


def compare(ref, got, e_per, x_per, mass, all_atoms):
    assert sorted(got) == [s for s in sorted(ref) if due(e_per, s) or due(x_per, s)]
    for s, r in got.items():
        a = ref[s]
        if due(e_per, s):
            assert np.array_equal(r["e_inter"], a["e_inter"]), s
            assert r["ke"] == a["ke"] and r["temp"] == a["temp"] and r["pe"] == a["pe"], s
            assert np.array_equal(r["e_nhc"], a["e_nhc"]), s
            assert r["sim_time"] == a["time"], s
            p, v, f = a["pos"], a["vel"], a["frc"]
            m = mass[all_atoms]
            tm = m.sum()
            mc = (m[:, None] * p[all_atoms]).sum(0) / tm
            mcv = (m[:, None] * v[all_atoms]).sum(0) / tm
            fs = f[all_atoms].sum(0)
            scale_v = (m[:, None] * np.abs(v[all_atoms])).sum(0) / tm
            assert np.allclose(r["mass_center"], mc, rtol=1e-12, atol=0), s
            assert np.all(np.abs(r["mass_center_vel"] - mcv) <= 1e-12 * scale_v), s
            assert np.all(np.abs(r["force_sum"] - fs) <= 1e-10 * np.abs(f[all_atoms]).sum(0) + 1e-300), s
            assert r["max_vel"] == pytest.approx(np.sqrt((v * v).sum(1).max()), rel=1e-14), s
        if due(x_per, s):
            assert np.array_equal(r["pos"], a["pos"]), s
            assert np.array_equal(r["vel"], a["vel"]), s
        else:
            assert "pos" not in r


def test_records_graphene_cu_nvt_mixed():
    s = configs.graphene_on_cu(cx=12, cy=7, gx=12, gy=7, layers=4, seed=71)
    print(f"[records gr/cu nvt] n={s.n}", flush=True)
    mk = lambda: gpu_ctx(s, capi.MIXED)  # noqa: E731
    # lists every 10 steps: blocks of 10; period 19 puts records on the last
    # step of some blocks, 30 and 50 frames overflow a ring of 2 slots
    e_per, x_per = (10, 19), (30, 50)
    ref = chunked(mk, "nvt", 0.5, 150, e_per, x_per)
    got, launched = recorded(mk, "nvt", 0.5, 150, e_per, x_per, slots=2)
    print(f"  {len(got)} records, {launched} graph launches", flush=True)
    assert launched >= 5
    compare(ref, got, e_per, x_per, s.mass, np.array(s.group_indices(s.all_atoms)) - 1)


def test_records_fcc_nve_fp64_uncaptured_and_split():
    s = configs.fcc_cu_rgl(nx=8, seed=72)
    print(f"[records fcc nve] n={s.n}", flush=True)
    mk = lambda: gpu_ctx(s, capi.FP64)  # noqa: E731
    e_per, x_per = (7,), (7, 12)
    ref = chunked(mk, "nve", 1.0, 60, e_per, x_per)
    got, launched = recorded(mk, "nve", 1.0, 60, e_per, x_per, slots=3, graphs=False, split=(0, 1, 33))
    assert launched == 0
    compare(ref, got, e_per, x_per, s.mass, np.array(s.group_indices(s.all_atoms)) - 1)


def test_records_nvms_until_converged():
    s = configs.fcc_cu_rgl(nx=6, seed=73)
    print(f"[records fcc nvms] n={s.n}", flush=True)
    mk = lambda: gpu_ctx(s, capi.FP64)  # noqa: E731
    e_per, x_per = (3,), (5,)
    ref = chunked(mk, "nvms", 1.0, 400, e_per, x_per, ms_de=1e-3)
    got, _ = recorded(mk, "nvms", 1.0, 400, e_per, x_per, slots=2, ms_de=1e-3)
    last = max(got)
    ref = {k: v for k, v in ref.items() if k <= last}
    print(f"  converged at or after step {last}", flush=True)
    compare(ref, got, e_per, x_per, s.mass, np.array(s.group_indices(s.all_atoms)) - 1)


def test_records_callback_error_stops_run_and_handle_recovers():
    s = configs.fcc_cu_rgl(nx=6, seed=74)
    c = gpu_ctx(s, capi.MIXED)

    class Stop(Exception):
        pass

    seen = []

    def cb(r):
        seen.append(r["step"])
        if r["step"] == 20:
            raise Stop()

    c.set_records((5,), (10,), 2, cb)
    with pytest.raises(Stop):
        c.run("nve", 1.0, 0, 60, False)
    assert seen[-1] == 20
    # records off: the handle runs on
    c.set_records()
    c.run("nve", 1.0, 61, 70, True)
    assert np.isfinite(c.energies()["pe"])
    c.close()


def test_records_refused_for_batches():
    reps = configs.replica_mix(n_rep=3, seed=75)
    b = system.merge_replicas(reps)
    c = system.build_batch_context(b, capi.MIXED)
    with pytest.raises(capi.PfmError):
        c.set_records((5,), (), 1, lambda r: None)
    c.close()
