#!/usr/bin/env python3
"""Benchmark of the PFMDS MD step on one MI355X.

Workload: bulk fcc Cu, 32x32x32 conventional cells (131,072 atoms, a=3.615 A),
Rosato-Guillope-Legrand (rjl) only, cutoff 6.0 A, list r_cut 6.5 A rebuilt
every 10 steps, seeded Gaussian displacements (0.02 A) and Maxwell velocities
(300 K), nve at dt = 1 fs.  One "step" is one full md() step (integrator
halves, list rebuild on its schedule, density + force passes).

`value` is the mixed-precision path (FP32 pair arithmetic, FP64 dr,
accumulation and integration) in atom-steps/s; the plain FP64 path, the
force RMSE and energy error of mixed against FP64 at identical positions, and
the NVE drift are reported beside it.  `roofline` is the dominant kernel's
algorithmic HBM bytes per launch over its mean HIP-event duration;
`cpu_baseline` times the CPU restatement of the reference algorithm (brute
force O(N^2) list build, stored distances) on one rebuild period of the same
system.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pfmds_amd import autotune, capi, configs, system  # noqa: E402

METRIC = "atom-steps/s on 1 MI355X: mixed/batched/surrogate vs plain FP64 + force RMSE"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


GRAPHS = True  # --no-graphs: every step launched directly (profiling)


def gpu_ctx(s, precision):
    return system.build_context(s, lambda b, p, v, m: capi.Context.gpu(b, p, v, m, precision))


def tuned(ctx, integrator, dt, precision, first):
    """Launch shape and launch mode from the autotuner, tried on the live
    handle from step `first` on; returns ("LxB[-fusion][:iKxL...][+graphs]",
    next step): lanes x workgroup size, separate launches instead of the
    co-scheduled ones, interaction K's own lane count L, graph launches."""
    lanes, block, graphs, _, nxt = autotune.tune(ctx, integrator, dt, precision, first, steps=100, repeats=3,
                                                 graphs=GRAPHS)
    shape = f"{lanes}x{block}" + ("" if getattr(ctx, "fused", True) else "-fusion")
    shape += "".join(f":i{k}x{v}" for k, v in sorted(getattr(ctx, "pass_lanes", {}).items()))
    return shape + ("+graphs" if graphs else ""), nxt


def kernels_only(which, steps=50, warmup=10):
    """A short mixed-precision run of one workload (c1: fcc Cu RGL 131,072
    atoms, nve; c2: graphene on Cu(111) 94,560 atoms, nvt; c3: the 512
    replicas, 3,374,724 atoms, nve at 0.5 fs; c4: the surrogate on the
    graphene/Cu system, nvt, with an untrained network scaled to the
    system's own descriptors) with a fixed launch shape (8 lanes x 256
    threads, co-scheduled passes), no autotuning and no FP64 leg: the program
    of the counter passes."""
    if which == "c3":
        b = system.merge_replicas(configs.replica_mix())
        s, integ, dt = b.system, "nve", 0.5
        ctx = system.build_batch_context(b, capi.MIXED)
    elif which == "c4":
        from pfmds_amd import surrogate as sg
        s, integ, dt = configs.graphene_on_cu(), "nvt", 0.5
        _, sur_s, _ = sg.split_system(s)
        probe = gpu_ctx(sg.training_system(sur_s), capi.MIXED)
        probe.sur_set_params(1, sg.init_theta(np.zeros(40), np.ones(40), 0.0))
        probe.run("nve", dt, 0, 0, False)
        G = probe.sur_descriptors(1, len(s.group_indices(1)))
        probe.close()
        ctx = sg.surrogate_context(sur_s, sg.init_theta(G.mean(0), 1.0 / (G.std(0) + 1e-3), -7.0, seed=5))
    else:
        s = configs.fcc_cu_rgl() if which == "c1" else configs.graphene_on_cu()
        integ, dt = ("nve", 1.0) if which == "c1" else ("nvt", 0.5)
        ctx = gpu_ctx(s, capi.MIXED)
    ctx.set_tuning(8, 256)
    ctx.run(integ, dt, 0, warmup, True)
    t0 = time.perf_counter()
    ctx.run(integ, dt, warmup + 1, warmup + steps, False)
    t = time.perf_counter() - t0
    return {"workload": which, "atoms": s.n, "steps": steps, "ms_per_step": t * 1e3 / steps}


def timed_run(ctx, s, steps, warmup, precision):
    dt = s.integrators[0].dt
    ctx.run("nve", dt, 0, warmup, True)
    shape, first = tuned(ctx, "nve", dt, precision, warmup + 1)
    e0 = ctx.energies()
    # pfm_run returns after the stream has finished (its last status check
    # synchronises); no energy evaluation inside the timed steps
    t0 = time.perf_counter()
    ctx.run("nve", dt, first, first + steps - 1, False)
    t1 = time.perf_counter()
    return t1 - t0, e0, shape


def mean_list_len(ctx):
    nnum, _, _ = ctx.get_list(1, 1, stride=128)
    return float(np.mean(nnum))


def rjl_counts(n_entries):
    """Algorithmic bytes and flops per atom of the two mixed-precision RGL passes.

    bytes (unique HBM traffic; neighbour records are gathered from L2):
      density: own position record 16 + nnum 4 + indices 4*n + rho**0.5 write 8
               + 1/sqrt(rho) into the record's w slot 4
      force:   own position record 16 + nnum 4 + indices 4*n + force write 32
    flops per list entry (each + - * / sqrt exp sin cos sign counts 1):
      distance with minimum image 30; density term 13; force term 40
    """
    b_density = 16 + 4 + 4 * n_entries + 8 + 4
    b_force = 16 + 4 + 4 * n_entries + 32
    f_density = (30 + 13) * n_entries
    f_force = (30 + 40) * n_entries
    return {"rjl_density": (b_density, f_density), "rjl_force": (b_force, f_force)}


def measured_traffic(kernel):
    """HBM bytes per launch of the mixed-precision `kernel` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain copy);
    None when no summary for it is committed."""
    path = os.path.join(ROOT, "profiles", "rjl_fcc_cu_131k_traffic.json")
    if not os.path.exists(path):
        return None
    data = json.load(open(path))["kernels"]
    for name, v in data.items():
        if f"k_{kernel}<" in name and ", float" in name and "true" not in name:
            return v["bytes_per_launch"]
    return None


def nve_drift(ctx, n_atoms, dt, first, ps, segments=20):
    """Slope (eV/atom/ns) of (E_tot(t) - E_tot(0))/N over `ps` picoseconds
    of nve from step `first` on (the handle has run step first - 1),
    sampled at `segments` equal segments; the steps run and the fitted span."""
    nseg = max(1, int(round(ps * 1000.0 / dt / segments)))
    ts, es = [], []
    step = first
    for k in range(segments + 1):
        last = step + (nseg if k else 1) - 1
        ctx.run("nve", dt, step, last, True)
        e = ctx.energies()
        ts.append(e["time"])
        es.append(e["pe"] + e["ke"])
        step = last + 1
    t = np.array(ts) * 1e-6  # fs -> ns
    slope = float(np.polyfit(t, (np.array(es) - es[0]) / n_atoms, 1)[0])
    return slope, step, (ts[-1] - ts[0]) * 1e-3


def graphene_cu(steps, warmup, drift_ps=20.0):
    """Graphene on Cu(111) (94,560 atoms: tb + morsec + rjl, nvt at 300 K,
    dt = 0.5 fs) in FP64 and mixed precision, the mixed force RMSE at
    identical positions, and the NVE drift of both paths from there (energy
    per atom of all 94,560 atoms; the frozen Cu layers do no work)."""
    s = configs.graphene_on_cu()
    dt = s.integrators[-1].dt
    out, ctxs = {}, {}
    for name, prec in (("fp64", capi.FP64), ("mixed", capi.MIXED)):
        ctx = gpu_ctx(s, prec)
        ctx.run("nvt", dt, 0, warmup, True)
        shape, first = tuned(ctx, "nvt", dt, prec, warmup + 1)
        t0 = time.perf_counter()
        ctx.run("nvt", dt, first, first + steps - 1, False)
        t = time.perf_counter() - t0
        out[name + "_launch_shape"] = shape
        out[name + "_ms_per_step"] = t * 1e3 / steps
        out[name + "_atom_steps_per_s"] = s.n * steps / t
        ctxs[name] = ctx
    pos, vel, _ = ctxs["mixed"].state()
    for ctx in ctxs.values():
        ctx.set_state(pos, vel)
        ctx.run("nvt", dt, 0, 0, True)
    f64, f32 = ctxs["fp64"].state()[2], ctxs["mixed"].state()[2]
    out["speedup_vs_fp64"] = out["mixed_atom_steps_per_s"] / out["fp64_atom_steps_per_s"]
    out["force_rmse_eV_per_A"] = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    out["atoms"] = s.n
    out["steps"] = steps
    if drift_ps > 0:
        d = {}
        for name, ctx in ctxs.items():
            d[name], _, span = nve_drift(ctx, s.n, dt, 1, drift_ps)
        d["fit_ps"] = span
        d["unit"] = "eV/atom/ns"
        d["start"] = "the mixed path's state after its timed nvt steps, both paths"
        out["nve_drift"] = d
        print(f"# graphene/Cu NVE drift over {span:.1f} ps: mixed {d['mixed']:.3e}, fp64 {d['fp64']:.3e} eV/atom/ns",
              file=sys.stderr, flush=True)
    print(f"# graphene/Cu: mixed {out['mixed_ms_per_step']:.4f} ms/step, fp64 {out['fp64_ms_per_step']:.4f} ms/step",
          file=sys.stderr, flush=True)
    return out


def output_path(steps=2000, log_every=10, traj_every=100, warmup=20):
    """md() with its outputs on graphene on Cu(111) (94,560 atoms, nvt, mixed
    precision): an energy record every `log_every` steps and a frame of every
    atom every `traj_every` steps, taken (a) by stopping the run at each
    output step (pfm_run to the step, pfm_energies, pfm_get_state: the
    reference's synchronous pattern) and (b) with output records
    (pfm_set_records), beside the same steps with no output.  Text
    formatting is left out of both; the callback copies what it is handed."""
    s = configs.graphene_on_cu()
    dt = s.integrators[-1].dt
    ctx = gpu_ctx(s, capi.MIXED)
    ctx.run("nvt", dt, 0, warmup, True)
    shape, step = tuned(ctx, "nvt", dt, capi.MIXED, warmup + 1)
    step = -(-step // traj_every) * traj_every + 1
    out = {"atoms": s.n, "steps": steps, "log_every": log_every, "traj_every": traj_every, "launch_shape": shape}
    t0 = time.perf_counter()
    ctx.run("nvt", dt, step, step + steps - 1, True)
    out["no_output_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
    step += steps
    t0 = time.perf_counter()
    end = step + steps - 1
    while step <= end:
        t = min(end, -(-step // log_every) * log_every)
        ctx.run("nvt", dt, step, t, True)
        ctx.energies()
        if t % traj_every == 0:
            ctx.state()
        step = t + 1
    out["stop_at_outputs_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
    got = []

    def keep(r):
        got.append(r["step"])

    ctx.set_records((log_every,), (traj_every,), 4, keep)
    t0 = time.perf_counter()
    ctx.run("nvt", dt, step, step + steps - 1, True)
    out["records_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
    step += steps
    out["records_delivered"] = len(got)
    # the same with captured blocks: the host issues one launch per block,
    # so the callback's host work overlaps the device's
    ctx.set_graphs(True)
    t0 = time.perf_counter()
    ctx.run("nvt", dt, step, step + steps - 1, True)
    out["records_graphs_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
    step += steps
    ctx.set_records()
    t0 = time.perf_counter()
    ctx.run("nvt", dt, step, step + steps - 1, True)
    out["no_output_graphs_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
    ctx.close()
    print(f"# output path: none {out['no_output_ms_per_step']:.4f}, stop {out['stop_at_outputs_ms_per_step']:.4f}, "
          f"records {out['records_ms_per_step']:.4f}, records+graphs {out['records_graphs_ms_per_step']:.4f}, "
          f"none+graphs {out['no_output_graphs_ms_per_step']:.4f} ms/step", file=sys.stderr, flush=True)
    return out


def moire_fit(steps=2000, cells=((30, 17, 31, 18), (32, 18, 33, 19)), jobs=8):
    """The moire fit's first search step (pfmds_amd/moire.py): the four
    starting points of a coordinate, eight md() runs of `steps` nvms steps on
    two graphene-on-Cu(111) cells, one run at a time and `jobs` at once."""
    import tempfile

    from pfmds_amd import moire
    out = {"cells": [list(c) for c in cells], "steps": steps, "evaluations": 4, "jobs": jobs}
    lines = {}
    for j in (1, jobs):
        with tempfile.TemporaryDirectory() as d:
            fs = moire.read_fitting_parameters(configs.moire_fit_case(os.path.join(d, "in"), cells=cells, steps=steps))
            fit = moire.Fitter(fs, out_period=10**9, jobs=j, out=open(os.devnull, "w"))
            mid = (fs.min_params + fs.max_params) / 2
            cand = []
            for x in (fs.min_params[0], fs.max_params[0], fs.min_params[0] + (fs.max_params[0] - fs.min_params[0]) * moire.GOLD,
                      fs.max_params[0] - (fs.max_params[0] - fs.min_params[0]) * moire.GOLD):
                q = mid.copy()
                q[0] = x
                cand.append((q, True))
            t0 = time.perf_counter()
            fit.evaluate(cand)
            out[f"jobs{j}_s_per_evaluation"] = (time.perf_counter() - t0) / 4
            lines[j] = open(os.path.join(fs.out_path, "fitfit_out.txt")).read()
            fit.close()
            atoms = []
            for i in (0, 1):
                with open(os.path.join(fs.input_path, fs.start_xyz[i])) as f:
                    atoms.append(int(f.readline()))
            out["atoms"] = atoms
    out["identical_results"] = lines[1] == lines[jobs]
    out["speedup"] = out["jobs1_s_per_evaluation"] / out[f"jobs{jobs}_s_per_evaluation"]
    print(f"# moire fit: {out['jobs1_s_per_evaluation']:.3f} s/evaluation one at a time, "
          f"{out[f'jobs{jobs}_s_per_evaluation']:.3f} with {jobs} at once", file=sys.stderr, flush=True)
    return out


def nvms_phase(steps=500, warmup=20):
    """The nvms relaxation that opens the graphene/Cu workload (94,560 atoms,
    mixed precision): the exit test on the host every step, against nvms
    blocks (pfm_set_speculation) that read energies back once per block."""
    s = configs.graphene_on_cu()
    dt = s.integrators[0].dt
    out = {"atoms": s.n, "steps": steps}
    for name, spec in (("per_step_test", False), ("blocks", True)):
        ctx = gpu_ctx(s, capi.MIXED)
        ctx.set_speculation(spec)
        ctx.run("nvms", dt, 0, warmup, True, 1e-30)
        t0 = time.perf_counter()
        ctx.run("nvms", dt, warmup + 1, warmup + steps, True, 1e-30)
        out[name + "_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
        ctx.close()
    out["speedup"] = out["per_step_test_ms_per_step"] / out["blocks_ms_per_step"]
    print(f"# nvms: {out['per_step_test_ms_per_step']:.4f} ms/step testing every step, "
          f"{out['blocks_ms_per_step']:.4f} in blocks", file=sys.stderr, flush=True)
    return out


def surrogate_c4(train_iters=20000, frames_full=125, frames_sub=875, steps=10000, judge_every=100,
                 threshold=0.05, mixed_ms=None, fp64_ms=None, log=None):
    """Learned surrogate on graphene/Cu (94,560 atoms): frames from plain nvt
    runs at 300-900 K on the full system and on a 1/16 sub-box (labels: the
    FP64 tb + morsec path), training on the GPU, held-out force error, and a
    judged surrogate nvt run at 300 K (rjl + sur, the plain potential every
    `judge_every` steps, fallback past `threshold` eV/A)."""
    from pfmds_amd import surrogate as sg
    log = log or (lambda m: print("# c4: " + m, file=sys.stderr, flush=True))
    temps = [300.0, 500.0, 700.0, 900.0]
    s = configs.graphene_on_cu()
    sub = configs.graphene_on_cu(cx=25, cy=15, gx=26, gy=16, seed=1014)
    t0 = time.perf_counter()
    fr_full = sg.generate_frames(s, temps, frames_full, every=50, warmup=200)
    fr_sub = sg.generate_frames(sub, temps, frames_sub, every=50, warmup=200)
    t_frames = time.perf_counter() - t0
    log(f"{len(fr_full)} + {len(fr_sub)} frames in {t_frames:.1f} s")
    # held out: every 25th full-system frame
    hold = np.arange(len(fr_full)) % 25 == 24
    test_f = fr_full.subset(np.nonzero(hold)[0])
    train_full = fr_full.subset(np.nonzero(~hold)[0])
    ctxs = []
    for sysm in (s, sub):
        _, sur_s, _ = sg.split_system(sysm)
        ctxs.append(gpu_ctx(sg.training_system(sur_s), capi.MIXED))
    n_c = len(s.group_indices(1))
    mu, isg = sg.feature_scaling(ctxs[0], 1, train_full, n_c)
    theta0 = sg.init_theta(mu, isg, float(np.mean(train_full.energy)) / n_c, seed=1004)
    g_all = s.group_indices(s.all_atoms) - 1
    ctxs[0].sur_set_params(1, theta0)
    rmse0 = sg.force_rmse(ctxs[0], test_f, g_all)
    t0 = time.perf_counter()
    theta, hist = sg.train(ctxs, 1, [train_full, fr_sub], theta0, train_iters, lr=3e-3, lr_end=5e-5,
                           seed=1004, log=log, log_every=max(train_iters // 6, 1))
    t_train = time.perf_counter() - t0
    rmse1 = sg.force_rmse(ctxs[0], test_f, g_all)
    t300 = test_f.subset(np.nonzero(test_f.temperature == 300.0)[0])
    rmse300 = sg.force_rmse(ctxs[0], t300, g_all) if len(t300) else None
    rms_label = float(np.sqrt(np.mean(test_f.force[:, g_all].astype(np.float64) ** 2)))
    log(f"trained {train_iters} iterations in {t_train:.1f} s: held-out force RMSE {rmse0:.4f} -> {rmse1:.4f} "
        f"eV/A (labels {rms_label:.4f})")
    for c in ctxs:
        c.close()
    # judged inference at 300 K
    _, sur_s, _ = sg.split_system(s)
    dt = s.integrators[-1].dt
    sc = sg.surrogate_context(sur_s, theta)
    pc = gpu_ctx(s, capi.MIXED)
    moving = s.group_indices(s.all_moving) - 1
    jr = sg.JudgedRun(sc, pc, moving, judge_every=judge_every, threshold=threshold, group=s.all_moving)
    jr.run("nvt", dt, 0, 20)
    sc.run("nvt", dt, 21, 21, False)
    t0 = time.perf_counter()
    end = jr.run("nvt", dt, 22, 21 + steps)
    sc.energies()
    t_run = time.perf_counter() - t0
    # the surrogate step alone (no judge), same handle
    t0 = time.perf_counter()
    if end is sc:
        sc.run("nvt", dt, 22 + steps, 21 + 2 * steps // 5 + steps, False)
        sc.energies()
    t_pure = (time.perf_counter() - t0) / (2 * steps // 5) if end is sc else None
    rm = [r for _, r in jr.judged]
    out = {
        "atoms": s.n, "carbon_atoms": n_c,
        "frames": {"full_system": len(fr_full), "sub_box": len(fr_sub), "sub_box_atoms": sub.n,
                   "temperatures_K": temps, "every_steps": 50, "seconds": t_frames},
        "train": {"iterations": train_iters, "seconds": t_train, "loss_first_100": float(hist[:100, 0].mean()),
                  "loss_last_100": float(hist[-100:, 0].mean())},
        "heldout_force_rmse_eV_per_A": {"trained": rmse1, "trained_300K": rmse300, "untrained": rmse0,
                                        "labels_rms": rms_label, "frames": int(hold.sum())},
        "run": {"steps": steps, "judge_every": judge_every, "threshold_eV_per_A": threshold,
                "judged_rmse_max": max(rm) if rm else None, "judged_rmse_mean": float(np.mean(rm)) if rm else None,
                "judges": len(rm), "fallback_step": jr.switch_step, "judge_seconds": jr.t_judge},
        "ms_per_step_with_judge": t_run * 1e3 / steps,
        "atom_steps_per_s_with_judge": s.n * steps / t_run,
        "ms_per_step_surrogate": t_pure * 1e3 if t_pure else None,
        "atom_steps_per_s_surrogate": s.n / t_pure if t_pure else None,
    }
    if mixed_ms:
        out["vs_mixed"] = mixed_ms / (t_run * 1e3 / steps)
    if fp64_ms:
        out["vs_fp64"] = fp64_ms / (t_run * 1e3 / steps)
    log(f"surrogate run: {out['ms_per_step_with_judge']:.4f} ms/step with judge, judged RMSE max "
        f"{out['run']['judged_rmse_max']}, fallback at {jr.switch_step}")
    sc.close()
    pc.close()
    return out


def replica_batch(steps, fp64_steps, check_steps=100, n_check=8):
    """512 independent replicas (binary LJ gas and fcc Cu, 256-32768 atoms
    each, about 3.4M atoms) stepped together with nve at dt = 0.5 fs, in mixed
    precision and in FP64, and the largest per-replica deviation of forces and
    potential energy from running the same replicas one at a time."""
    reps = configs.replica_mix()
    b = system.merge_replicas(reps)
    out = {"replicas": b.n_rep, "atoms": int(b.system.n)}
    for name, prec, n in (("mixed", capi.MIXED, steps), ("fp64", capi.FP64, fp64_steps)):
        g = system.build_batch_context(b, prec)
        g.run("nve", 0.5, 0, 20, True)
        shape, first = tuned(g, "nve", 0.5, prec, 21)
        t0 = time.perf_counter()
        g.run("nve", 0.5, first, first + n - 1, False)
        t = time.perf_counter() - t0
        out[name + "_launch_shape"] = shape
        out[name + "_ms_per_step"] = t * 1e3 / n
        out[name + "_atom_steps_per_s"] = b.system.n * n / t
        out[name + "_steps"] = n
        g.close()
    out["speedup_vs_fp64"] = out["mixed_atom_steps_per_s"] / out["fp64_atom_steps_per_s"]
    # per-replica deviation from single runs, mixed precision, on a sample
    g = system.build_batch_context(b, capi.MIXED)
    g.run("nve", 0.5, 0, check_steps, True)
    _, _, frc = g.state()
    pe = g.replica_energies()["pe"]
    idx = np.linspace(0, b.n_rep - 1, n_check).astype(int)
    fdev, edev = 0.0, 0.0
    for r in idx:
        c = gpu_ctx(reps[r], capi.MIXED)
        c.run("nve", 0.5, 0, check_steps, True)
        f1 = c.state()[2]
        sl = b.replica_slice(r)
        fdev = max(fdev, float(np.abs(frc[sl] - f1).max() / max(np.abs(f1).max(), 1e-30)))
        edev = max(edev, abs(pe[r] - c.energies()["pe"]) / max(abs(c.energies()["pe"]), 1e-30))
        c.close()
    out["check"] = {"replicas": [int(x) for x in idx], "steps": check_steps,
                    "max_force_dev_rel": fdev, "max_pe_dev_rel": edev}
    print(f"# replicas: mixed {out['mixed_ms_per_step']:.3f} ms/step, fp64 {out['fp64_ms_per_step']:.3f} ms/step",
          file=sys.stderr, flush=True)
    return out


def host_cpu():
    """The host's CPU model (lscpu), its physical cores, the CPUs this process
    may run on, the cgroup's CPU quota and the CPU share OMP_NUM_THREADS
    grants this job (a node that gives a job one of its GPUs gives it 16
    CPUs); `threads` = the
    physical cores among the usable CPUs, capped by the quota and the share."""
    info = {"model": None, "physical_cores": None, "usable_cpus": len(os.sched_getaffinity(0)),
            "quota_cpus": None}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
        rows = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                              timeout=20).stdout.splitlines()
        rows = [r.split(",") for r in rows if r and not r.startswith("#")]
        info["physical_cores"] = len({(r[1], r[2]) for r in rows})
        usable = os.sched_getaffinity(0)
        usable_cores = len({(r[1], r[2]) for r in rows if int(r[0]) in usable})
    except (OSError, ValueError, subprocess.SubprocessError):
        usable_cores = info["usable_cpus"]
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            info["quota_cpus"] = float(q) / float(per)
    except (OSError, ValueError):
        pass
    threads = usable_cores
    if info["quota_cpus"]:
        threads = min(threads, max(1, int(info["quota_cpus"])))
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        info["share_cpus"] = int(share)
        threads = min(threads, int(share))
    info["threads"] = max(1, threads)
    return info


def _oracle(threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from tests.harness import oracle_ctx, oracle_lib

    lib = oracle_lib()
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_set_threads(threads)
    return oracle_ctx


def _cpu_leg_system(which, threads):
    """One CPU leg in this (child) process: the oracle, which restates the
    reference algorithm (brute-force O(N^2) list search, stored distances),
    on one list-rebuild period (steps 0-9) of C1 (nve) or C2 (nvt)."""
    oracle_ctx = _oracle(threads)
    s = configs.fcc_cu_rgl() if which == "c1" else configs.graphene_on_cu()
    integ, dt = ("nve", 1.0) if which == "c1" else ("nvt", 0.5)
    o = oracle_ctx(s)
    period = s.interactions[0].lists[0][4]
    t0 = time.perf_counter()
    o.run(integ, dt, 0, period - 1, True)
    t = time.perf_counter() - t0
    return {"value": s.n * period / t, "atoms": s.n, "steps": period, "seconds": t}


def _farm_one(k):
    s = configs.replica_mix(only=k)[k]
    oracle_ctx = _oracle(1)
    o = oracle_ctx(s)
    o.run("nve", 0.5, 0, 19, True)
    return s.n * 20


def _cpu_leg_farm(threads, per_worker=2):
    """The reference's farm of independent runs (run_md_simulation_mpi.f90:68-88)
    modelled on this host: one replica per core (OMP_NUM_THREADS=1), a
    size-stratified sample of the 512 replicas (per_worker per core), 20 steps
    each (one LJ rebuild period, two Cu ones)."""
    import multiprocessing as mp

    reps = configs.replica_mix()
    order = np.argsort([r.n for r in reps], kind="stable")
    pick = [int(order[i]) for i in np.linspace(0, len(reps) - 1, per_worker * threads).astype(int)]
    env_threads = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "1"
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(threads) as pool:
        work = sum(pool.map(_farm_one, pick, chunksize=1))
    t = time.perf_counter() - t0
    if env_threads is not None:
        os.environ["OMP_NUM_THREADS"] = env_threads
    return {"value": work / t, "replicas": pick, "atoms": int(sum(reps[k].n for k in pick)), "steps": 20,
            "seconds": t}


def cpu_baselines():
    """The CPU legs, each in a child process with OMP_NUM_THREADS = the usable
    physical cores, OMP_PLACES=cores, OMP_PROC_BIND=close."""
    info = host_cpu()
    n = info["threads"]
    env = dict(os.environ, OMP_NUM_THREADS=str(n), OMP_PLACES="cores", OMP_PROC_BIND="close")
    out = {}
    for leg in ("c1", "c2", "farm"):
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", leg, "--cpu-threads", str(n)],
                           env=env, capture_output=True, text=True, timeout=900)
        if p.returncode != 0:
            raise RuntimeError(f"cpu leg {leg} failed: {p.stderr[-2000:]}")
        out[leg] = json.loads(p.stdout.strip().splitlines()[-1])
        print(f"# cpu {leg}: {out[leg]['value']:.4e} atom-steps/s ({out[leg]['seconds']:.1f} s, {n} threads)",
              file=sys.stderr, flush=True)
    c1 = out["c1"]
    return {"value": c1["value"], "unit": "atom-steps/s", "cores": n, "kind": "port",
            "sample": f"steps 0-{c1['steps'] - 1} (one list-rebuild period) of the same {c1['atoms']}-atom system, "
                      f"{c1['seconds']:.1f} s wall, OpenMP on {n} threads (OMP_PLACES=cores)",
            "cpu": info,
            "graphene_cu_nvt": {"value": out["c2"]["value"], "unit": "atom-steps/s",
                                "sample": f"nvt steps 0-{out['c2']['steps'] - 1} of the {out['c2']['atoms']}-atom "
                                          f"graphene/Cu system, {out['c2']['seconds']:.1f} s wall"},
            "replica_farm": {"value": out["farm"]["value"], "unit": "atom-steps/s",
                             "sample": f"{len(out['farm']['replicas'])} of the 512 replicas (size-stratified, "
                                       f"{out['farm']['atoms']} atoms), 20 nve steps each, one replica per "
                                       f"core (OMP_NUM_THREADS=1), {out['farm']['seconds']:.1f} s wall"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--drift-ps", type=float, default=100.0)
    ap.add_argument("--no-c2", action="store_true", help="skip the graphene/Cu workload")
    ap.add_argument("--c2-only", action="store_true", help="time only the graphene/Cu workload (profiling)")
    ap.add_argument("--no-c3", action="store_true", help="skip the replica batch")
    ap.add_argument("--c3-only", action="store_true", help="time only the replica batch (profiling)")
    ap.add_argument("--no-c4", action="store_true", help="skip the learned surrogate")
    ap.add_argument("--c4-only", action="store_true", help="run only the learned surrogate")
    ap.add_argument("--c4-iters", type=int, default=20000, help="surrogate training iterations")
    ap.add_argument("--c4-small", action="store_true",
                    help="surrogate with 10 + 40 frames and a 2,000-step run (profiling)")
    ap.add_argument("--fit", action="store_true", help="time only the moire fit's first search step")
    ap.add_argument("--nvms", action="store_true", help="time only the graphene/Cu nvms relaxation")
    ap.add_argument("--output-path", action="store_true",
                    help="time only the run with outputs: stopping at each output step vs output records")
    ap.add_argument("--kernels", choices=("c1", "c2", "c3", "c4"),
                    help="only a short fixed-shape mixed run of one workload (counter passes)")
    ap.add_argument("--no-graphs", action="store_true", help="launch every step directly (profiling)")
    ap.add_argument("--cpu-leg", choices=("c1", "c2", "farm"), help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", type=int, default=1, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg:
        # one CPU leg, run by cpu_baselines() in a child process
        leg = _cpu_leg_farm(args.cpu_threads) if args.cpu_leg == "farm" else _cpu_leg_system(args.cpu_leg,
                                                                                              args.cpu_threads)
        print(json.dumps(leg))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the step path has no multi-GPU decomposition "
              "(replicas or spatial domains would shard it; see the design notes). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.no_graphs:
        global GRAPHS
        GRAPHS = False
        capi.GRAPHS_DEFAULT = False
    if args.kernels:
        print(json.dumps({"kernels": kernels_only(args.kernels)}))
        return
    if args.c2_only:
        print(json.dumps({"graphene_cu_nvt": graphene_cu(args.steps, args.warmup, drift_ps=0)}))
        return
    if args.nvms:
        print(json.dumps({"nvms_phase": nvms_phase(min(args.steps, 2000), args.warmup)}))
        return
    if args.fit:
        print(json.dumps({"moire_fit": moire_fit(args.steps)}))
        return
    if args.output_path:
        print(json.dumps({"output_path": output_path(args.steps, warmup=args.warmup)}))
        return
    if args.c3_only:
        print(json.dumps({"replica_batch": replica_batch(args.steps, args.steps, check_steps=10, n_check=2)}))
        return
    if args.c4_only:
        c2 = graphene_cu(200, 20, drift_ps=0)
        kw = dict(frames_full=10, frames_sub=40, steps=2000) if args.c4_small else {}
        print(json.dumps({"surrogate": surrogate_c4(args.c4_iters, mixed_ms=c2["mixed_ms_per_step"],
                                                    fp64_ms=c2["fp64_ms_per_step"], **kw), "graphene_cu_nvt": c2}))
        return
    s = configs.fcc_cu_rgl(nx=args.cells, steps=args.steps)
    res = {}
    ctxs = {}
    for name, prec in (("fp64", capi.FP64), ("mixed", capi.MIXED)):
        ctx = gpu_ctx(s, prec)
        t, e0, shape = timed_run(ctx, s, args.steps, args.warmup, prec)
        res[name] = dict(seconds=t, value=s.n * args.steps / t, e0=e0, shape=shape)
        ctxs[name] = ctx
        print(f"# {name}: {t * 1e3 / args.steps:.4f} ms/step, {s.n * args.steps / t:.4e} atom-steps/s", file=sys.stderr,
              flush=True)

    # accuracy of the mixed path at identical positions
    pos, vel, _ = ctxs["mixed"].state()
    for name in ("fp64", "mixed"):
        ctxs[name].set_state(pos, vel)
        ctxs[name].run("nve", 1.0, 0, 0, True)
    f64 = ctxs["fp64"].state()[2]
    f32 = ctxs["mixed"].state()[2]
    rmse = float(np.sqrt(np.mean((f32 - f64) ** 2)))
    pe64, pe32 = ctxs["fp64"].energies()["pe"], ctxs["mixed"].energies()["pe"]
    nbar = mean_list_len(ctxs["mixed"])

    # kernel durations with HIP events on the library's stream
    kt = ctxs["mixed"]
    kt.set_timing(True)
    kt.run("nve", 1.0, 1, 100, False)
    counts = rjl_counts(nbar)
    kern = {}
    for k in ("rjl_force", "rjl_density", "kick_drift", "kick2", "build_list"):
        c, ms = kt.kernel_time(k)
        kern[k] = (c, ms)
    kt.set_timing(False)
    dom = max(("rjl_force", "rjl_density"), key=lambda k: kern[k][0] * kern[k][1])
    cnt, ms = kern[dom]
    bytes_launch = counts[dom][0] * s.n
    flops_launch = counts[dom][1] * s.n
    achieved = bytes_launch / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom), "kernel": dom, "mean_us": ms * 1e3,
                "launches": cnt, "bytes_per_launch": bytes_launch,
                "valu": {"achieved_tflops": flops_launch / (ms * 1e-3) / 1e12, "peak_tflops": FP32_PEAK_TFLOPS,
                         "frac": flops_launch / (ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS}}

    # NVE drift: fitted slope of (E - E0)/N over drift_ps, after a 20 ps
    # equilibration of the displaced lattice (its energy settles by about
    # 1e-4 eV/atom/ns in the first 20 ps in both precisions)
    drift = None
    if args.drift_ps > 0:
        drift = {}
        for name, prec in (("mixed", capi.MIXED), ("fp64", capi.FP64)):
            d = gpu_ctx(s, prec)
            d.run("nve", 1.0, 0, 20000, True)
            e = d.energies()
            ts, es = [e["time"]], [e["pe"] + e["ke"]]
            nseg = 50
            seg = max(1, int(args.drift_ps * 1000 / nseg))
            for k in range(nseg):
                d.run("nve", 1.0, 20001 + k * seg, 20000 + (k + 1) * seg, True)
                e = d.energies()
                ts.append(e["time"])
                es.append(e["pe"] + e["ke"])
            drift[name] = float(np.polyfit(np.array(ts) * 1e-6, (np.array(es) - es[0]) / s.n, 1)[0])
            d.close()
        drift["fit_ps"] = args.drift_ps
        drift["equilibration_ps"] = 20.0

    c2 = None if args.no_c2 else graphene_cu(200, 20)
    c3 = None if args.no_c3 else replica_batch(2000, 500)
    c4 = None
    if not args.no_c4:
        c4 = surrogate_c4(args.c4_iters, mixed_ms=c2["mixed_ms_per_step"] if c2 else None,
                          fp64_ms=c2["fp64_ms_per_step"] if c2 else None)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines()

    mixed, fp64 = res["mixed"], res["fp64"]
    line = {
        "metric": METRIC,
        "value": mixed["value"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["seconds"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic, f64 dr/accumulation/integration",
        "data": "synthetic",
        "config": {"workload": f"fcc_cu_rgl {args.cells}^3 cells ({s.n} atoms), rjl R1=5.5 R2=6.0 A, "
                               "r_cut 6.5 A / 10 steps, nve dt=1 fs",
                   "atoms": s.n, "mean_list_entries": nbar},
        "fp64_value": fp64["value"],
        "fp64_ms_per_step": fp64["seconds"] * 1e3 / args.steps,
        "speedup_vs_fp64": mixed["value"] / fp64["value"],
        "launch_shape": {"mixed": mixed["shape"], "fp64": fp64["shape"],
                         "note": "lanes per row x workgroup size of the pair passes, and +graphs when steps "
                                 "launched from captured HIP graphs were faster, picked by pfmds_amd.autotune "
                                 "on the live handle before the timed steps"},
        "force_rmse_eV_per_A": rmse,
        "pe_err_eV_per_atom": abs(pe32 - pe64) / s.n,
        "nve_drift_eV_per_atom_per_ns": drift,
        "kernels_us": {k: v[1] * 1e3 for k, v in kern.items()},
        "roofline": roofline,
        "graphene_cu_nvt": c2,
        "replica_batch": c3,
        "surrogate": c4,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
