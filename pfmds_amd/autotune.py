"""Launch-shape autotuner for the pair passes of libpfmds.

The pair passes (rjl density and force, lj, lj1g, ljc/morsec graphene and
metal rows) walk each atom row with T lanes inside workgroups of B threads
(`pfm_set_tuning`).  The best shape depends on the row length, the precision
and the system size.  `tune` times a short run of the live handle for every
candidate shape, keeps the fastest, then times the steps launched from
captured graphs against launched one by one (`pfm_set_graphs`: graphs pay
off where launches dominate, small systems) and the co-scheduled passes of
graphene on a metal against separate launches (`pfm_set_fusion`); with
separate launches, the lanes per row of each interaction's own passes
(`pfm_set_pass_lanes`, one interaction at a time); and leaves the faster set;
results are cached in a JSON file keyed by a signature of the workload.  No counterpart in the
reference, which tunes only its OpenMP thread count
(perfomance_settings.f90:13-25).
"""
import json
import os
import time

# (lanes per row, workgroup size); FP64 passes are limited to 256 threads
CANDIDATES = [(4, 256), (8, 128), (8, 256), (8, 512), (16, 256), (16, 512), (32, 256)]
# lanes per row tried for each interaction's own passes, and the gain a
# per-interaction count must show over the shared one to be kept
PASS_LANES = (4, 8, 16)
PASS_MARGIN = 0.02
CACHE = os.environ.get("PFMDS_TUNING_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "pfmds_amd", "tuning.json"))


def signature(s, precision, n_rep=1):
    """Key of a workload: atom count, precision, and per interaction its kind
    and list shapes (group sizes, nmax, r_cut)."""
    parts = [f"n{s.n}", f"p{int(precision)}", f"r{n_rep}"]
    for it in s.interactions:
        for (g1, g2, nmax, rc, per) in it.lists:
            parts.append(f"{it.name}:{len(s.group_indices(g1))}x{len(s.group_indices(g2))}:{nmax}:{rc:.3f}:{per}")
    return "|".join(parts)


def valid(lanes, block, precision):
    return block <= 256 or int(precision) != 0


def time_shape(ctx, integrator, dt, first, steps):
    """Wall time of steps first..first+steps-1 (energy at the end forces the
    device to finish before the clock stops)."""
    t0 = time.perf_counter()
    ctx.run(integrator, dt, first, first + steps - 1, True)
    return time.perf_counter() - t0


def tune(ctx, integrator, dt, precision, first_step=1, steps=40, repeats=2, candidates=CANDIDATES, graphs=True):
    """Try every candidate shape on the handle, from step `first_step` on
    (the handle must have run step first_step - 1), set the fastest, then the
    faster of co-scheduled and separate launches, with separate launches the
    lanes of each interaction's passes, then the faster of graph and direct
    launches (direct only with graphs=False), and return (lanes, block,
    graphs, {setting: seconds}, next step); the choices of fusion and pass
    lanes are left on the handle as ctx.fused and ctx.pass_lanes."""
    table = {}
    step = first_step

    def best_of():
        nonlocal step
        best = None
        for _ in range(repeats):
            t = time_shape(ctx, integrator, dt, step, steps)
            step += steps
            best = t if best is None else min(best, t)
        return best

    ctx.set_graphs(False)
    for lanes, block in candidates:
        if not valid(lanes, block, precision):
            continue
        ctx.set_tuning(lanes, block)
        table[f"{lanes}x{block}"] = best_of()
    pick = min(table, key=table.get)
    lanes, block = (int(x) for x in pick.split("x"))
    ctx.set_tuning(lanes, block)
    # the co-scheduled passes of graphene on a metal (on by default) against
    # separate launches, at the picked shape
    ctx.set_fusion(False)
    table[f"{pick}-fusion"] = best_of()
    fused = table[pick] <= table[f"{pick}-fusion"]
    ctx.set_fusion(fused)
    pass_lanes = {}
    n_inter = getattr(ctx, "n_inter", 0)
    if not fused and n_inter > 1:
        best = table[f"{pick}-fusion"]
        for k in range(1, n_inter + 1):
            keep, keep_t = 0, best
            for pl in PASS_LANES:
                if pl == lanes:
                    continue
                ctx.set_pass_lanes(k, pl)
                t = best_of()
                table[f"{pick}-fusion:i{k}x{pl}"] = t
                if t < keep_t * (1.0 - PASS_MARGIN):
                    keep, keep_t = pl, t
            ctx.set_pass_lanes(k, keep)
            if keep:
                pass_lanes[k] = keep
                best = keep_t
    if graphs:
        ctx.set_graphs(True)
        table[f"{pick}+graphs"] = best_of()
        graphs = table[f"{pick}+graphs"] < table[pick]
    ctx.set_graphs(graphs)
    ctx.fused = fused
    ctx.pass_lanes = pass_lanes
    return lanes, block, graphs, table, step


def load(path=CACHE):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(key, lanes, block, graphs, table, path=CACHE, fusion=True, pass_lanes=None):
    data = load(path)
    data[key] = {"lanes": lanes, "block": block, "graphs": bool(graphs), "fusion": bool(fusion),
                 "pass_lanes": {str(k): int(v) for k, v in (pass_lanes or {}).items()}, "seconds": table}
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply_cached(ctx, key, path=CACHE):
    """Set the cached shape of `key` on the handle; False when none is cached."""
    e = load(path).get(key)
    if not e:
        return False
    ctx.set_tuning(int(e["lanes"]), int(e["block"]))
    ctx.set_graphs(bool(e.get("graphs", False)))
    ctx.set_fusion(bool(e.get("fusion", True)))
    for k, v in e.get("pass_lanes", {}).items():
        ctx.set_pass_lanes(int(k), int(v))
    return True


def autotune(ctx, s, integrator, dt, precision, first_step=1, n_rep=1, path=CACHE, steps=40, retune=False):
    """Cached shape for this workload if there is one, else tune and cache.
    Returns (lanes, block, graphs, source, next step) with source 'cache' or
    'tuned'."""
    key = signature(s, precision, n_rep)
    if not retune and apply_cached(ctx, key, path):
        e = load(path)[key]
        return int(e["lanes"]), int(e["block"]), bool(e.get("graphs", False)), "cache", first_step
    lanes, block, graphs, table, step = tune(ctx, integrator, dt, precision, first_step, steps)
    save(key, lanes, block, graphs, table, path, fusion=getattr(ctx, "fused", True),
         pass_lanes=getattr(ctx, "pass_lanes", None))
    return lanes, block, graphs, "tuned", step
