"""System description, the reference's file formats, and md()'s step schedule.

`System` holds what one PFMDS settings file describes (md_simulation.f90:48-93,
md_interactions.f90:38-136): atoms from the xyz file, groups by species name,
group roles, the integrator table, Nose-Hoover chains and interactions with
their neighbour-list rows.  `write_files`/`read_files` use the reference's
formats (settings grammar, extended xyz, one parameter file per interaction).

`run_md` restates the control flow of md()'s loop (md_simulation.f90:114-243):
integrator switching, the nvms exit, and the steps on which energies, log
lines, snapshots and trajectory frames happen.  It hands each run of steps
between two such events to a `capi.Context` (the HIP library, or in tests the
CPU oracle), exactly as the Fortran front end does.
"""
import dataclasses
import os
from typing import List, Optional

import numpy as np

from . import capi

# parameter counts and list rows of each interaction (md_interactions.f90:70-121)
PARAM_LINES = {
    "lj": [2, 2],
    "lj1g": [2, 2],
    "ljc": [3, 2, 1],
    "morsec": [4, 2, 1],
    "rjl": [5, 2],
    "tb": [8, 2],
    "rebosc": [3, 3, 3, 1, 6, 2],
}
NL_N = {"lj": 2, "lj1g": 1, "ljc": 3, "morsec": 3, "rjl": 1, "tb": 1, "rebosc": 1}
def _f(x):
    return repr(float(x))


MASS_COEF = 1.6605389217 / 1.6021765654 * 100.0
KB = 1.3806488 / 1.6021765654 * (1.0 / 10000.0)


@dataclasses.dataclass
class Interaction:
    name: str
    params: List[float]
    lists: List[tuple]  # (g1, g2, nmax, r_cut, update_period) per list row
    param_file: str = ""


@dataclasses.dataclass
class Integrator:
    name: str
    dt: float
    length: int
    period_snapshot: int
    period_log: int


@dataclasses.dataclass
class NHC:
    group: int
    temperature: float
    M: int
    q1: float


@dataclasses.dataclass
class System:
    box: np.ndarray
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    species: List[str]
    groups: List[List[str]]  # group g (1-based) = groups[g-1]
    interactions: List[Interaction]
    integrators: List[Integrator]
    nhcs: List[NHC] = dataclasses.field(default_factory=list)
    all_moving: int = 1
    xyz_moving: int = 1
    z_moving: int = 1
    all_atoms: int = 1
    traj: int = 1
    period_traj: int = 10**9
    zero_momentum_period: int = 10**9
    invert_z: bool = False
    new_velocities: bool = False
    ms_de: float = 1e-8
    initial_temperature: float = 0.0
    md_step_limit: int = 0
    logfilename: str = "log.txt"
    init_xyz_filename: str = "init.xyz"
    # runtime group changes (md_simulation.f90:63-70): (from, to, ts1, ts2, frec)
    changes: List[tuple] = dataclasses.field(default_factory=list)

    @property
    def n(self):
        return len(self.mass)

    def group_indices(self, g):
        """create_particle_group (md_general.f90:57-80): by listed name, then
        ascending atom index; 1-based."""
        sp = np.asarray(self.species)
        out = [np.nonzero(sp == name)[0] + 1 for name in self.groups[g - 1]]
        return np.concatenate(out).astype(np.int32) if out else np.zeros(0, np.int32)

    # -- file formats ----------------------------------------------------------
    def write_files(self, directory, settings="md_run_settings.txt", digits=17):
        os.makedirs(directory, exist_ok=True)
        write_xyz(os.path.join(directory, self.init_xyz_filename), self, digits)
        ntypes = max(len(g) for g in self.groups)
        lines = [
            f"md_step_limit {self.md_step_limit}",
            f"logfilename {self.logfilename}",
            f"init_xyz_filename {self.init_xyz_filename}",
            f"new_velocities {'T' if self.new_velocities else 'F'}",
            f"zero_momentum_period {self.zero_momentum_period}",
            f"particle_types_num {ntypes}",
            f"groups_num {len(self.groups)}",
        ]
        for i, g in enumerate(self.groups):
            names = list(g) + ["#"] * (ntypes - len(g))
            lines.append(f"group{i + 1} " + " ".join(names))
        lines += [
            f"all_moving_atoms_group_num {self.all_moving}",
            f"xyz_moving_atoms_group_num {self.xyz_moving}",
            f"z_moving_atoms_group_num {self.z_moving}",
            f"all_atoms_group_num {self.all_atoms}",
            f"traj_group_num {self.traj}",
            f"period_traj {self.period_traj}",
            f"change_group_num {len(self.changes)}",
        ]
        for (g_from, g_to, ts1, ts2, frec) in self.changes:
            lines += [f"group_change {g_from} {g_to}", f"change_steps {ts1} {ts2} {frec}"]
        lines += [
            f"invert_z_vel {'T' if self.invert_z else 'F'}",
            f"integrators_num {len(self.integrators)}",
            "name dt steps period_snapshot period_log",
        ]
        for it in self.integrators:
            lines.append(f"{it.name} {_f(it.dt)} {it.length} {it.period_snapshot} {it.period_log}")
        lines.append(f"ms_de {_f(self.ms_de)}")
        lines.append(f"nhc_num {len(self.nhcs)}")
        for c in self.nhcs:
            lines.append(f"{c.group} {_f(c.temperature)} {c.M} {_f(c.q1)}")
        lines.append(f"initial_temperature {_f(self.initial_temperature)}")
        lines.append(f"interactions_num {len(self.interactions)}")
        for k, it in enumerate(self.interactions):
            pf = it.param_file or f"{it.name}_{k + 1}.txt"
            lines.append(f"{it.name} {pf}")
            for (g1, g2, nmax, rc, per) in it.lists:
                lines.append(f"{g1} {g2} {nmax} {_f(rc)} {per}")
            write_params(os.path.join(directory, pf), it)
        with open(os.path.join(directory, settings), "w") as f:
            f.write("\n".join(lines) + "\n")
        return os.path.join(directory, settings)


def write_params(path, it):
    """Parameter file as read_*_parameters reads it (one line per read)."""
    vals = list(it.params)
    out = []
    for cnt in PARAM_LINES[it.name]:
        chunk, vals = vals[:cnt], vals[cnt:]
        if it.name in ("ljc", "morsec") and len(out) == 2:
            out.append("T" if chunk[0] else "F")
        else:
            out.append(" ".join(_f(v) for v in chunk))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_xyz(path, s, digits=17):
    """Extended xyz as read_box_size/read_particles read it (md_read_write.f90:22-61)."""
    b = s.box
    fmt = f"%.{digits}g"
    with open(path, "w") as f:
        f.write(f"{s.n}\n")
        f.write(f'Lattice=" {_f(b[0])} 0.0 0.0 0.0 {_f(b[1])} 0.0 0.0 0.0 {_f(b[2])} " '
                "Properties=pos:R:3:vel:R:3:mass:R:1:species:S:1\n")
        rows = np.column_stack([s.pos, s.vel, s.mass])
        for r, sp in zip(rows, s.species):
            f.write(" ".join(fmt % v for v in r) + f" {sp}\n")


def read_xyz(path):
    with open(path) as f:
        n = int(f.readline().split()[0])
        toks = f.readline().split()
        box = np.array([float(toks[1]), float(toks[5]), float(toks[9])])
        data = [f.readline().split() for _ in range(n)]
    arr = np.array([[float(x) for x in d[:7]] for d in data])
    return box, arr[:, 0:3].copy(), arr[:, 3:6].copy(), arr[:, 6].copy(), [d[7] for d in data]


def _fortran_logical(tok):
    t = tok.strip().lstrip(".").upper()
    return t.startswith("T")


def read_files(directory, settings="md_run_settings.txt"):
    """Parse a settings file the way md() reads it (md_simulation.f90:48-93);
    only line order matters, the leading name of each line is ignored."""
    with open(os.path.join(directory, settings)) as f:
        lines = [ln.rstrip("\n") for ln in f]
    it = iter(lines)

    def val():
        return next(it).split()[1:]

    md_step_limit = int(val()[0])
    logfilename = val()[0]
    init_xyz = val()[0]
    box, pos, vel, mass, species = read_xyz(os.path.join(directory, init_xyz))
    new_vel = _fortran_logical(val()[0])
    zmp = int(val()[0])
    ntypes = int(val()[0])
    ngroups = int(val()[0])
    groups = []
    for _ in range(ngroups):
        names = next(it).split()[1:1 + ntypes]
        groups.append([n for n in names])
    roles = [int(val()[0]) for _ in range(6)]
    nchange = int(val()[0])
    changes = []
    for _ in range(nchange):
        g_from, g_to = (int(x) for x in val()[:2])
        ts1, ts2, frec = (int(x) for x in val()[:3])
        changes.append((g_from, g_to, ts1, ts2, frec))
    invert = _fortran_logical(val()[0])
    nint = int(val()[0])
    next(it)
    integrators = []
    for _ in range(nint):
        t = next(it).split()
        integrators.append(Integrator(t[0], float(t[1]), int(t[2]), int(t[3]), int(t[4])))
    ms_de = float(val()[0])
    nnhc = int(val()[0])
    nhcs = []
    for _ in range(nnhc):
        t = next(it).split()
        nhcs.append(NHC(int(t[0]), float(t[1]), int(t[2]), float(t[3])))
    t0 = float(val()[0])
    ninter = int(val()[0])
    inters = []
    for _ in range(ninter):
        name, pf = next(it).split()[:2]
        rows = []
        for _ in range(NL_N[name]):
            t = next(it).split()
            rows.append((int(t[0]), int(t[1]), int(t[2]), float(t[3]), int(t[4])))
        with open(os.path.join(directory, pf)) as f:
            toks = f.read().split()
        params = [1.0 if _fortran_logical(x) else 0.0 if x.strip(".").upper().startswith("F") else float(x)
                  for x in toks]
        inters.append(Interaction(name, params, rows, pf))
    return System(box=box, pos=pos, vel=vel, mass=mass, species=species, groups=groups, interactions=inters,
                  integrators=integrators, nhcs=nhcs, all_moving=roles[0], xyz_moving=roles[1],
                  z_moving=roles[2], all_atoms=roles[3], traj=roles[4], period_traj=roles[5],
                  zero_momentum_period=zmp, invert_z=invert, new_velocities=new_vel, ms_de=ms_de,
                  initial_temperature=t0, md_step_limit=md_step_limit, logfilename=logfilename,
                  init_xyz_filename=init_xyz, changes=changes)


# -- building a context ---------------------------------------------------------

def build_context(s: System, make):
    """Create a Context with `make(box, pos, vel, mass)` and load groups, roles,
    chains and interactions in settings order (md_simulation.f90:53-97)."""
    if s.new_velocities:
        raise ValueError("new_velocities=T: write velocities into the xyz file (the reference's "
                         "rand()-based initialisation is not reproducible)")
    ctx = make(s.box, s.pos, s.vel, s.mass)
    for g in range(1, len(s.groups) + 1):
        ctx.set_group(g, s.group_indices(g))
    ctx.set_roles(s.all_moving, s.xyz_moving, s.z_moving, s.all_atoms, s.zero_momentum_period, s.invert_z)
    for c in s.nhcs:
        ctx.add_nhc(c.group, c.temperature, c.M, c.q1)
    for it in s.interactions:
        ctx.add_interaction(it.name, it.params, it.lists)
    for c in s.changes:
        ctx.add_group_change(*c)
    return ctx


@dataclasses.dataclass
class Batch:
    """Independent replicas stepped together in one context: `system` holds
    the concatenated atoms with batch-wide groups and interactions, replica r
    has n_atoms[r] atoms in box boxes[r]."""
    system: System
    n_atoms: np.ndarray
    boxes: np.ndarray

    @property
    def n_rep(self):
        return len(self.n_atoms)

    def replica_slice(self, r):
        a = int(np.sum(self.n_atoms[:r]))
        return slice(a, a + int(self.n_atoms[r]))


def merge_replicas(systems):
    """Concatenate systems into one Batch.  Groups are matched by their species
    lists, so a group of the batch holds the matching atoms of every replica;
    interactions equal in name, parameters and lists are shared.  Role groups
    that differ between replicas become the union of their species lists.
    All replicas must share the integrator schedule and periods; chains are
    not allowed (a chain's group would span replicas).  Group changes must be
    the same lines in every replica (after matching the groups); the batch
    applies them to each replica's own counts."""
    first = systems[0]
    gid = {}
    groups = []

    def group_of(names):
        key = tuple(names)
        if key not in gid:
            groups.append(list(names))
            gid[key] = len(groups)
        return gid[key]

    inter_keys, inters = {}, []
    for s in systems:
        if s.nhcs:
            raise ValueError("replicas of a batch cannot carry Nose-Hoover chains")
        if [(i.name, i.dt) for i in s.integrators] != [(i.name, i.dt) for i in first.integrators]:
            raise ValueError("replicas of a batch share one integrator schedule")
        if s.zero_momentum_period != first.zero_momentum_period or s.invert_z != first.invert_z:
            raise ValueError("replicas of a batch share zero_momentum_period and invert_z")
        for it in s.interactions:
            lists = tuple((group_of(s.groups[g1 - 1]), group_of(s.groups[g2 - 1]), nm, float(rc), per)
                          for g1, g2, nm, rc, per in it.lists)
            key = (it.name, tuple(float(x) for x in it.params), lists)
            if key not in inter_keys:
                inter_keys[key] = len(inters)
                inters.append(Interaction(it.name, list(it.params), [list(x) for x in lists]))

    def role(attr):
        names = []
        for s in systems:
            for n in s.groups[getattr(s, attr) - 1]:
                if n not in names:
                    names.append(n)
        return group_of(names)

    changes = None
    for s in systems:
        ch = [(group_of(s.groups[f - 1]), group_of(s.groups[t - 1]), ts1, ts2, fr) for f, t, ts1, ts2, fr in s.changes]
        if changes is None:
            changes = ch
        elif ch != changes:
            raise ValueError("replicas of a batch share their group change lines")
    roles = {a: role(a) for a in ("all_moving", "xyz_moving", "z_moving", "all_atoms", "traj")}
    batch = System(box=np.array(first.box, dtype=float), pos=np.concatenate([s.pos for s in systems]),
                   vel=np.concatenate([s.vel for s in systems]), mass=np.concatenate([s.mass for s in systems]),
                   species=[x for s in systems for x in s.species], groups=groups, interactions=inters,
                   integrators=list(first.integrators), period_traj=first.period_traj,
                   zero_momentum_period=first.zero_momentum_period, invert_z=first.invert_z, ms_de=first.ms_de,
                   md_step_limit=first.md_step_limit, changes=changes or [], **roles)
    return Batch(batch, np.array([s.n for s in systems], dtype=np.int32),
                 np.array([np.asarray(s.box, dtype=float) for s in systems]))


def build_batch_context(b: Batch, precision):
    """GPU context of a Batch (the C-ABI's pfm_create_batch)."""
    from . import capi
    return build_context(b.system, lambda box, p, v, m: capi.Context.gpu_batch(b.n_atoms, b.boxes, p, v, m,
                                                                              precision))


def _next_multiple(step, period):
    return ((step + period - 1) // period) * period


@dataclasses.dataclass
class LogRecord:
    integrator: str
    step: int
    time: float
    conserved: float
    e_nhc: float
    total: float
    pe: float
    ke: float
    temp: float
    e_inter: np.ndarray
    e_chains: np.ndarray


def run_md(s: System, ctx, out_period=1, on_event=None, step_limit=None):
    """md()'s loop (md_simulation.f90:114-243) over a Context.  Returns
    (log_records, final_step) where final_step is md()'s md_step at exit.
    `on_event(step, kinds)` is called after each event step with the set of
    outputs due there ({'log','out','snapshot','traj'})."""
    limit = s.md_step_limit if step_limit is None else step_limit
    ints = [Integrator("none", 0.0, 0, 1, 1)] + list(s.integrators)
    cum = np.cumsum([x.length for x in ints])
    idx = 0
    name, dt = "none", 0.0
    records = []
    step = 0
    while step <= limit:
        # integrator switch (md_simulation.f90:121-136)
        if idx == 0 or step - 1 == cum[idx]:
            idx += 1
            i = idx
            while i < len(ints) and ints[i].length <= 0:
                i += 1
            if i >= len(ints):
                return records, step  # 'no more integrators'
            idx = i
            name, dt = ints[idx].name, ints[idx].dt
        cur = ints[idx]
        end = max(min(int(cum[idx]), limit), step)
        cands = [end,
                 _next_multiple(step, cur.period_log),
                 _next_multiple(step, out_period),
                 _next_multiple(step, cur.period_snapshot),
                 _next_multiple(step, s.period_traj)]
        t = min(cands)
        energy = (t % cur.period_log == 0) or (t % out_period == 0)
        last, ev = ctx.run(name, dt, step, t, energy, s.ms_de)
        if ev == capi.EV_CONVERGED:
            return records, last
        kinds = set()
        if t % cur.period_log == 0:
            kinds.add("log")
        if t % out_period == 0:
            kinds.add("out")
        if t % cur.period_snapshot == 0 and t != 0:
            kinds.add("snapshot")
        if t % s.period_traj == 0:
            kinds.add("traj")
        if "log" in kinds:
            e = ctx.energies()
            enhc = float(np.sum(e["e_nhc"])) if len(e["e_nhc"]) else 0.0
            total = e["pe"] + e["ke"]
            records.append(LogRecord(name, t, e["time"], total + enhc, enhc, total, e["pe"], e["ke"], e["temp"],
                                     e["e_inter"].copy(), e["e_nhc"].copy()))
        if on_event and kinds:
            on_event(t, kinds)
        step = t + 1
    return records, step
