"""A deposition-style system for the runtime group growth tests: substrate
atoms A and atoms D that join the moving group one at a time
(change_particle_group_N, md_general.f90:82-94).  Group 2 (A then D, in
index order) grows from group 1's count: one more atom at step ts1, then one
every frec steps before ts2.  Group 2 is all_moving and xyz_moving and the
group of an lj interaction, so atoms that have not joined neither move nor
interact."""
import numpy as np

from pfmds_amd import configs
from pfmds_amd.system import Integrator, Interaction, System


def deposition(n_side=4, spacing=4.2, seed=91, ts1=5, ts2=40, frec=7, period=10, temperature=150.0,
               kind="lj"):
    rng = configs.SplitMix64(seed)
    g = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (g + 0.5) * spacing + 0.15 * (2 * rng.uniform(3 * len(g)).reshape(-1, 3) - 1)
    n = len(pos)
    L = n_side * spacing
    species = ["A" if i < n // 2 else "D" for i in range(n)]
    if kind == "lj":
        mass = np.full(n, 39.95)
        inter = Interaction("lj", [0.0104, 3.40, 6.5, 7.5], [(2, 2, 96, 8.5, period), (2, 2, 96, 8.5, period)])
    elif kind == "lj1g":
        mass = np.full(n, 39.95)
        inter = Interaction("lj1g", [0.0104, 3.40, 6.5, 7.5], [(2, 2, 96, 8.5, period)])
    else:
        mass = np.full(n, 63.546)
        inter = Interaction("rjl", list(configs.RJL_CU), [(2, 2, 96, configs.RJL_CU[6] + 0.5, period)])
    vel = configs.maxwell(rng, mass, temperature)
    return System(box=np.array([L, L, L]), pos=np.mod(pos, L), vel=vel, mass=mass, species=species,
                  groups=[["A"], ["A", "D"], ["A", "D"], ["#"]], interactions=[inter],
                  integrators=[Integrator("nve", 1.0, 60, 10**9, 10**9)], all_moving=2, xyz_moving=2, z_moving=4,
                  all_atoms=3, traj=2, period_traj=10**9, zero_momentum_period=10**9, md_step_limit=60,
                  changes=[(1, 2, ts1, ts2, frec)])


def expected_counts(steps, n_from, size, ts1, ts2, frec):
    """Group 2's active count after each step, written out from the rule:
    n_from up to ts1 - 1, n_from + 1 at ts1, then +1 at ts1 + k*frec < ts2."""
    out = []
    for s in steps:
        if s < ts1:
            c = n_from
        else:
            c = n_from + 1 + len([t for t in range(ts1 + frec, min(s, ts2 - 1) + 1, frec)])
        out.append(min(c, size))
    return out


def carbon_growth_tb(ts1=3, ts2=40, frec=4, seed=95):
    """A free graphene sheet with tb only; its second half (species C2, after
    C1 in group order) joins group 1 one atom at a time.  Group 1 is the tb
    group and the moving group, so atoms that have not joined neither move nor
    bond."""
    g = configs.graphene_on_cu(cx=6, cy=4, gx=6, gy=4, layers=3, seed=seed, temperature=200.0)
    c = np.array([sp == "C" for sp in g.species])
    pos, vel, mass = g.pos[c], g.vel[c], g.mass[c]
    n = len(pos)
    species = ["C1" if i < n // 2 else "C2" for i in range(n)]
    inter = Interaction("tb", list(configs.TB_BRENNER_I), [(1, 1, 8, configs.TB_BRENNER_I[9] + 0.3, 10)])
    return System(box=g.box.copy(), pos=pos, vel=vel, mass=mass, species=species,
                  groups=[["C1", "C2"], ["C1"], ["#"]], interactions=[inter],
                  integrators=[Integrator("nve", 0.5, 60, 10**9, 10**9)], all_moving=1, xyz_moving=1, z_moving=3,
                  all_atoms=1, traj=1, period_traj=10**9, zero_momentum_period=10**9, md_step_limit=60,
                  changes=[(2, 1, ts1, ts2, frec)])


def metal_adatoms(ts1=3, ts2=40, frec=4, seed=96, simplified=False):
    """Graphene on Cu(111) with rjl, tb and morsec, whose top Cu layer
    (species Cu_ad, last in group 2's order) joins the metal group one atom at
    a time: group 2 is the rjl group and the metal group of morsec.  The
    adatoms stay out of the moving group."""
    s = configs.graphene_on_cu(cx=6, cy=4, gx=6, gy=4, layers=4, seed=seed, temperature=200.0)
    z = s.pos[:, 2]
    cu = np.array([sp != "C" for sp in s.species])
    top = cu & (z > z[cu].max() - 0.1)
    s.species = ["Cu_ad" if top[i] else s.species[i] for i in range(s.n)]
    # groups: 1 C, 2 Cu_fixed Cu Cu_ad (grows from 7), 3 moving, 4 all, 5 Cu, 6 #, 7 Cu_fixed Cu
    s.groups = [["C"], ["Cu_fixed", "Cu", "Cu_ad"], ["Cu", "C"], ["Cu_fixed", "Cu", "Cu_ad", "C"], ["Cu"], ["#"],
                ["Cu_fixed", "Cu"]]
    if simplified:
        s.interactions[2].params[6] = 1.0
    s.integrators = [Integrator("nve", 0.5, 60, 10**9, 10**9)]
    s.nhcs = []
    s.md_step_limit = 60
    s.changes = [(7, 2, ts1, ts2, frec)]
    return s


def carbon_growth_with_cos(seed=97):
    """Graphene on Cu with tb and morsec whose graphene group grows from an
    empty group: the members that have not joined have no nearest neighbours
    in the tb list, which stops the run at step 0 (graphenenorm.f90:30)."""
    s = configs.graphene_on_cu(cx=6, cy=4, gx=6, gy=4, layers=3, seed=seed)
    s.integrators = [Integrator("nve", 0.5, 20, 10**9, 10**9)]
    s.nhcs = []
    s.md_step_limit = 20
    s.changes = [(6, 1, 2, 20, 3)]
    return s
