"""Seeded generators of the benchmark and test systems, in PFMDS's own terms.

Random numbers come from SplitMix64 (53-bit uniforms), Gaussians from
Box-Muller; velocities are written into the system so `new_velocities=F`
(the reference's rand()-based initialisation, md_general.f90:114-139, is not
reproducible).

Workloads (sizes as named in the project's benchmark table):
  binary_lj_gas   4,000-atom A/B Lennard-Jones gas, 60 A periodic cube
                  (lj1g A-A, lj1g B-B, lj A-B), nvt / nve / nvms schedule
  fcc_cu_rgl      fcc Cu, nx*ny*nz conventional cells (32^3 = 131,072 atoms),
                  Rosato-Guillope-Legrand (rjl), nve at 1 fs
  graphene_on_cu  Cu(111) slab (6 layers of 100x58 rectangular cells, bottom
                  two frozen) + 104x60-cell graphene sheet 3.2 A above it:
                  rjl (Cu), tb (C-C), morsec (C-Cu)
Each takes scale arguments so tests can build small instances of the same
construction.

Parameters the workload description leaves open, fixed here:
  rjl   R1 = 5.5 A (switching over the fifth fcc shell), R2 = 6.0 A
  tb    Brenner (1990) parameter set I: D=6.325 eV, S=1.29, beta=1.5 1/A,
        r0=1.315 A, delta=0.80469, a0=0.011304, c0=19, d0=2.5, R1=1.7, R2=2.0 A
  morsec delta = 2, R1 = 6.0 A, R2 = 7.0 A, simplified = F
  lists r_cut = R2 + skin, update_period 10 (20 for the gas)
"""
import os

import numpy as np

from .system import KB, MASS_COEF, NHC, Integrator, Interaction, System

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based SplitMix64: u_k = mix(seed + (k+1)*golden)."""

    def __init__(self, seed):
        self.state = np.uint64(seed)
        self.count = np.uint64(0)

    def uniform(self, n):
        k = np.arange(1, n + 1, dtype=np.uint64) + self.count
        self.count += np.uint64(n)
        with np.errstate(over="ignore"):
            z = self.state + k * GOLDEN
            z = (z ^ (z >> np.uint64(30))) * M1
            z = (z ^ (z >> np.uint64(27))) * M2
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        u1 = np.where(u1 <= 0.0, 1e-300, u1)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def maxwell(rng, mass, temperature, moving=None):
    """Maxwell velocities (A/fs), zero total momentum of the moving atoms, scaled
    to `temperature` with the reference's 3N degrees of freedom (md_general.f90:301-311)."""
    n = len(mass)
    moving = np.ones(n, bool) if moving is None else moving
    v = rng.normal(3 * n).reshape(n, 3) * np.sqrt(KB * temperature / (mass * MASS_COEF))[:, None]
    v[~moving] = 0.0
    if temperature <= 0:
        return np.zeros((n, 3))
    m = mass[moving]
    v[moving] -= (m[:, None] * v[moving]).sum(0) / m.sum()
    ke = 0.5 * MASS_COEF * (m[:, None] * v[moving] ** 2).sum()
    t = 2 * ke / KB / (3 * moving.sum())
    v[moving] *= np.sqrt(temperature / t)
    return v


# ----------------------------------------------------------------------- LJ gas

def binary_lj_gas(n=4000, box=60.0, seed=1000, rmin=3.0, temperature=100.0, r1=6.5, r2=7.5, skin=1.0,
                  nmax=160, period=20, schedule=((("nvt", 0.5, 1000, 10**9, 100)), ("nve", 0.5, 1000, 10**9, 100),
                                                 ("nvms", 0.5, 100000, 10**9, 100))):
    rng = SplitMix64(seed)
    L = float(box)
    pts = np.zeros((n, 3))
    cnt = 0
    while cnt < n:
        c = rng.uniform(3) * L
        if cnt:
            d = pts[:cnt] - c
            d -= L * np.round(d / L)
            if (d * d).sum(1).min() < rmin * rmin:
                continue
        pts[cnt] = c
        cnt += 1
    species = ["A" if i < n // 2 else "B" for i in range(n)]
    mass = np.where(np.array(species) == "A", 39.95, 83.80)
    vel = maxwell(rng, mass, temperature)
    eA, sA, eB, sB = 0.0104, 3.40, 0.0140, 3.65
    eAB, sAB = float(np.sqrt(eA * eB)), 0.5 * (sA + sB)
    rc = r2 + skin
    inters = [
        Interaction("lj1g", [eA, sA, r1, r2], [(1, 1, nmax, rc, period)]),
        Interaction("lj1g", [eB, sB, r1, r2], [(2, 2, nmax, rc, period)]),
        Interaction("lj", [eAB, sAB, r1, r2], [(1, 2, nmax, rc, period), (2, 1, nmax, rc, period)]),
    ]
    q1 = 3 * n * KB * temperature * 100.0 ** 2
    ints = [Integrator(*x) for x in schedule]
    return System(box=np.array([L, L, L]), pos=pts, vel=vel, mass=mass, species=species,
                  groups=[["A"], ["B"], ["A", "B"], ["#"]], interactions=inters, integrators=ints,
                  nhcs=[NHC(3, temperature, 3, q1)], all_moving=3, xyz_moving=3, z_moving=4, all_atoms=3,
                  traj=3, period_traj=10**9, zero_momentum_period=10**9, ms_de=1e-8,
                  md_step_limit=sum(x.length for x in ints), init_xyz_filename="binary_lj_gas.xyz")


# ------------------------------------------------------------------------ fcc Cu

RJL_CU = [0.0855, 1.224, 10.960, 2.278, 2.5562, 5.5, 6.0]


def fcc_cu_rgl(nx=32, ny=None, nz=None, a=3.615, seed=1001, sigma=0.02, temperature=300.0, skin=0.5, nmax=96,
               period=10, steps=1000, dt=1.0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    rng = SplitMix64(seed)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    pos = ((g[:, None, :] + base[None]) * a).reshape(-1, 3)
    box = np.array([nx * a, ny * a, nz * a])
    pos = pos + sigma * rng.normal(pos.size).reshape(-1, 3)
    pos = np.mod(pos, box)
    n = len(pos)
    mass = np.full(n, 63.546)
    vel = maxwell(rng, mass, temperature)
    inters = [Interaction("rjl", list(RJL_CU), [(1, 1, nmax, RJL_CU[6] + skin, period)])]
    return System(box=box, pos=pos, vel=vel, mass=mass, species=["Cu"] * n, groups=[["Cu"], ["#"]],
                  interactions=inters, z_moving=2, integrators=[Integrator("nve", dt, steps, 10**9, 10**9)],
                  md_step_limit=steps, init_xyz_filename="fcc_cu.xyz")


# ---------------------------------------------------------------- graphene on Cu

TB_BRENNER_I = [6.325, 1.29, 1.5, 1.315, 0.80469, 0.011304, 19.0, 2.5, 1.7, 2.0]
MORSEC_CCU = [0.025, 3.05, 1.30, 2.0, 6.0, 7.0, 0.0]
# rebosc (REBOsolidcarbon.f90:17-22): A Q alpha / B(3) / beta(3) / T / g(6) / R1 R2.
# The C-C pair terms of Brenner's second-generation REBO (J. Phys.: Condens.
# Matter 14, 783 (2002), Table 2), its torsion prefactor, and the quintic
# g(cos) through Brenner's G_C knots (-1, -2/3, -1/2, -1/3, 1) and g(0) = 0.375.
REBO_C = [10953.544162170, 0.3134602960833, 4.7465390606595,
          12388.79197798, 17.56740646509, 30.71493208065,
          4.7204523127, 1.4332132499, 1.3826912506,
          -0.004048375,
          0.375, 1.493091717, 2.689174625, 2.427981333, 0.930825375, 0.08392695,
          1.7, 2.0]


def graphene_on_cu(cx=100, cy=58, layers=6, frozen=2, gx=104, gy=60, seed=1002, acu=2.5562, gap=3.2, lz=40.0,
                   z0=2.0, jitter=0.05, temperature=300.0, period=10, nvt_steps=5000, dt=0.5):
    """Cu(111) slab of `layers` layers of cx*cy rectangular (a x a*sqrt3) cells
    with 2 atoms each, ABC stacked, and a graphene sheet of gx*gy rectangular
    4-atom cells (zigzag along x) strained to the same box."""
    rng = SplitMix64(seed)
    ay = acu * np.sqrt(3.0)
    Lx, Ly = cx * acu, cy * ay
    dz = acu * np.sqrt(2.0 / 3.0)  # (111) spacing = a_fcc/sqrt(3)
    cell = np.array([[0.0, 0.0], [0.5 * acu, 0.5 * ay]])
    shifts = [np.array([0.0, 0.0]), np.array([0.5 * acu, ay / 6.0]), np.array([0.0, ay / 3.0])]
    ij = np.stack(np.meshgrid(np.arange(cx), np.arange(cy), indexing="ij"), -1).reshape(-1, 2)
    layer_xy = (ij[:, None, :] * np.array([acu, ay]) + cell[None]).reshape(-1, 2)
    cu, cu_species = [], []
    for k in range(layers):
        xy = np.mod(layer_xy + shifts[k % 3], [Lx, Ly])
        cu.append(np.column_stack([xy, np.full(len(xy), z0 + k * dz)]))
        cu_species += ["Cu_fixed" if k < frozen else "Cu"] * len(xy)
    cu = np.concatenate(cu)
    # graphene: rectangular cell (sqrt3*acc) x (3*acc), strained to (Lx/gx, Ly/gy)
    sx, sy = Lx / gx, Ly / gy
    gcell = np.array([[0.0, 0.0], [0.5, 1.0 / 6.0], [0.5, 0.5], [0.0, 2.0 / 3.0]]) * np.array([sx, sy])
    gij = np.stack(np.meshgrid(np.arange(gx), np.arange(gy), indexing="ij"), -1).reshape(-1, 2)
    gxy = (gij[:, None, :] * np.array([sx, sy]) + gcell[None]).reshape(-1, 2)
    off = rng.uniform(2) * np.array([sx, sy])
    gxy = np.mod(gxy + off, [Lx, Ly])
    ztop = z0 + (layers - 1) * dz
    gz = ztop + gap + jitter * (2.0 * rng.uniform(len(gxy)) - 1.0)
    gr = np.column_stack([gxy, gz])
    pos = np.concatenate([cu, gr])
    species = cu_species + ["C"] * len(gr)
    mass = np.array([12.011 if s == "C" else 63.546 for s in species])
    moving = np.array([s != "Cu_fixed" for s in species])
    vel = maxwell(rng, mass, temperature, moving)
    box = np.array([Lx, Ly, lz])
    groups = [["C"], ["Cu_fixed", "Cu"], ["Cu", "C"], ["Cu_fixed", "Cu", "C"], ["Cu"], ["#"]]
    inters = [
        # fcc Cu holds 86 neighbours within 6.5 A at rest; thermal motion in
        # long runs takes single atoms past 96
        Interaction("rjl", list(RJL_CU), [(2, 2, 128, RJL_CU[6] + 0.5, period)]),
        Interaction("tb", list(TB_BRENNER_I), [(1, 1, 8, TB_BRENNER_I[9] + 0.3, period)]),
        Interaction("morsec", list(MORSEC_CCU),
                    [(1, 2, 96, MORSEC_CCU[5] + 0.5, period), (2, 1, 128, MORSEC_CCU[5] + 0.5, period),
                     (1, 1, 3, 1.8, period)]),
    ]
    nmov = int(moving.sum())
    q1 = 3 * nmov * KB * temperature * 100.0 ** 2
    ints = [Integrator("nvms", dt, 100000, 10**9, 10**9), Integrator("nvt", dt, nvt_steps, 10**9, 10**9)]
    return System(box=box, pos=pos, vel=vel, mass=mass, species=species, groups=groups, interactions=inters,
                  integrators=ints, nhcs=[NHC(3, temperature, 3, q1)], all_moving=3, xyz_moving=3, z_moving=6,
                  all_atoms=4, traj=1, period_traj=10**9, zero_momentum_period=10**9, ms_de=1e-8,
                  md_step_limit=100000 + nvt_steps, init_xyz_filename="graphene_on_cu.xyz")


def moire_fit_case(directory, cells=((10, 6, 10, 6), (12, 6, 12, 6)), layers=4, steps=200, period=10, be0=-0.05,
                   grd0=(0.2, 0.2), spread=0.1, seed=1006):
    """Input folder of a morsec moire fit (fit_gr_moire.f90:106-184) over two
    graphene-on-Cu(111) cells (cx, cy, gx, gy each), relaxed by `steps` steps
    of nvms: settings, start xyz and morsec parameter files per system, the
    bounds (MORSEC_CCU times 1 -+ spread) and the fitting-parameters file,
    whose path is returned.  The metal is named CU / CU_fixed as the analysis
    expects; zero energy levels are 0."""
    os.makedirs(directory, exist_ok=True)
    rename = {"Cu": "CU", "Cu_fixed": "CU_fixed"}
    vals = []
    for k, (cx, cy, gx, gy) in enumerate(cells):
        s = graphene_on_cu(cx=cx, cy=cy, gx=gx, gy=gy, layers=layers, seed=seed + k, period=period)
        s.species = [rename.get(x, x) for x in s.species]
        s.groups = [[rename.get(x, x) for x in g] for g in s.groups]
        s.integrators = [Integrator("nvms", 0.5, steps, 10**9, max(1, steps // 4))]
        s.nhcs = []
        s.md_step_limit = steps
        s.init_xyz_filename = f"gr_cu_{k + 1}.xyz"
        s.logfilename = f"log_{k + 1}.txt"
        for it in s.interactions:
            it.param_file = f"{it.name}.txt"
        s.write_files(directory, settings=f"settings_{k + 1}.txt")
        os.replace(os.path.join(directory, s.init_xyz_filename), os.path.join(directory, f"start_{k + 1}.xyz"))
        vals.append((f"settings_{k + 1}.txt", f"start_{k + 1}.xyz", s.species.count("C"), 0.0, f"final_{k + 1}.txt"))
    d, r, a, delt, r1, r2, _ = MORSEC_CCU
    for name, f in (("morsec_min.txt", 1 - spread), ("morsec_max.txt", 1 + spread)):
        with open(os.path.join(directory, name), "w") as fh:
            fh.write(f"{d * f!r} {r * f!r} {a!r} {delt * f!r}\n{r1!r} {r2!r}\nF\n")
    path = os.path.join(directory, "fitting_parameters.txt")
    with open(path, "w") as fh:
        fh.write(f"{'input_path:':<16}{os.path.join(directory, '')}\n{'out_path:':<16}{os.path.join(directory, '')}\n")
        fh.write("interaction_name morsec\noutput_prefix fit\n")
        for k, (st, x, c, z0, fin) in enumerate(vals):
            fh.write(f"settings_{k + 1} {st}\nstart_xyz_{k + 1} {x}\nc_num_{k + 1} {c}\nzero_energy_{k + 1} {z0!r}\n"
                     f"final_file_{k + 1} {fin}\n")
        fh.write(f"min_param_file morsec_min.txt\nmax_param_file morsec_max.txt\nz 0.0\nbe0 {be0!r}\n"
                 f"grd0_1 {grd0[0]!r}\ngrd0_2 {grd0[1]!r}\n")
    return path


def graphene_rebo(gx=8, gy=5, acc=1.42, lz=20.0, jitter=0.0, buckle=0.0, seed=1005, temperature=300.0, dt=0.5,
                  period=10):
    """Free-standing graphene sheet of gx*gy rectangular 4-atom cells with
    rebosc: seeded in-plane jitter and out-of-plane buckling amplitudes (A),
    Maxwell velocities."""
    rng = SplitMix64(seed)
    sx, sy = np.sqrt(3.0) * acc, 3.0 * acc
    cell = np.array([[0.0, 0.0], [0.5, 1.0 / 6.0], [0.5, 0.5], [0.0, 2.0 / 3.0]]) * np.array([sx, sy])
    ij = np.stack(np.meshgrid(np.arange(gx), np.arange(gy), indexing="ij"), -1).reshape(-1, 2)
    xy = (ij[:, None, :] * np.array([sx, sy]) + cell[None]).reshape(-1, 2) + 0.3
    n = len(xy)
    pos = np.column_stack([xy, np.full(n, 0.5 * lz)])
    pos[:, :2] += jitter * (2 * rng.uniform(2 * n).reshape(n, 2) - 1)
    pos[:, 2] += buckle * (2 * rng.uniform(n) - 1)
    box = np.array([gx * sx, gy * sy, lz])
    mass = np.full(n, 12.011)
    vel = maxwell(rng, mass, temperature)
    return System(box=box, pos=np.mod(pos, box), vel=vel, mass=mass, species=["C"] * n, groups=[["C"], ["#"]],
                  interactions=[Interaction("rebosc", list(REBO_C), [(1, 1, 8, REBO_C[17] + 0.3, period)])],
                  integrators=[Integrator("nve", dt, 1000, 10**9, 10**9)], z_moving=2)


# ------------------------------------------------------------- replica batches

def lj_gas_lattice(n, density, seed, temperature, dt=0.5, steps=2000, r1=6.5, r2=7.5, skin=1.0, nmax=160,
                   period=20, jitter=0.1):
    """Binary LJ gas (the interactions of binary_lj_gas) of n atoms at
    `density` (A^-3), on a jittered simple cubic lattice whose sites the two
    species take in random order: spacing
    (1/density)**(1/3) - 2*jitter stays above binary_lj_gas's 3.0 A minimum
    for density <= 0.03, and the start costs O(n)."""
    rng = SplitMix64(seed)
    L = float((n / density) ** (1.0 / 3.0))
    m = int(np.ceil(n ** (1.0 / 3.0)))
    a = L / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    pick = np.argsort(rng.uniform(len(g)), kind="stable")[:n]
    pos = (g[np.sort(pick)] + 0.5) * a + jitter * (2.0 * rng.uniform(3 * n).reshape(n, 3) - 1.0)
    pos = np.mod(pos, L)
    # the lattice sites in a seeded random order: the first n/2 atoms (A)
    # and the rest (B) are mixed through the box, as in binary_lj_gas
    pos = pos[np.argsort(rng.uniform(n), kind="stable")]
    species = ["A" if i < n // 2 else "B" for i in range(n)]
    mass = np.where(np.array(species) == "A", 39.95, 83.80)
    vel = maxwell(rng, mass, temperature)
    eA, sA, eB, sB = 0.0104, 3.40, 0.0140, 3.65
    eAB, sAB = float(np.sqrt(eA * eB)), 0.5 * (sA + sB)
    rc = r2 + skin
    inters = [
        Interaction("lj1g", [eA, sA, r1, r2], [(1, 1, nmax, rc, period)]),
        Interaction("lj1g", [eB, sB, r1, r2], [(2, 2, nmax, rc, period)]),
        Interaction("lj", [eAB, sAB, r1, r2], [(1, 2, nmax, rc, period), (2, 1, nmax, rc, period)]),
    ]
    return System(box=np.array([L, L, L]), pos=pos, vel=vel, mass=mass, species=species,
                  groups=[["A"], ["B"], ["A", "B"], ["#"]], interactions=inters,
                  integrators=[Integrator("nve", dt, steps, 10**9, 10**9)], all_moving=3, xyz_moving=3, z_moving=4,
                  all_atoms=3, traj=3, md_step_limit=steps, init_xyz_filename="lj_gas.xyz")


def replica_mix(n_rep=512, seed=1003, n_min=256, n_max=32768, steps=2000, dt=0.5, only=None):
    """Independent replicas of mixed size and kind: N_k = round(exp U(ln n_min,
    ln n_max)); even k a binary LJ gas at T ~ U(50, 300) K and density
    ~ U(0.01, 0.03) A^-3, odd k fcc Cu (RGL) of nx*ny*nz*4 ~ N_k atoms with at
    least four cells per side (list nmax 128) at T ~ U(300, 1200) K; replica k is seeded
    seed*10**4 + k.  All step with nve at dt.  Returns the list of Systems
    (with `only` = k, replica k alone and None for the others)."""
    rng = SplitMix64(seed)
    u = rng.uniform(4 * n_rep).reshape(n_rep, 4)
    out = []
    for k in range(n_rep):
        if only is not None and k != only:
            out.append(None)
            continue
        nk = int(round(np.exp(np.log(n_min) + u[k, 0] * (np.log(n_max) - np.log(n_min)))))
        sk = seed * 10**4 + k
        if k % 2 == 0:
            t = 50.0 + 250.0 * u[k, 1]
            rho = 0.01 + 0.02 * u[k, 2]
            s = lj_gas_lattice(max(nk, 2), rho, sk, t, dt=dt, steps=steps)
        else:
            t = 300.0 + 900.0 * u[k, 1]
            c = max(4, int(round((nk / 4.0) ** (1.0 / 3.0))))
            nz = max(4, int(round(nk / (4.0 * c * c))))
            # nmax 128: up to 1200 K the density fluctuations put more than
            # 96 atoms inside the 6.5 A list cutoff
            s = fcc_cu_rgl(nx=c, ny=c, nz=nz, seed=sk, temperature=t, steps=steps, dt=dt, nmax=128)
        out.append(s)
    return out
