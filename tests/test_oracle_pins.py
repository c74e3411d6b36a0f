"""Pins of the CPU oracle (oracle/liborc.so) before it is trusted.

The reference ships no golden vectors, fixtures or tests, and it cannot be
built here (it needs the GNU rand/srand extension that amdflang lacks), so
every pin comes from what the mathematics fixes:
  - neighbour sets = brute force over all pairs with numpy; minimum image =
    nearest of the 27 periodic images;
  - forces = -grad E by torch autograd of independently written energies
    (tests/ref_torch.py), for every potential; lj1g's cutoff quirk as the
    linear combination A + (R2-R1) B of two autograd terms;
  - energies = closed forms (LJ dimer, RGL perfect fcc shell sums, TB
    perfect graphene) and the autograd energies;
  - integrators: NVE time reversibility and energy conservation, the NHC
    conserved quantity, nvms convergence to the analytic LJ dimer minimum;
  - unit constants.
A mistake planted in the oracle (tests/test_oracle_mutants.py) fails these.
"""
import numpy as np
import pytest
import torch

from pfmds_amd import configs
from pfmds_amd.system import KB, MASS_COEF, NHC, Integrator, Interaction, System
from tests import ref_torch as R
from tests.harness import list_sets, oracle_ctx, rel_inf


def graphene_sheet(gx=6, gy=4, acc=1.42, lz=20.0, jitter=0.0, seed=5):
    sx, sy = np.sqrt(3.0) * acc, 3.0 * acc
    cell = np.array([[0.0, 0.0], [0.5, 1.0 / 6.0], [0.5, 0.5], [0.0, 2.0 / 3.0]]) * np.array([sx, sy])
    ij = np.stack(np.meshgrid(np.arange(gx), np.arange(gy), indexing="ij"), -1).reshape(-1, 2)
    xy = (ij[:, None, :] * np.array([sx, sy]) + cell[None]).reshape(-1, 2) + 0.3
    rng = configs.SplitMix64(seed)
    pos = np.column_stack([xy, np.full(len(xy), 0.5 * lz)])
    pos += jitter * (2 * rng.uniform(pos.size).reshape(-1, 3) - 1)
    n = len(pos)
    box = np.array([gx * sx, gy * sy, lz])
    return System(box=box, pos=np.mod(pos, box), vel=np.zeros((n, 3)), mass=np.full(n, 12.011),
                  species=["C"] * n, groups=[["C"], ["#"]],
                  interactions=[Interaction("tb", list(configs.TB_BRENNER_I), [(1, 1, 8, 2.3, 10)])],
                  integrators=[Integrator("nve", 0.5, 10, 10**9, 10**9)], z_moving=2)


def step0(s):
    o = oracle_ctx(s)
    o.run(s.integrators[0].name, s.integrators[0].dt, 0, 0, True)
    return o


# ------------------------------------------------------------------ constants

def test_unit_constants():
    # 1 u*A^2/fs^2 in eV and k_B in eV/K, from the SI values of CODATA 2010
    # (m_u = 1.660538921e-27 kg, e = 1.602176565e-19 C, k_B = 1.3806488e-23 J/K)
    m_u, e, kb = 1.660538921e-27, 1.602176565e-19, 1.3806488e-23
    assert MASS_COEF == pytest.approx(m_u * 1e-20 / 1e-30 / e, rel=1e-9)
    assert KB == pytest.approx(kb / e, rel=1e-9)


# -------------------------------------------------------------- neighbour sets

def brute_sets(pos, box, g1, g2, rc):
    """{j : j in g2, j != i, dr2 < rc^2} per i in g1, group order, numpy."""
    g1, g2 = np.asarray(g1), np.asarray(g2)
    d = pos[g2][None, :, :] - pos[g1][:, None, :]
    h = 0.5 * box
    d = d - h * (np.copysign(1.0, d - h) + np.copysign(1.0, d + h))
    d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
    ok = (d2 < rc * rc) & (g1[:, None] != g2[None, :])
    return [tuple((np.nonzero(ok[a])[0] + 1).tolist()) for a in range(len(g1))]


def test_min_image_is_nearest_image():
    rng = np.random.default_rng(3)
    box = np.array([11.0, 13.0, 17.0])
    a = rng.uniform(0, 1, (400, 3)) * box
    b = rng.uniform(0, 1, (400, 3)) * box
    d = b - a
    h = 0.5 * box
    dm = d - h * (np.copysign(1.0, d - h) + np.copysign(1.0, d + h))
    shifts = np.stack(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij"), -1).reshape(-1, 3)
    best = np.min(((d[:, None, :] + shifts[None] * box) ** 2).sum(-1), axis=1)
    np.testing.assert_allclose((dm ** 2).sum(-1), best, rtol=1e-14)
    # ties: dr = +L/2 maps to -L/2, dr = -L/2 stays (sign(1,0) = +1)
    hh = h[0]
    assert hh - hh * (np.copysign(1.0, 0.0) + np.copysign(1.0, 2 * hh)) == -hh
    assert -hh - hh * (np.copysign(1.0, -2 * hh) + np.copysign(1.0, 0.0)) == -hh


@pytest.mark.parametrize("which", ["lj_gas", "cu", "gr_cu"])
def test_lists_brute_force(which):
    s = {"lj_gas": lambda: configs.binary_lj_gas(n=300, box=26.0, seed=21),
         "cu": lambda: configs.fcc_cu_rgl(nx=4, seed=22),
         "gr_cu": lambda: configs.graphene_on_cu(cx=8, cy=5, gx=8, gy=5, layers=3, seed=23)}[which]()
    o = step0(s)
    for k, it in enumerate(s.interactions, start=1):
        g1, g2, nmax, rc, _ = it.lists[0]
        sets, less = list_sets(o, k, 1)
        want = brute_sets(s.pos, s.box, s.group_indices(g1) - 1, s.group_indices(g2) - 1, rc)
        assert sets == want
        # lessnnum: entries before the first j with a larger atom index
        gi, gj = s.group_indices(g1), s.group_indices(g2)
        for r, row in enumerate(sets):
            atoms = [gj[e - 1] for e in row]
            first = next((p for p, a in enumerate(atoms) if a > gi[r]), len(atoms))
            assert less[r] == first
        if it.name in ("lj", "morsec"):
            # converse list = the same relation seen from group 2 (md_neighbours.f90:127)
            sets2, _ = list_sets(o, k, 2)
            assert sets2 == brute_sets(s.pos, s.box, gj - 1, gi - 1, rc)
        if it.name == "morsec":
            nn, _ = list_sets(o, k, 3)
            rnn = it.lists[2][3]
            assert all(len(x) == 3 for x in nn)
            assert nn == brute_sets(s.pos, s.box, gi - 1, gi - 1, rnn)


# --------------------------------------------------------------- closed forms

def test_lj_dimer_energy():
    eps, sig = 0.0104, 3.40
    for r in (3.3, 3.8, 5.0, 6.9, 7.2):
        pos = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
        s = System(box=np.array([30.0, 30.0, 30.0]), pos=pos, vel=np.zeros((2, 3)), mass=np.full(2, 40.0),
                   species=["A", "B"], groups=[["A"], ["B"], ["A", "B"], ["#"]],
                   interactions=[Interaction("lj", [eps, sig, 6.5, 7.5], [(1, 2, 4, 8.5, 10), (2, 1, 4, 8.5, 10)]),
                                 Interaction("lj1g", [eps, sig, 6.5, 7.5], [(3, 3, 4, 8.5, 10)])],
                   integrators=[Integrator("nve", 1.0, 1, 10**9, 10**9)], all_moving=3, xyz_moving=3, z_moving=4,
                   all_atoms=3)
        o = step0(s)
        fc = 1.0 if r < 6.5 else 0.5 * (1 + np.cos(R.PI_LIT * (r - 6.5)))
        V = (sig / r) ** 6
        e = o.energies()["e_inter"]
        assert e[0] == pytest.approx(4 * eps * V * (V - 1) * fc, rel=1e-13)
        x = min(max((r - 6.5) / 1.0, 0), 1)
        fp = 1 - 10 * x ** 3 + 15 * x ** 4 - 6 * x ** 5
        U = r ** -6
        assert e[1] == pytest.approx(U * (4 * eps * sig ** 12 * U - 4 * eps * sig ** 6) * fp, rel=1e-13)


def test_rjl_perfect_fcc_closed_form():
    s = configs.fcc_cu_rgl(nx=5, sigma=0.0, temperature=0.0)
    o = step0(s)
    A0, xi, p, q, r0, R1, R2 = configs.RJL_CU
    a = 3.615
    rep = att = 0.0
    for k, cnt in ((1, 12), (2, 6), (3, 24), (4, 12), (5, 24), (6, 8), (7, 48)):
        r = a / np.sqrt(2.0) * np.sqrt(k)
        if r > R2:
            continue
        fc = 1.0 if r < R1 else 0.5 * (1 + np.cos(R.PI_LIT * (r - R1) / (R2 - R1)))
        rep += cnt * np.exp(-p * (r / r0 - 1)) * fc
        att += cnt * np.exp(-2 * q * (r / r0 - 1)) * fc
    e_atom = A0 * rep - xi * np.sqrt(att)
    assert o.energies()["pe"] / s.n == pytest.approx(e_atom, rel=1e-12)
    assert np.abs(o.state()[2]).max() < 1e-11  # perfect lattice: zero force


def test_tb_perfect_graphene_closed_form():
    s = graphene_sheet()
    o = step0(s)
    D, S, b, r0, delt, a0, c0, d0, R1, R2 = configs.TB_BRENNER_I
    G = 1 + c0 ** 2 / d0 ** 2 - c0 ** 2 / (d0 ** 2 + (1 - 0.5) ** 2)
    B = (1 + a0 * 2 * G) ** (-delt)
    r = 1.42
    e_bond = D / (S - 1) * (np.exp(-np.sqrt(2 * S) * b * (r - r0)) - B * S * np.exp(-np.sqrt(2 / S) * b * (r - r0)))
    assert o.energies()["pe"] / s.n == pytest.approx(1.5 * e_bond, rel=1e-11)
    assert np.abs(o.state()[2]).max() < 1e-10


# ------------------------------------------------------- forces = -grad E

def _cmp_forces(s, efun, rtol=1e-9):
    o = step0(s)
    f_or = o.state()[2]
    e_ref, f_ref = R.forces(efun, s.pos)
    assert rel_inf(f_or, f_ref) < rtol
    return o, e_ref


def test_forces_lj_two_groups():
    s = configs.binary_lj_gas(n=200, box=24.0, seed=31)
    s.interactions = [s.interactions[2]]
    box = torch.tensor(s.box)
    gA, gB = s.group_indices(1) - 1, s.group_indices(2) - 1
    o, e = _cmp_forces(s, lambda x: R.e_lj(x, box, gA, gB, *s.interactions[0].params, s.pos))
    assert o.energies()["e_inter"][0] == pytest.approx(e, rel=1e-11)


def test_forces_lj1g_unit_switch_width():
    s = configs.binary_lj_gas(n=200, box=24.0, seed=32)
    s.interactions = [s.interactions[0]]  # R1=6.5, R2=7.5: R2-R1 = 1
    box = torch.tensor(s.box)
    g = s.group_indices(1) - 1
    o, e = _cmp_forces(s, lambda x: R.e_lj1g(x, box, g, *s.interactions[0].params, s.pos))
    assert o.energies()["e_inter"][0] == pytest.approx(e, rel=1e-11)


def test_forces_lj1g_cutoff_quirk():
    """f_dfr_cut omits 1/(R2-R1) (cut_off_poly.f90:41): the force is
    -grad[phi f]|_f fixed - (R2-R1) grad[phi f]|_phi fixed, not -grad E."""
    R1, R2 = 5.5, 7.5
    s = configs.binary_lj_gas(n=200, box=24.0, seed=33, r1=R1, r2=R2)
    s.interactions = [s.interactions[0]]
    box = torch.tensor(s.box)
    g = s.group_indices(1) - 1
    eps, sig = s.interactions[0].params[:2]
    o = step0(s)
    f_or = o.state()[2]

    def split(x, which):
        i, j = R.pairs_within(s.pos, s.box, g, g, R2)
        keep = i < j
        d = R.disp(x, box, i[keep], j[keep])
        r = torch.sqrt((d * d).sum(-1))
        U = r ** -6
        phi = U * (4 * eps * sig ** 12 * U - 4 * eps * sig ** 6)
        f = R.fc_poly(r, R1, R2)
        return (phi * f.detach()).sum() if which == "phi" else (phi.detach() * f).sum()

    phi_part = lambda x: split(x, "phi")  # noqa: E731
    cut_part = lambda x: split(x, "cut")  # noqa: E731
    _, fa = R.forces(phi_part, s.pos)
    _, fb = R.forces(cut_part, s.pos)
    assert rel_inf(f_or, fa + (R2 - R1) * fb) < 1e-9
    assert rel_inf(f_or, fa + fb) > 1e-3  # and it really is not -grad E


def test_forces_rjl():
    s = configs.fcc_cu_rgl(nx=4, seed=34, sigma=0.08)
    box = torch.tensor(s.box)
    g = s.group_indices(1) - 1
    o, e = _cmp_forces(s, lambda x: R.e_rjl(x, box, g, *configs.RJL_CU, s.pos))
    assert o.energies()["pe"] == pytest.approx(e, rel=1e-12)


def test_forces_tb():
    s = graphene_sheet(jitter=0.12, seed=35)
    box = torch.tensor(s.box)
    g = s.group_indices(1) - 1
    o, e = _cmp_forces(s, lambda x: R.e_tb(x, box, g, *configs.TB_BRENNER_I, s.pos))
    assert o.energies()["pe"] == pytest.approx(e, rel=1e-12)


def test_forces_tb_switching_zone():
    # bonds stretched into R1 <= r < R2 exercise the cutoff derivative terms
    s = graphene_sheet(acc=1.72, jitter=0.1, seed=36)
    s.interactions[0].lists[0] = (1, 1, 8, 2.3, 10)
    box = torch.tensor(s.box)
    g = s.group_indices(1) - 1
    _cmp_forces(s, lambda x: R.e_tb(x, box, g, *configs.TB_BRENNER_I, s.pos))


@pytest.mark.parametrize("kind,simplified", [("morsec", False), ("morsec", True), ("ljc", False), ("ljc", True)])
def test_forces_graphene_metal(kind, simplified):
    s = configs.graphene_on_cu(cx=8, cy=5, gx=8, gy=5, layers=3, seed=37, jitter=0.15)
    s.pos[s.mass < 13] += 0.05 * (np.random.default_rng(2).uniform(-1, 1, ((s.mass < 13).sum(), 3)))
    s.pos = np.mod(s.pos, s.box)
    mor = s.interactions[2]
    if kind == "ljc":
        params = [0.005, 3.2, 2.0, 6.0, 7.0, 1.0 if simplified else 0.0]
    else:
        params = list(configs.MORSEC_CCU[:6]) + [1.0 if simplified else 0.0]
    s.interactions = [s.interactions[1], Interaction(kind, params, mor.lists)]
    box = torch.tensor(s.box)
    gc, gm = s.group_indices(1) - 1, s.group_indices(2) - 1
    rnn = mor.lists[2][3]

    def efun(x):
        return (R.e_tb(x, box, gc, *configs.TB_BRENNER_I, s.pos) +
                R.e_cos(x, box, gc, gm, params, kind == "morsec", rnn, s.pos, simplified))

    o, e = _cmp_forces(s, efun)
    assert o.energies()["pe"] == pytest.approx(e, rel=1e-12)
    if not simplified:
        n_ref = R.graphene_normals(torch.tensor(s.pos), box, gc, rnn, s.pos).numpy()
        np.testing.assert_allclose(o.normals(2, len(gc)), n_ref, atol=1e-13)


# ----------------------------------------------------------------- dynamics

def test_nve_time_reversible_and_conserving():
    # velocity Verlet: the energy error is O(dt^2) (a wrong half-kick breaks this)
    errs = []
    for dt in (0.5, 0.25):
        s = configs.fcc_cu_rgl(nx=4, seed=41, temperature=600.0)
        o = oracle_ctx(s)
        o.run("nve", dt, 0, 0, True)
        e0 = o.energies()
        n = int(100 / dt)
        mx = 0.0
        for k in range(10):
            o.run("nve", dt, 1 + k * n // 10, (k + 1) * n // 10, True)
            e = o.energies()
            mx = max(mx, abs(e["pe"] + e["ke"] - e0["pe"] - e0["ke"]))
        errs.append(mx / s.n)
    assert errs[1] < 1e-5 and 3.5 < errs[0] / errs[1] < 4.5
    s = configs.fcc_cu_rgl(nx=4, seed=41, temperature=600.0)
    o = oracle_ctx(s)
    o.run("nve", 1.0, 0, 0, True)
    p0, v0, _ = o.state()
    o.run("nve", 1.0, 1, 200, True)
    p1, v1, _ = o.state()
    o.set_state(p1, -v1)
    o.run("nve", 1.0, 0, 0, True)  # refresh lists and forces at the reversed state
    o.run("nve", 1.0, 1, 200, True)
    p2, v2, _ = o.state()
    dp = p2 - p0
    dp -= s.box * np.round(dp / s.box)
    assert np.abs(dp).max() < 1e-9
    assert np.abs(v2 + v0).max() < 1e-10


@pytest.mark.parametrize("M", [1, 3])
def test_nhc_conserved_quantity(M):
    """E + E_NHC is conserved to O(dt^2) while the thermostat moves ~0.1
    eV/atom of energy (md_integrators.f90:200-260, md_simulation.f90:201)."""
    res = []
    for dt in (1.0, 0.5):
        s = configs.fcc_cu_rgl(nx=4, seed=42, temperature=300.0)
        s.nhcs = [NHC(1, 600.0, M, 3 * s.n * KB * 600.0 * 20.0 ** 2)]
        s.interactions[0].lists[0] = (1, 1, 160, 6.5, 5)
        o = oracle_ctx(s)
        o.run("nvt", dt, 0, 0, True)
        n = int(400 / dt)
        cons, phys = [], []
        for k in range(10):
            o.run("nvt", dt, 1 + n // 10 * k, n // 10 * (k + 1), True)
            e = o.energies()
            cons.append(e["pe"] + e["ke"] + e["e_nhc"].sum())
            phys.append(e["pe"] + e["ke"])
        res.append((np.ptp(cons) / s.n, np.ptp(phys) / s.n, e["temp"]))
    assert 3.5 < res[0][0] / res[1][0] < 4.5
    assert res[1][0] < 2e-3 * res[1][1]
    assert res[1][2] > 450.0  # driven from 300 K towards 600 K


def test_nvms_lj_dimer_minimum():
    eps, sig = 0.0104, 3.40
    pos = np.array([[5.0, 5.0, 5.0], [9.3, 5.2, 5.1]])
    s = System(box=np.array([30.0, 30.0, 30.0]), pos=pos, vel=np.zeros((2, 3)), mass=np.full(2, 40.0),
               species=["A", "A"], groups=[["A"], ["#"]],
               interactions=[Interaction("lj1g", [eps, sig, 6.5, 7.5], [(1, 1, 4, 8.5, 10)])],
               integrators=[Integrator("nvms", 2.0, 100000, 10**9, 10**9)], z_moving=2, ms_de=1e-14)
    o = oracle_ctx(s)
    o.run("nvms", 2.0, 0, 0, True, 1e-14)
    last, ev = o.run("nvms", 2.0, 1, 100000, True, 1e-14)
    assert ev == 1 and last < 100000
    p, _, _ = o.state()
    r = np.linalg.norm(p[1] - p[0])
    assert r == pytest.approx(2 ** (1 / 6) * sig, rel=1e-5)


def test_nvms_velocity_rule():
    """One nvms step against its definition (md_integrators.f90:99-145):
    after the velocity-Verlet kicks, an xyz atom keeps only the projection
    (v.F/F.F) F of its velocity when v.F > 0 and is stopped otherwise; a
    z-only atom keeps its velocity when v.F > 0 and is stopped otherwise. The
    forces at the new positions come from a fresh oracle context."""
    eps, sig, dt = 0.0104, 3.40, 2.0
    pos = np.array([[5.0, 5.0, 5.0], [8.7, 5.4, 5.3], [5.3, 8.9, 4.6], [8.5, 8.6, 8.9]])
    box = np.array([30.0, 30.0, 30.0])
    mass = np.array([40.0, 39.0, 41.0, 38.0])

    def dimer_system(p):
        return System(box=box, pos=p.copy(), vel=np.zeros((4, 3)), mass=mass, species=["A", "A", "B", "B"],
                      groups=[["A"], ["B"], ["A", "B"]],
                      interactions=[Interaction("lj1g", [eps, sig, 6.5, 7.5], [(3, 3, 4, 8.5, 10)])],
                      integrators=[Integrator("nvms", dt, 10, 10**9, 10**9)], all_moving=3, xyz_moving=1,
                      z_moving=2, all_atoms=3)

    s = dimer_system(pos)
    o = step0(s)
    f0 = o.state()[2]
    fh = f0 / np.linalg.norm(f0, axis=1)[:, None]
    # atoms 0 and 2 move along their force, atoms 1 and 3 against it
    v0 = 0.02 * np.array([fh[0], -fh[1] + [0.0, 0.3, 0.0], fh[2], -fh[3]])
    o.set_state(pos, v0)
    o.run("nvms", dt, 1, 1, True, -1.0)
    p1, v1, f1 = o.state()

    a = dt / 2 / MASS_COEF / mass[:, None]
    vh = v0 + f0 * a
    xyz = np.array([True, True, False, False])
    want_p = pos.copy()
    want_p[xyz] += vh[xyz] * dt
    want_p[~xyz, 2] += vh[~xyz, 2] * dt
    np.testing.assert_allclose(p1, want_p, rtol=0, atol=1e-13)
    f1_ref = step0(dimer_system(p1)).state()[2]
    np.testing.assert_allclose(f1, f1_ref, rtol=1e-13, atol=1e-16)
    vk = vh + f1 * a
    vk[~xyz, :2] = v0[~xyz, :2]
    want_v = np.zeros((4, 3))
    for i in range(4):
        fv = vk[i] @ f1[i]
        if xyz[i] and fv > 0:
            want_v[i] = fv / (f1[i] @ f1[i]) * f1[i]
        elif not xyz[i] and fv > 0:
            want_v[i] = vk[i]
    assert (np.abs(want_v[[0, 2]]).max(1) > 0).all() and (want_v[[1, 3]] == 0).all()
    np.testing.assert_allclose(v1, want_v, rtol=1e-12, atol=1e-18)


# ------------------------------------------------------------------ rebosc

def test_rebo_perfect_graphene_closed_form():
    """Planar sheet: every dihedral term vanishes (sin^2 = 0) and each bond
    sees two others at cos = -1/2, so E/atom = 1.5 [(1+Q/r)A e^(-alpha r) -
    (1 + 2 g(-1/2))^(-1/2) sum_k B_k e^(-beta_k r)] at r = 1.42 < R1; the
    central-difference forces vanish to their rounding."""
    s = configs.graphene_rebo(gx=6, gy=4)
    o = step0(s)
    P = configs.REBO_C
    r = 1.42
    g = sum(P[10 + k] * (-0.5) ** k for k in range(6))
    eb = (1 + P[1] / r) * P[0] * np.exp(-P[2] * r) - (1 + 2 * g) ** -0.5 * sum(
        P[3 + k] * np.exp(-P[6 + k] * r) for k in range(3))
    assert o.energies()["pe"] / s.n == pytest.approx(1.5 * eb, rel=1e-12)
    assert np.abs(o.state()[2]).max() < 1e-6


def test_forces_rebo_buckled():
    """Jittered and buckled sheet (dihedral and cutoff terms active): the
    energy equals an independent restatement to 1e-12 and the reference's
    central differences (dx = 1e-6 A over truncated lists) equal -grad E by
    autograd to 1e-7."""
    s = configs.graphene_rebo(gx=5, gy=4, jitter=0.1, buckle=0.2, seed=72)
    box = torch.tensor(s.box)
    g = s.group_indices(1) - 1
    o = step0(s)
    e, f = R.forces(lambda x: R.e_rebo(x, box, g, configs.REBO_C, s.pos), s.pos)
    assert o.energies()["pe"] == pytest.approx(e, rel=1e-12)
    assert rel_inf(o.state()[2], f) < 1e-7
