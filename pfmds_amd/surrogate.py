"""Learned surrogate for the graphene terms of a graphene-on-metal system.

The interaction 'sur' (include/pfm.h, PFM_SUR) stands in for the carbon-carbon
Tersoff-Brenner term and the carbon-metal morsec term: every carbon atom gets
an energy from a 40-64-64-1 tanh network of 40 descriptors of its carbon and
metal neighbourhood within r_c = 6 A.  The metal's own rjl term stays.  This
module is the host side of it:

  - `split_system`: from a plain system (rjl + tb + morsec), the label system
    (tb + morsec only: the reference energies and forces the network learns)
    and the surrogate system (rjl + sur);
  - `generate_frames`: snapshots of plain nvt runs at several temperatures,
    labelled by the FP64 path;
  - `train`: Adam over frames, the loss and its gradient computed on the GPU by
    pfm_sur_loss_grad (forward, force and training kernels on MFMA);
  - `JudgedRun`: surrogate MD in which the plain potential is evaluated every
    `judge_every` steps at the same positions; when the force RMSE against it
    passes `threshold` (eV/A), the run continues on the plain potential.

No counterpart in the reference; the plain path it replaces is TB_forces and
MorseC_forces (TersoffBrenner.f90:74-150, MorseCosine.f90:53-139).
"""
import copy
import dataclasses
import time
from typing import List

import numpy as np

from . import capi, system

SUR_RC = 6.0
NTHETA = capi.SUR_NTHETA


def split_system(s, g_all=None, skin=1.0, nmax=(160, 112), period=10):
    """(label system, surrogate system, index of 'sur' among the surrogate
    system's interactions, 1-based) of a plain graphene/metal system whose
    interactions include tb (carbon group) and morsec."""
    names = [it.name for it in s.interactions]
    if "tb" not in names or "morsec" not in names:
        raise ValueError("the surrogate replaces a tb + morsec pair of interactions")
    g_c = s.interactions[names.index("tb")].lists[0][0]
    g_all = g_all if g_all is not None else s.all_atoms
    lab = copy.deepcopy(s)
    lab.interactions = [it for it in s.interactions if it.name in ("tb", "morsec")]
    sur = copy.deepcopy(s)
    rc = SUR_RC + skin
    si = system.Interaction("sur", [SUR_RC], [(g_c, g_all, nmax[0], rc, period), (g_all, g_c, nmax[1], rc, period)])
    inters = []
    for it in s.interactions:
        if it.name == "tb":
            inters.append(si)
        elif it.name != "morsec":
            inters.append(it)
    sur.interactions = inters
    return lab, sur, inters.index(si) + 1


@dataclasses.dataclass
class Frames:
    pos: np.ndarray      # [F, N, 3] (float32)
    force: np.ndarray    # [F, N, 3] tb + morsec forces (float32)
    energy: np.ndarray   # [F] tb + morsec energy (eV)
    temperature: np.ndarray
    box: np.ndarray

    def __len__(self):
        return len(self.energy)

    def subset(self, idx):
        idx = np.asarray(idx)
        return Frames(self.pos[idx], self.force[idx], self.energy[idx], self.temperature[idx], self.box)

    def positions(self, k):
        """Frame k in float64, wrapped into [0, L) (float32 storage can round
        a coordinate next to L past it)."""
        return np.mod(self.pos[k].astype(np.float64), self.box)


def label(lab_ctx, pos):
    """tb + morsec energy and forces at `pos` (FP64 path)."""
    lab_ctx.set_state(pos, np.zeros_like(pos))
    lab_ctx.run("nve", 0.5, 0, 0, True)
    e = float(np.sum(lab_ctx.energies()["e_inter"]))
    return e, lab_ctx.state()[2]


def generate_frames(s, temperatures, frames_per_temperature, every=50, warmup=200, dt=0.5,
                    precision=capi.MIXED, log=None):
    """Snapshots of plain nvt runs of `s` (one run per temperature), labelled
    with the FP64 tb + morsec path."""
    lab_s, _, _ = split_system(s)
    lab = system.build_context(lab_s, lambda b, p, v, m: capi.Context.gpu(b, p, v, m, capi.FP64))
    P, F, E, T = [], [], [], []
    for temp in temperatures:
        st = copy.deepcopy(s)
        for c in st.nhcs:
            c.temperature = float(temp)
        ctx = system.build_context(st, lambda b, p, v, m: capi.Context.gpu(b, p, v, m, precision))
        ctx.run("nvt", dt, 0, warmup, False)
        step = warmup
        for _ in range(frames_per_temperature):
            ctx.run("nvt", dt, step + 1, step + every, False)
            step += every
            pos = ctx.state()[0]
            e, f = label(lab, pos)
            P.append(pos.astype(np.float32))
            F.append(f.astype(np.float32))
            E.append(e)
            T.append(temp)
        ctx.close()
        if log:
            log(f"frames at {temp} K: {frames_per_temperature}")
    lab.close()
    return Frames(np.array(P), np.array(F), np.array(E), np.array(T, dtype=float), np.asarray(s.box, dtype=float))


def surrogate_context(sur_s, theta, precision=capi.MIXED):
    ctx = system.build_context(sur_s, lambda b, p, v, m: capi.Context.gpu(b, p, v, m, precision))
    ctx.sur_set_params(sur_index(sur_s), theta)
    return ctx


def sur_index(sur_s):
    return [it.name for it in sur_s.interactions].index("sur") + 1


def training_system(sur_s):
    """The surrogate system with the 'sur' interaction alone (its forces are
    then the handle's forces)."""
    t = copy.deepcopy(sur_s)
    t.interactions = [it for it in sur_s.interactions if it.name == "sur"]
    t.nhcs = []
    return t


def feature_scaling(ctx, inter, frames, n_rows, max_frames=8):
    """Mean and inverse standard deviation of the descriptors over frames
    (sets placeholder parameters on the handle first)."""
    ctx.sur_set_params(inter, init_theta(np.zeros(40), np.ones(40), 0.0))
    G = []
    for k in np.unique(np.linspace(0, len(frames) - 1, max_frames).astype(int)):
        pos = frames.positions(k)
        ctx.set_state(pos, np.zeros_like(pos))
        ctx.run("nve", 0.5, 0, 0, False)
        G.append(ctx.sur_descriptors(inter, n_rows))
    G = np.concatenate(G)
    mu = G.mean(0)
    sd = G.std(0)
    return mu, 1.0 / np.maximum(sd, 1e-3 * np.maximum(np.abs(mu), 1e-6))


def init_theta(mu, isg, b3, seed=0):
    """Fan-in scaled uniform weights, b3 = mean label energy per carbon."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (64, 40)) / np.sqrt(40)
    W2 = rng.uniform(-1, 1, (64, 64)) / np.sqrt(64)
    w3 = rng.uniform(-1, 1, 64) / np.sqrt(64) * 0.1
    return np.concatenate([mu, isg, W1.ravel(), np.zeros(64), W2.ravel(), np.zeros(64), w3, [b3]])


class Adam:
    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, theta, g, lr=None):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return theta - (lr or self.lr) * mh / (np.sqrt(vh) + self.eps)


def train(ctx, inter, frames, theta, iters, lr=2e-3, lr_end=2e-4, w_e=1.0, w_f=1.0, seed=0, log=None,
          log_every=100, batch=1):
    """Adam over randomly drawn frames, `batch` frames (gradients averaged)
    per iteration; the loss and gradient of each frame come from
    pfm_sur_loss_grad.  `ctx`/`frames` may be lists (one handle per system
    size); frames are drawn over all sets in proportion to their counts.
    Returns the final parameters and the per-iteration losses."""
    ctxs = ctx if isinstance(ctx, (list, tuple)) else [ctx]
    sets = frames if isinstance(frames, (list, tuple)) else [frames]
    sizes = np.array([len(f) for f in sets], dtype=float)
    rng = np.random.default_rng(seed)
    opt = Adam(len(theta), lr)
    hist = []
    theta = np.array(theta, dtype=np.float64)
    for c in ctxs:
        c.sur_set_params(inter, theta)
    for it in range(iters):
        loss, g = 0.0, 0.0
        for _ in range(batch):
            w = int(rng.choice(len(sets), p=sizes / sizes.sum()))
            c, fr = ctxs[w], sets[w]
            k = int(rng.integers(len(fr)))
            pos = fr.positions(k)
            c.set_state(pos, np.zeros_like(pos))
            c.run("nve", 0.5, 0, 0, False)
            lk, gk = c.sur_loss_grad(inter, fr.force[k].astype(np.float64), fr.energy[k], w_e, w_f)
            loss = loss + lk / batch
            g = g + gk / batch
        hist.append(loss)
        a = lr * (lr_end / lr) ** (it / max(iters - 1, 1))
        theta = opt.step(theta, g, a)
        for cc in ctxs:
            cc.sur_set_params(inter, theta)
        if log and (it + 1) % log_every == 0:
            h = np.array(hist[-log_every:])
            log(f"iter {it + 1}: loss {h[:, 0].mean():.4e} (energy {h[:, 1].mean():.3e}, force {h[:, 2].mean():.3e})")
    return theta, np.array(hist)


def force_rmse(ctx, frames, atoms=None):
    """RMSE (eV/A) of the handle's forces against the frames' labels, over
    `atoms` (default: all)."""
    err = []
    for k in range(len(frames)):
        pos = frames.positions(k)
        ctx.set_state(pos, np.zeros_like(pos))
        ctx.run("nve", 0.5, 0, 0, False)
        d = ctx.state()[2] - frames.force[k]
        err.append(d if atoms is None else d[atoms])
    e = np.concatenate(err)
    return float(np.sqrt(np.mean(e * e)))


class JudgedRun:
    """Surrogate MD with the plain potential as judge.  `sur_ctx` steps the
    system; every `judge_every` steps `plain_ctx` evaluates the plain forces
    at the same positions, and the force RMSE over the moving atoms is
    recorded; past `threshold` eV/A the run moves to `plain_ctx` (positions,
    velocities and the Nose-Hoover chain states).  With `group` (the group
    number of the moving atoms) the state copy and the RMSE stay on the GPU
    (pfm_copy_state, pfm_force_rmse); without it they go through the host."""

    def __init__(self, sur_ctx, plain_ctx, moving, judge_every=100, threshold=0.05, group=None):
        self.sur, self.plain = sur_ctx, plain_ctx
        self.group = group
        self.moving = np.asarray(moving)
        self.judge_every, self.threshold = judge_every, threshold
        self.on_plain = False
        self.judged: List[tuple] = []
        self.switch_step = None
        self.t_judge = 0.0

    def judge(self, step):
        t0 = time.perf_counter()
        if self.group is not None:
            self.plain.copy_state_from(self.sur)
            self.plain.run("nvt", 0.5, 0, 0, False)
            rmse = self.plain.force_rmse(self.sur, self.group)
        else:
            p, v, f = self.sur.state()
            self.plain.set_state(p, v)
            self.plain.run("nvt", 0.5, 0, 0, False)
            fp = self.plain.state()[2]
            d = (f - fp)[self.moving]
            rmse = float(np.sqrt(np.mean(d * d)))
        self.judged.append((step, rmse))
        self.t_judge += time.perf_counter() - t0
        if rmse > self.threshold:
            self.on_plain = True
            self.switch_step = step
            for c, M in enumerate(self.sur.nhc_M, 1):
                self.plain.set_chain(c, *self.sur.get_chain(c, M))
        return rmse

    def run(self, integrator, dt, first, last):
        """Steps first..last; step 0 (when first == 0) builds lists and forces."""
        s = first
        if s == 0:
            self.sur.run(integrator, dt, 0, 0, False)
            s = 1
        while s <= last:
            # chunks end on the judge steps (multiples of judge_every)
            e = min(last, -(-s // self.judge_every) * self.judge_every)
            ctx = self.plain if self.on_plain else self.sur
            ctx.run(integrator, dt, s, e, False)
            if not self.on_plain and e % self.judge_every == 0:
                self.judge(e)
            s = e + 1
        return self.plain if self.on_plain else self.sur
