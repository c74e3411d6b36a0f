"""CPU definition of the learned surrogate (interaction 'sur'), in float64
torch, for the tests: descriptors by their textbook sums (the angular ones
over explicit neighbour pairs j != k), the 40-64-64-1 tanh perceptron, forces
by autograd and the training loss with its gradient by double backward.

The reference has no surrogate; these definitions state the model of
include/pfm.h (PFM_SUR) independently of the GPU kernels' factorisations
(moment tensors, the hand-written forward-over-reverse gradient), which
`moments_g5` and `manual_grad` restate in numpy so the CPU tests can pin
those factorisations against the textbook forms.
"""
import numpy as np
import torch

NF, NH = 40, 64
NTHETA = 6929
ETA2 = 0.01 * 500.0 ** (np.arange(12) / 11.0)
ETAA = (0.05, 0.5)
PI = np.pi


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64) if not torch.is_tensor(theta) else theta
    o = 0
    out = {}
    for name, shape in (("mu", (NF,)), ("isg", (NF,)), ("W1", (NH, NF)), ("b1", (NH,)), ("W2", (NH, NH)),
                        ("b2", (NH,)), ("w3", (NH,)), ("b3", ())):
        n = int(np.prod(shape)) if shape else 1
        out[name] = t[o:o + n].reshape(shape) if shape else t[o]
        o += n
    assert o == NTHETA
    return out


def init_theta(mu, isg, seed=0, scale=1.0, b3=0.0):
    """Random weights (uniform, fan-in scaled) with the given feature scaling."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (NH, NF)) * scale / np.sqrt(NF)
    W2 = rng.uniform(-1, 1, (NH, NH)) * scale / np.sqrt(NH)
    w3 = rng.uniform(-1, 1, NH) * scale / np.sqrt(NH)
    b1 = rng.uniform(-0.1, 0.1, NH)
    b2 = rng.uniform(-0.1, 0.1, NH)
    return np.concatenate([mu, isg, W1.ravel(), b1, W2.ravel(), b2, w3, [b3]])


def bf16_round(x):
    """Round float64 values to the nearest BF16 (ties to even) through float32."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def round_weights(theta):
    """The parameters as the device holds them: W1, W2 in BF16, the rest in FP32."""
    t = np.array(theta, dtype=np.float64)
    p = unpack(t)
    out = [np.asarray(p["mu"], np.float32), np.asarray(p["isg"], np.float32), bf16_round(p["W1"]).ravel(),
           np.asarray(p["b1"], np.float32), bf16_round(p["W2"]).ravel(), np.asarray(p["b2"], np.float32),
           np.asarray(p["w3"], np.float32), [np.float32(p["b3"])]]
    return np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in out])


# ------------------------------------------------------------------ geometry

def pairs(pos, box, g1, g2, rc):
    """(row index into g1, atom i, atom j) for j in g2, j != i, r_ij < rc; the
    minimum image fixed at the given positions."""
    pos = np.asarray(pos)
    g1, g2 = np.asarray(g1), np.asarray(g2)
    d = pos[g2][None, :, :] - pos[g1][:, None, :]
    d -= box * np.round(d / box)
    r2 = (d * d).sum(-1)
    ok = (r2 < rc * rc) & (g1[:, None] != g2[None, :])
    ri, cj = np.nonzero(ok)
    shift = np.round((pos[g2][None, :, :] - pos[g1][:, None, :]) / box)[ri, cj]
    return ri, g1[ri], g2[cj], shift


def fc(r, rc):
    return 0.5 * (1.0 + torch.cos(PI * r / rc))


def descriptors(x, box, g1, g2, rc, P=None, brute=True):
    """G[N_g1, 40] of the positions x (torch float64, may require grad)."""
    g1 = np.asarray(g1)
    if P is None:
        P = pairs(x.detach().numpy(), box, g1, g2, rc)
    ri, ia, ja, sh = P
    boxt = torch.tensor(np.asarray(box, dtype=np.float64))
    d = x[ja] - x[ia] - torch.tensor(sh) * boxt
    r2 = (d * d).sum(-1)
    r = torch.sqrt(r2)
    f = fc(r, rc)
    in1 = np.zeros(x.shape[0], dtype=bool)
    in1[g1] = True
    sj = torch.tensor(np.where(in1[ja], 0, 1))
    n1 = len(g1)
    rit = torch.tensor(ri)
    cols = []
    for s in (0, 1):
        for k in range(12):
            v = torch.exp(-ETA2[k] * r2) * f * (sj == s)
            cols.append(torch.zeros(n1, dtype=x.dtype).index_add(0, rit, v))
    G2 = torch.stack(cols, 1)
    G5 = torch.zeros(n1, 16, dtype=x.dtype)
    u = d / r[:, None]
    for e, eta in enumerate(ETAA):
        w = torch.exp(-eta * r2) * f
        if brute:
            g5 = _g5_brute(ri, u, w, sj, n1)
        else:
            g5 = _g5_moments(ri, u, w, sj, n1)
        for p in range(2):
            for m in range(4):
                G5[:, p * 8 + e * 4 + m] = g5[p][m]
    return torch.cat([G2, G5], 1)


def _g5_from_cos(c, wjk):
    return [(wjk * (1 + c)).sum(), (wjk * (1 - c)).sum(), 0.125 * (wjk * (1 + c) ** 4).sum(),
            0.125 * (wjk * (1 - c) ** 4).sum()]


def _g5_brute(ri, u, w, sj, n1):
    """Explicit sums over ordered pairs (j, k), j != k, of the neighbours of each row."""
    out = [[[] for _ in range(4)] for _ in range(2)]
    order = np.argsort(ri, kind="stable")
    starts = np.searchsorted(ri[order], np.arange(n1 + 1))
    for a in range(n1):
        idx = torch.tensor(order[starts[a]:starts[a + 1]])
        ua, wa, sa = u[idx], w[idx], sj[idx]
        c = ua @ ua.T
        ww = wa[:, None] * wa[None, :]
        cc = (sa[:, None] == 0) & (sa[None, :] == 0)
        cc = cc & ~torch.eye(len(idx), dtype=torch.bool)
        cx = (sa[:, None] == 0) & (sa[None, :] == 1)
        for p, m in ((0, cc), (1, cx)):
            vals = _g5_from_cos(c[m], ww[m])
            for z in range(4):
                out[p][z].append(vals[z])
    return [[torch.stack(out[p][z]) for z in range(4)] for p in range(2)]


MONOS = [(a, b, n - a - b) for n in range(5) for a in range(n, -1, -1) for b in range(n - a, -1, -1)]


def _g5_moments(ri, u, w, sj, n1):
    """The O(neighbours) form of the GPU kernels: degree sums
    S_n = <M^{s1}_n, M^{s2}_n> of the moment tensors M^s_n = sum_j w_j u_j^(x)n."""
    from math import factorial
    mono = torch.stack([u[:, 0] ** a * u[:, 1] ** b * u[:, 2] ** c for a, b, c in MONOS], 1)
    mul = torch.tensor([factorial(a + b + c) / (factorial(a) * factorial(b) * factorial(c)) for a, b, c in MONOS])
    deg = np.array([a + b + c for a, b, c in MONOS])
    rit = torch.tensor(ri)
    M = [torch.zeros(n1, 35, dtype=u.dtype).index_add(0, rit, (w * (sj == s))[:, None] * mono) for s in (0, 1)]
    D = torch.zeros(n1, dtype=u.dtype).index_add(0, rit, w * w * (sj == 0))
    out = []
    for p in range(2):
        prod = M[0] * (M[0] if p == 0 else M[1]) * mul
        S = [prod[:, deg == n].sum(1) - (D if p == 0 else 0.0) for n in range(5)]
        out.append([S[0] + S[1], S[0] - S[1], 0.125 * (S[0] + 4 * S[1] + 6 * S[2] + 4 * S[3] + S[4]),
                    0.125 * (S[0] - 4 * S[1] + 6 * S[2] - 4 * S[3] + S[4])])
    return out


# ------------------------------------------------------------------ network

def network(G, theta):
    """Per-atom energies E_i (torch float64, differentiable in G and theta)."""
    p = unpack(theta)
    x = (G - p["mu"]) * p["isg"]
    a1 = torch.tanh(x @ p["W1"].T + p["b1"])
    a2 = torch.tanh(a1 @ p["W2"].T + p["b2"])
    return a2 @ p["w3"] + p["b3"]


def energy_forces(pos, box, g1, g2, rc, theta, P=None):
    """Total energy and forces [N,3] of the surrogate (float64 throughout)."""
    x = torch.tensor(np.asarray(pos, dtype=np.float64), requires_grad=True)
    th = torch.tensor(np.asarray(theta, dtype=np.float64))
    E = network(descriptors(x, box, g1, g2, rc, P, brute=False), th).sum()
    (g,) = torch.autograd.grad(E, x)
    return float(E.detach()), -g.numpy()


def loss_and_grad(pos, box, g1, g2, rc, theta, f_ref, e_ref, w_e, w_f, P=None):
    """The training loss of pfm_sur_loss_grad and its gradient by double
    backward (mu and isg held fixed)."""
    x = torch.tensor(np.asarray(pos, dtype=np.float64), requires_grad=True)
    th = torch.tensor(np.asarray(theta, dtype=np.float64), requires_grad=True)
    G = descriptors(x, box, g1, g2, rc, P, brute=False)
    E = network(G, th).sum()
    (gx,) = torch.autograd.grad(E, x, create_graph=True)
    F = -gx
    in2 = np.zeros(len(pos), dtype=bool)
    in2[np.asarray(g2)] = True
    m = torch.tensor(in2)
    n1, n2 = len(g1), int(in2.sum())
    r = (F - torch.tensor(f_ref))[m]
    le = w_e * ((E - e_ref) / n1) ** 2
    lf = w_f * (r * r).sum() / (3 * n2)
    L = le + lf
    (gt,) = torch.autograd.grad(L, th)
    g = gt.detach().numpy().copy()
    g[:80] = 0.0
    return np.array([float(L.detach()), float(le.detach()), float(lf.detach())]), g


def manual_grad(G, t, theta, cE, cF):
    """The weight gradient the training kernel forms, restated in numpy:
    cE * sum_i dE_i/dtheta + cF * sum_i d/dtheta [grad_x NN(x_i) . t_i],
    x_i = (G_i - mu) isg, t_i the scaled tangent descriptors."""
    p = {k: np.asarray(v, dtype=np.float64) for k, v in unpack(np.asarray(theta)).items()}
    x = (G - p["mu"]) * p["isg"]
    z1 = x @ p["W1"].T + p["b1"]
    a1 = np.tanh(z1)
    z1d = t @ p["W1"].T
    s1 = 1 - a1 * a1
    a1d = s1 * z1d
    a2 = np.tanh(a1 @ p["W2"].T + p["b2"])
    z2d = a1d @ p["W2"].T
    s2 = 1 - a2 * a2
    a2d = s2 * z2d
    w3 = p["w3"]
    adj_a2 = cE * w3 - 2 * cF * w3 * a2 * z2d
    adz2 = cF * w3 * s2
    az2 = adj_a2 * s2
    ad = adz2 @ p["W2"]
    adj_a1 = az2 @ p["W2"] - 2 * ad * a1 * z1d
    adz1 = ad * s1
    az1 = adj_a1 * s1
    dW1 = az1.T @ x + adz1.T @ t
    dW2 = az2.T @ a1 + adz2.T @ a1d
    g = np.concatenate([np.zeros(80), dW1.ravel(), az1.sum(0), dW2.ravel(), az2.sum(0),
                        (cE * a2 + cF * a2d).sum(0), [cE * len(G)]])
    return g
