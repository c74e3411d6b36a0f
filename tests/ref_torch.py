"""Energies written from the potentials' mathematical definitions, vectorised
over explicit pair/triplet index arrays in torch float64, used to pin the CPU
oracle: its forces must equal -grad E computed by autograd, its energies
must equal these sums.  Nothing here uses the oracle's loop structure
(neighbour lists, stored distances, the reference's force formulas).

Definitions (reference files give the same functions):
  cosine cutoff   f = 1 (r<R1), (1+cos(pi (r-R1)/(R2-R1)))/2 (R1<=r<R2), 0
  poly cutoff     f = 1 - 10x^3 + 15x^4 - 6x^5, x = (r-R1)/(R2-R1) clamped
  lj              4 eps V (V-1) f,  V = (sig/r)^6          per pair of g1 x g2
  lj1g            U (c12 U - c6) f_poly, U = r^-6          per unordered pair
  rjl             sum_i A0 sum_j exp(-p(r/r0-1)) f - xi sqrt(sum_j exp(-2q(r/r0-1)) f)
  tb              sum_{i<j} f D/(S-1) [e^{-sqrt(2S) b (r-r0)} - (B_ij+B_ji)/2 S e^{-sqrt(2/S) b (r-r0)}]
                  B_ij = (1 + a0 sum_{k!=j} f(r_ik) G(cos theta_jik))^-delta,
                  G(c) = 1 + c0^2/d0^2 - c0^2/(d0^2 + (1+c)^2)
  morsec          sum_{i in C, j in M} D (V2^2 - 2 V2 V3) f, V2 = exp(-a(r-r0)), V3 = (|n_i.d|/r)^delta
  ljc             sum 4 eps (V2^2 - V2 V3) f, V2 = (sig/r)^6
  graphene normal n_i = +-(d2-d1)x(d1-d3)/|..| with n_z >= 0, d_k to the 3 nearest C
"""
import numpy as np
import torch

PI_LIT = 3.14159265358979  # the cutoff module's pi literal (part of the function's definition)


def disp(pos, box, i, j):
    d = pos[j] - pos[i]
    return d - box * torch.round(d / box)


def pairs_within(pos_np, box, g1, g2, rc):
    """All (i, j), i in g1, j in g2, i != j, |d| < rc (numpy, all pairs)."""
    g1 = np.asarray(g1)
    g2 = np.asarray(g2)
    d = pos_np[g2][None, :, :] - pos_np[g1][:, None, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d * d).sum(-1))
    a, b = np.nonzero((r < rc) & (g1[:, None] != g2[None, :]))
    return g1[a], g2[b]


def fc_cos(r, R1, R2):
    x = torch.clamp((r - R1) / (R2 - R1), 0.0, 1.0)
    return torch.where(r < R1, torch.ones_like(r),
                       torch.where(r < R2, 0.5 * (1.0 + torch.cos(PI_LIT * x)), torch.zeros_like(r)))


def fc_poly(r, R1, R2):
    x = torch.clamp((r - R1) / (R2 - R1), 0.0, 1.0)
    return 1.0 - 10.0 * x ** 3 + 15.0 * x ** 4 - 6.0 * x ** 5


def e_lj(pos, box, g1, g2, eps, sig, R1, R2, pos_np):
    i, j = pairs_within(pos_np, box.numpy(), g1, g2, R2)
    d = disp(pos, box, i, j)
    r = torch.sqrt((d * d).sum(-1))
    V = (sig / r) ** 6
    return (4 * eps * V * (V - 1.0) * fc_cos(r, R1, R2)).sum()


def e_lj1g(pos, box, g, eps, sig, R1, R2, pos_np, cut_poly=None):
    """cut_poly: optional callable r -> f to replace the cutoff (used to split
    the energy into terms for the lj1g force quirk)."""
    i, j = pairs_within(pos_np, box.numpy(), g, g, R2)
    keep = i < j
    i, j = i[keep], j[keep]
    d = disp(pos, box, i, j)
    r = torch.sqrt((d * d).sum(-1))
    U = r ** -6
    c6, c12 = 4 * eps * sig ** 6, 4 * eps * sig ** 12
    f = fc_poly(r, R1, R2) if cut_poly is None else cut_poly(r)
    return (U * (c12 * U - c6) * f).sum()


def e_rjl(pos, box, g, A0, xi, p, q, r0, R1, R2, pos_np):
    i, j = pairs_within(pos_np, box.numpy(), g, g, R2)
    d = disp(pos, box, i, j)
    r = torch.sqrt((d * d).sum(-1))
    f = fc_cos(r, R1, R2)
    n = pos.shape[0]
    rep = torch.zeros(n, dtype=pos.dtype).index_add(0, torch.as_tensor(i), A0 * torch.exp(-p * (r / r0 - 1.0)) * f)
    rho = torch.zeros(n, dtype=pos.dtype).index_add(0, torch.as_tensor(i), torch.exp(-2 * q * (r / r0 - 1.0)) * f)
    gi = torch.as_tensor(np.asarray(g))
    return (rep[gi] - xi * torch.sqrt(rho[gi])).sum()


def e_tb(pos, box, g, D, S, b, r0, delt, a0, c0, d0, R1, R2, pos_np):
    i, j = pairs_within(pos_np, box.numpy(), g, g, R2)
    n_b = len(i)
    d = disp(pos, box, i, j)
    r = torch.sqrt((d * d).sum(-1))
    f = fc_cos(r, R1, R2)
    # triplets: bond e=(i,j) and bond e2=(i,k), k != j
    by_i = {}
    for e, a in enumerate(i):
        by_i.setdefault(int(a), []).append(e)
    t1, t2 = [], []
    for e in range(n_b):
        for e2 in by_i[int(i[e])]:
            if e2 != e:
                t1.append(e)
                t2.append(e2)
    t1 = torch.as_tensor(t1, dtype=torch.long)
    t2 = torch.as_tensor(t2, dtype=torch.long)
    cos = (d[t1] * d[t2]).sum(-1) / (r[t1] * r[t2])
    G = 1.0 + c0 ** 2 / d0 ** 2 - c0 ** 2 / (d0 ** 2 + (1.0 + cos) ** 2)
    zeta = torch.zeros(n_b, dtype=pos.dtype).index_add(0, t1, f[t2] * G)
    B = (1.0 + a0 * zeta) ** (-delt)
    # B_ji: the bond (j,i)
    key = {(int(a), int(c)): e for e, (a, c) in enumerate(zip(i, j))}
    rev = torch.as_tensor([key[(int(c), int(a))] for a, c in zip(i, j)], dtype=torch.long)
    av = -np.sqrt(2.0 * S) * b * (r - r0)
    pair = f * D / (S - 1.0) * (torch.exp(av) - (B + B[rev]) / 2 * S * torch.exp(av / S))
    return 0.5 * pair.sum()


def e_rebo(pos, box, g, params, pos_np):
    """REBO solid-carbon energy (REBOsolidcarbon.f90:27-92) over explicit
    neighbour sets: per directed bond (i, j) the sigma-pi term
    (1 + sum_{k != j} fc(r_ik) g(cos theta_jik))^-1/2 and the dihedral term
    T sum_{k != j} sum_{l != i} fc(r_ik) fc(r_jl) sin^2(dihedral kijl), the bond
    energy counted once from its lower-index atom."""
    A, Q, al = params[0:3]
    B, be = params[3:6], params[6:9]
    T, gc = params[9], params[10:16]
    R1, R2 = params[16], params[17]
    i, j = pairs_within(pos_np, box.numpy(), g, g, R2)
    d = disp(pos, box, i, j)
    r = torch.sqrt((d * d).sum(-1))
    f = fc_cos(r, R1, R2)
    bonds = {}
    for e, (a, c) in enumerate(zip(i, j)):
        bonds.setdefault(int(a), []).append(e)
    E = torch.zeros((), dtype=pos.dtype)
    for e, (a, c) in enumerate(zip(i, j)):
        a, c = int(a), int(c)
        if c < a:
            continue
        def bsp(e0, atom):
            s = torch.zeros((), dtype=pos.dtype)
            for e2 in bonds[atom]:
                if e2 != e0:
                    cs = (d[e0] * d[e2]).sum() / (r[e0] * r[e2])
                    s = s + f[e2] * sum(gc[k] * cs ** k for k in range(6))
            return (1.0 + s) ** -0.5
        back = [e2 for e2 in bonds[c] if int(j[e2]) == a][0]
        bdh = torch.zeros((), dtype=pos.dtype)
        for eq in bonds[a]:
            if eq == e:
                continue
            for el in bonds[c]:
                if int(j[el]) == a:
                    continue
                aa = (d[e] * d[e]).sum()
                ab = (d[e] * d[eq]).sum()
                ac = (d[e] * d[el]).sum()
                bc = (d[eq] * d[el]).sum()
                bdh = bdh + f[eq] * f[el] * (1.0 - (aa * bc - ab * ac) ** 2 / (aa * r[eq] ** 2 - ab ** 2)
                                                   / (aa * r[el] ** 2 - ac ** 2))
        rr = r[e]
        att = sum(B[k] * torch.exp(-be[k] * rr) for k in range(3))
        E = E + f[e] * ((1 + Q / rr) * A * torch.exp(-al * rr) - ((bsp(e, a) + bsp(back, c)) / 2 + T * bdh) * att)
    return E


def graphene_normals(pos, box, gc, rnn, pos_np):
    """n_i of graphenenorm: three nearest neighbours within rnn, fixed topology."""
    i, j = pairs_within(pos_np, box.numpy(), gc, gc, rnn)
    nn = {}
    for a, c in zip(i, j):
        nn.setdefault(int(a), []).append(int(c))
    idx = np.array([nn[int(a)] for a in gc])
    assert idx.shape[1] == 3
    gc_t = torch.as_tensor(np.asarray(gc))
    d = [disp(pos, box, gc_t, torch.as_tensor(idx[:, k])) for k in range(3)]
    n = torch.cross(d[1] - d[0], d[0] - d[2], dim=-1)
    sgn = torch.where(n[:, 2] < 0, -torch.ones_like(n[:, 2]), torch.ones_like(n[:, 2])).detach()
    n = n * sgn[:, None]
    return n / torch.sqrt((n * n).sum(-1, keepdim=True))


def e_cos(pos, box, gc, gm, params, morse, rnn, pos_np, simplified=False):
    if morse:
        D, r0, a, delt, R1, R2 = params[:6]
    else:
        eps, sig, delt, R1, R2 = params[:5]
    gc = np.asarray(gc)
    if simplified:
        nrm = torch.zeros(len(gc), 3, dtype=pos.dtype)
        nrm[:, 2] = 1.0
    else:
        nrm = graphene_normals(pos, box, gc, rnn, pos_np)
    row = {int(a): k for k, a in enumerate(gc)}
    i, j = pairs_within(pos_np, box.numpy(), gc, gm, R2)
    d = disp(pos, box, i, j)
    r = torch.sqrt((d * d).sum(-1))
    nv = nrm[torch.as_tensor([row[int(a)] for a in i], dtype=torch.long)]
    V3 = (torch.abs((nv * d).sum(-1)) / r) ** delt
    f = fc_cos(r, R1, R2)
    if morse:
        V2 = torch.exp(-a * (r - r0))
        return (D * (V2 * V2 - 2.0 * V2 * V3) * f).sum()
    V2 = (sig / r) ** 6
    return (4 * eps * (V2 * V2 - V2 * V3) * f).sum()


def forces(efun, pos_np):
    pos = torch.tensor(pos_np, dtype=torch.float64, requires_grad=True)
    e = efun(pos)
    (g,) = torch.autograd.grad(e, pos)
    return float(e.detach()), -g.numpy()
