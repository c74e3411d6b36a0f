"""CPU pins of the learned surrogate's mathematics (no GPU):
  - the O(neighbours) moment-tensor form of the angular descriptors (the form
    the GPU kernels use) equals the textbook sum over neighbour pairs j != k;
  - the hand-derived forward-over-reverse weight gradient (the form the
    training kernel uses) equals torch's double backward of the force loss,
    and its energy part the plain backward.
"""
import numpy as np
import torch

from pfmds_amd import configs
from tests import sur_ref as S


def small_slab(seed=3):
    s = configs.graphene_on_cu(cx=5, cy=3, gx=5, gy=3, layers=2, seed=seed, jitter=0.2)
    g1 = s.group_indices(1) - 1
    g2 = s.group_indices(4) - 1
    return s, g1, g2


def test_angular_moments_equal_pair_sums():
    s, g1, g2 = small_slab()
    x = torch.tensor(s.pos)
    P = S.pairs(s.pos, s.box, g1, g2, 6.0)
    a = S.descriptors(x, s.box, g1, g2, 6.0, P, brute=True)
    b = S.descriptors(x, s.box, g1, g2, 6.0, P, brute=False)
    assert a.shape == (len(g1), 40)
    assert float(a[:, 24:].abs().min()) >= 0.0 and float(a[:, 24:].abs().max()) > 1.0
    np.testing.assert_allclose(b.numpy(), a.numpy(), rtol=1e-12, atol=1e-12)


def test_descriptors_invariant_under_rotation_and_permutation():
    s, g1, g2 = small_slab(seed=4)
    # a rotation about z by 90 degrees maps the periodic box onto a box with
    # x and y swapped; descriptors are per-atom invariants
    P = S.pairs(s.pos, s.box, g1, g2, 6.0)
    G = S.descriptors(torch.tensor(s.pos), s.box, g1, g2, 6.0, P, brute=False).numpy()
    rot = s.pos[:, [1, 0, 2]] * np.array([-1.0, 1.0, 1.0])
    box_r = s.box[[1, 0, 2]]
    rot = np.mod(rot, box_r)
    Pr = S.pairs(rot, box_r, g1, g2, 6.0)
    Gr = S.descriptors(torch.tensor(rot), box_r, g1, g2, 6.0, Pr, brute=False).numpy()
    np.testing.assert_allclose(Gr, G, rtol=1e-10, atol=1e-12)


def test_manual_weight_gradient_matches_double_backward():
    rng = np.random.default_rng(5)
    n = 37
    G = rng.normal(size=(n, 40))
    t = rng.normal(size=(n, 40))
    mu, isg = rng.normal(size=40) * 0.1, 1.0 + 0.1 * rng.uniform(size=40)
    theta = S.init_theta(mu, isg, seed=6, b3=-1.0)
    cE, cF = 0.37, -0.81
    th = torch.tensor(theta, requires_grad=True)
    Gt = torch.tensor(G, requires_grad=True)
    E = S.network(Gt, th)
    (gG,) = torch.autograd.grad(E.sum(), Gt, create_graph=True)
    # t is a tangent in normalised coordinates: grad_x NN . t = gG/isg . t
    target = cE * E.sum() + cF * (gG / torch.tensor(isg) * torch.tensor(t)).sum()
    (want,) = torch.autograd.grad(target, th)
    want = want.numpy().copy()
    want[:80] = 0.0
    got = S.manual_grad(G, t, theta, cE, cF)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)


def test_loss_gradient_factorisation_on_a_slab():
    """The force-loss gradient through the tangent descriptors equals the
    double backward of the loss over positions (small slab, float64)."""
    s, g1, g2 = small_slab(seed=7)
    P = S.pairs(s.pos, s.box, g1, g2, 6.0)
    x = torch.tensor(s.pos)
    G = S.descriptors(x, s.box, g1, g2, 6.0, P, brute=False).numpy()
    mu, isg = G.mean(0), 1.0 / (G.std(0) + 1e-3)
    theta = S.init_theta(mu, isg, seed=8, b3=-7.0)
    rng = np.random.default_rng(9)
    f_ref = rng.normal(size=s.pos.shape) * 0.3
    e_ref = -7.3 * len(g1)
    w_e, w_f = 0.5, 2.0
    loss, want = S.loss_and_grad(s.pos, s.box, g1, g2, 6.0, theta, f_ref, e_ref, w_e, w_f, P)
    # the kernels' route: forces, residual, tangent descriptors, manual gradient
    E, F = S.energy_forces(s.pos, s.box, g1, g2, 6.0, theta, P)
    n1, n2 = len(g1), len(g2)
    res = np.zeros_like(F)
    res[g2] = F[g2] - f_ref[g2]
    xt = torch.tensor(s.pos, requires_grad=True)
    Gt = S.descriptors(xt, s.box, g1, g2, 6.0, P, brute=False)
    # dG along the displacement field res: forward-mode by a double vjp
    v = torch.zeros_like(Gt, requires_grad=True)
    (gv,) = torch.autograd.grad(Gt, xt, grad_outputs=v, create_graph=True)
    (dG,) = torch.autograd.grad(gv, v, grad_outputs=torch.tensor(res))
    t = dG.numpy() * isg
    cE = 2 * w_e * (E - e_ref) / n1 / n1
    cF = -2 * w_f / (3 * n2)
    got = S.manual_grad(G, t, theta, cE, cF)
    np.testing.assert_allclose(got, want, rtol=1e-8, atol=1e-10 * np.abs(want).max())
    assert loss[2] > 0 and loss[1] > 0
