"""The learned surrogate on the GPU against its CPU definition
(tests/sur_ref.py) on a small graphene/Cu slab.

Bars: both lists equal brute-force neighbour sets (the converse list is the
transpose of the first); descriptors within 1e-4 of each feature's largest
value over the slab (FP32 arithmetic; the (1 - cos)^4 angular functions
combine degree sums that cancel to a few parts in 1e4 on the small metal
features, far below the BF16 rounding of the network's scaled inputs);
energy within 1e-5 and forces within 1e-3 (relative RMSE) of a CPU replay
that rounds to BF16 where the kernels do (network inputs, hidden activations
and the backward signals), forces within 3e-2 of the float64 network (the
BF16 rounding itself); the sum of the surrogate forces zero to 1e-10 of
sum |F| (every d_ij enters twice with opposite signs); the training loss and
gradient of pfm_sur_loss_grad within 3e-2 of torch's double backward
(cosine above 0.999), zero for the fixed scaling, and the same bits on a
repeated call; training on frames of the plain potential cuts the
held-out force error below half of both the untrained network's and the
error of leaving the graphene terms out; the judge moves a run onto the
plain potential when the force error passes the threshold.
"""
import numpy as np
import pytest
import torch

from pfmds_amd import capi, configs, surrogate, system
from tests import sur_ref as S

pytestmark = pytest.mark.gpu

RC = surrogate.SUR_RC


def slab(seed=51, jitter=0.15):
    return configs.graphene_on_cu(cx=10, cy=6, gx=10, gy=6, layers=3, seed=seed, jitter=jitter)


def sur_ctx(s, theta, precision=capi.MIXED):
    _, sur_s, _ = surrogate.split_system(s)
    tr = surrogate.training_system(sur_s)
    ctx = system.build_context(tr, lambda b, p, v, m: capi.Context.gpu(b, p, v, m, precision))
    ctx.sur_set_params(1, theta)
    ctx.run("nve", 0.5, 0, 0, True)
    return ctx, tr


def groups(s):
    return s.group_indices(1) - 1, s.group_indices(4) - 1


def ref_descriptors(s, brute=True):
    g1, g2 = groups(s)
    P = S.pairs(s.pos, s.box, g1, g2, RC)
    return S.descriptors(torch.tensor(s.pos), s.box, g1, g2, RC, P, brute=brute).numpy(), P


def theta_for(G, seed=3, b3=-7.0, scale=1.5):
    mu, isg = G.mean(0), 1.0 / (G.std(0) + 1e-3)
    return S.init_theta(mu, isg, seed=seed, scale=scale, b3=b3)


def emulate(G, theta):
    """E_i and dE_i/dG_i with the kernels' BF16 rounding points."""
    p = S.unpack(S.round_weights(theta))
    f32 = np.float32
    x = S.bf16_round((G.astype(f32) - p["mu"].astype(f32)) * p["isg"].astype(f32))
    a1 = np.tanh(x @ p["W1"].T + p["b1"])
    A1 = S.bf16_round(a1)
    a2 = np.tanh(A1 @ p["W2"].T + p["b2"])
    E = a2 @ p["w3"] + p["b3"]
    g2 = S.bf16_round(p["w3"] * (1 - a2 * a2))
    gz1 = S.bf16_round((g2 @ p["W2"]) * (1 - a1 * a1))
    gG = (gz1 @ p["W1"]) * p["isg"]
    return E, gG


def forces_from_gG(s, P, gG):
    g1, g2 = groups(s)
    x = torch.tensor(s.pos, requires_grad=True)
    G = S.descriptors(x, s.box, g1, g2, RC, P, brute=False)
    (g,) = torch.autograd.grad(G, x, grad_outputs=torch.tensor(gG))
    return -g.numpy()


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_descriptors_match_textbook_sums(precision):
    s = slab()
    G, _ = ref_descriptors(s)
    ctx, _ = sur_ctx(s, theta_for(G), precision)
    g1, _ = groups(s)
    Gg = ctx.sur_descriptors(1, len(g1))
    print(f"[sur desc] n={s.n} rows={len(g1)} max|G|={np.abs(G).max():.3g}", flush=True)
    scale = np.abs(G).max(0)
    err = np.abs(Gg - G) / np.maximum(scale, 1e-12)
    print(f"  max relative error {err.max():.2e} (feature {int(np.argmax(err.max(0)))})", flush=True)
    assert err.max() < 1e-4


def test_lists_are_brute_force_sets_and_converse():
    from tests.test_oracle_pins import brute_sets
    from tests.harness import list_sets
    s = slab(seed=56)
    G, _ = ref_descriptors(s, brute=False)
    ctx, tr = sur_ctx(s, theta_for(G))
    (g1n, g2n, _, rc, _), _ = tr.interactions[0].lists
    g1, g2 = s.group_indices(g1n), s.group_indices(g2n)
    l1, _ = list_sets(ctx, 1, 1)
    l2, _ = list_sets(ctx, 1, 2)
    assert l1 == brute_sets(s.pos, s.box, g1 - 1, g2 - 1, rc)
    assert l2 == brute_sets(s.pos, s.box, g2 - 1, g1 - 1, rc)
    print(f"[sur lists] rows {len(l1)} / {len(l2)}, entries {sum(map(len, l1))} = {sum(map(len, l2))}", flush=True)


def test_energy_and_forces_against_cpu():
    s = slab(seed=52)
    G, P = ref_descriptors(s, brute=False)
    theta = theta_for(G, seed=4)
    ctx, _ = sur_ctx(s, theta)
    g1, g2 = groups(s)
    Gg = ctx.sur_descriptors(1, len(g1))
    e_gpu = ctx.energies()["e_inter"][0]
    f_gpu = ctx.state()[2]
    Ei, gG = emulate(Gg, theta)
    f_emu = forces_from_gG(s, P, gG)
    b3 = S.unpack(theta)["b3"]
    print(f"[sur E/F] E gpu {e_gpu:.8f} emu {Ei.sum():.8f}", flush=True)
    assert abs(e_gpu - Ei.sum()) < 1e-5 * np.abs(Ei - b3).sum()
    rel = np.sqrt(np.mean((f_gpu - f_emu) ** 2)) / np.sqrt(np.mean(f_emu ** 2))
    e64, f64 = S.energy_forces(s.pos, s.box, g1, g2, RC, S.round_weights(theta), P)
    rel64 = np.sqrt(np.mean((f_gpu - f64) ** 2)) / np.sqrt(np.mean(f64 ** 2))
    print(f"  force rel RMSE vs BF16 replay {rel:.2e}, vs float64 network {rel64:.2e}", flush=True)
    assert rel < 1e-3
    assert rel64 < 3e-2
    assert abs(e_gpu - e64) < 3e-2 * np.abs(Ei - b3).sum()
    # Newton's third law of the gather formulation
    assert np.linalg.norm(f_gpu.sum(0)) < 1e-10 * np.linalg.norm(f_gpu, axis=1).sum()


def test_loss_gradient_against_double_backward():
    s = slab(seed=53)
    G, P = ref_descriptors(s, brute=False)
    theta = theta_for(G, seed=5)
    ctx, _ = sur_ctx(s, theta)
    g1, g2 = groups(s)
    rng = np.random.default_rng(6)
    f_ref = rng.normal(size=s.pos.shape) * 0.5
    e_ref = float(S.unpack(theta)["b3"]) * len(g1) + 3.0
    w_e, w_f = 0.7, 1.3
    loss, grad = ctx.sur_loss_grad(1, f_ref, e_ref, w_e, w_f)
    lref, gref = S.loss_and_grad(s.pos, s.box, g1, g2, RC, S.round_weights(theta), f_ref, e_ref, w_e, w_f, P)
    cos = grad @ gref / (np.linalg.norm(grad) * np.linalg.norm(gref))
    rel = np.linalg.norm(grad - gref) / np.linalg.norm(gref)
    print(f"[sur grad] loss gpu {loss} ref {lref} cos {cos:.6f} rel {rel:.2e}", flush=True)
    np.testing.assert_allclose(loss, lref, rtol=3e-2)
    assert cos > 0.999 and rel < 3e-2
    # each block of parameters on its own
    for name, a, b in (("W1", 80, 2640), ("b1", 2640, 2704), ("W2", 2704, 6800), ("b2", 6800, 6864),
                       ("w3", 6864, 6928), ("b3", 6928, 6929)):
        rb = np.linalg.norm(grad[a:b] - gref[a:b]) / np.linalg.norm(gref[a:b])
        print(f"  {name}: rel {rb:.2e}", flush=True)
        assert rb < 5e-2, name
    # mu and isg are fixed: zero gradient
    assert not grad[:80].any()
    # every partial sum is reduced in a fixed order: a second call gives the same bits
    loss2, grad2 = ctx.sur_loss_grad(1, f_ref, e_ref, w_e, w_f)
    assert np.array_equal(loss, loss2) and np.array_equal(grad, grad2)


def test_training_gains_over_untrained_and_over_no_term():
    s = slab(seed=54, jitter=0.05)
    print(f"[sur train] n={s.n}", flush=True)
    frames = surrogate.generate_frames(s, [300.0, 600.0, 900.0], 12, every=20, warmup=100)
    test_idx = np.arange(len(frames)) % 6 == 5
    train_f, test_f = frames.subset(np.nonzero(~test_idx)[0]), frames.subset(np.nonzero(test_idx)[0])
    _, sur_s, _ = surrogate.split_system(s)
    tr = surrogate.training_system(sur_s)
    ctx = system.build_context(tr, lambda b, p, v, m: capi.Context.gpu(b, p, v, m, capi.MIXED))
    g1, g2 = groups(s)
    mu, isg = surrogate.feature_scaling(ctx, 1, train_f, len(g1))
    theta0 = surrogate.init_theta(mu, isg, float(np.mean(train_f.energy)) / len(g1), seed=1)
    ctx.sur_set_params(1, theta0)
    rms_label = float(np.sqrt(np.mean(test_f.force[:, g2] ** 2)))
    e0 = surrogate.force_rmse(ctx, test_f, g2)
    theta, hist = surrogate.train(ctx, 1, train_f, theta0, 400, lr=3e-3, lr_end=3e-4)
    e1 = surrogate.force_rmse(ctx, test_f, g2)
    print(f"  held-out force RMSE: untrained {e0:.4f}, trained {e1:.4f}, no term {rms_label:.4f} eV/A; "
          f"loss {hist[:20, 0].mean():.3e} -> {hist[-20:, 0].mean():.3e}", flush=True)
    assert e1 < 0.5 * e0 and e1 < 0.5 * rms_label


def test_judge_switches_to_plain_potential():
    s = slab(seed=55)
    G, _ = ref_descriptors(s, brute=False)
    _, sur_s, k = surrogate.split_system(s)
    # untrained and weak: the graphene keeps its shape over the run, while the
    # forces stay far from the plain ones
    theta = theta_for(G, seed=7, scale=0.3)
    runs = []
    for thr, group in ((1e-3, None), (1e6, None), (1e-3, s.all_moving), (1e6, s.all_moving)):
        sc = surrogate.surrogate_context(sur_s, theta)
        pc = system.build_context(s, lambda b, p, v, m: capi.Context.gpu(b, p, v, m, capi.MIXED))
        moving = s.group_indices(s.all_moving) - 1
        jr = surrogate.JudgedRun(sc, pc, moving, judge_every=10, threshold=thr, group=group)
        end = jr.run("nvt", 0.5, 0, 35)
        runs.append(jr)
        print(f"[judge] threshold {thr} group {group}: judged {jr.judged} switch {jr.switch_step}", flush=True)
        assert [j[0] for j in jr.judged] == ([10] if thr < 1 else [10, 20, 30])
        assert (end is pc) == (thr < 1)
    assert runs[0].switch_step == 10 and runs[1].switch_step is None
    # the judge on the GPU (state copy and RMSE) matches the one through the host
    for a, b in ((runs[0], runs[2]), (runs[1], runs[3])):
        assert a.switch_step == b.switch_step
        for (sa, ra), (sb, rb) in zip(a.judged, b.judged):
            assert sa == sb and rb == pytest.approx(ra, rel=1e-12)


def test_surrogate_runs_are_bit_identical():
    # the converse rows are filled in a varying order; the gather's fixed-point
    # sums make the forces independent of it
    s = slab(seed=57)
    G, _ = ref_descriptors(s, brute=False)
    _, sur_s, _ = surrogate.split_system(s)
    theta = theta_for(G, seed=9, scale=0.3)
    out = []
    for graphs in (True, True, False):
        c = surrogate.surrogate_context(sur_s, theta)
        c.set_graphs(graphs)
        c.run("nvt", 0.5, 0, 30, True)
        out.append(c.state())
        assert (c.graph_count()[1] > 0) == graphs
        c.close()
    # twice with captured blocks, once launched step by step
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64])
def test_fused_inference_equals_separate_kernels(precision):
    """The one-kernel inference pass (k_sur_fused) against the descriptor,
    network and backward kernels (pfm_set_fusion(h, 0)): the same bits for
    forces, energies and 30 nvt steps."""
    s = slab(seed=58)
    G, _ = ref_descriptors(s, brute=False)
    _, sur_s, _ = surrogate.split_system(s)
    theta = theta_for(G, seed=11, scale=0.3)
    out = []
    for fuse in (True, False):
        c = surrogate.surrogate_context(sur_s, theta, precision)
        c.set_fusion(fuse)
        c.run("nvt", 0.5, 0, 0, True)
        f0, e0 = c.state()[2], c.energies()["e_inter"]
        c.run("nvt", 0.5, 1, 30, True)
        out.append((f0, e0, c.state(), c.energies()["e_inter"]))
        c.close()
    (fa, ea, sa, ea2), (fb, eb, sb, eb2) = out
    assert np.array_equal(fa, fb) and np.array_equal(ea, eb) and np.array_equal(ea2, eb2)
    for x, y in zip(sa, sb):
        assert np.array_equal(x, y)
