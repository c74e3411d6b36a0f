"""pfm_force_rmse_noalloc, the entry point that allocates nothing, against
pfm_force_rmse through the C-ABI.

Bar: bit-identical results on two calls in a row, the second on a smaller
system than the first (the shared device array keeps the first call's
contents), and a third on the first system again; no device memory taken
by the call (free device memory before and after a call on handles that
never staged a comparison, within 1 MiB); more atoms than the array holds
refused with PFM_E_INVALID and a message naming the limit.
"""
import ctypes as C

import numpy as np
import pytest

from pfmds_amd import capi, configs
from tests.harness import gpu_ctx

pytestmark = pytest.mark.gpu


def free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def pair(s):
    """FP64 and mixed handles of one system after step 0 (their forces differ
    by the mixed path's rounding)."""
    a, b = gpu_ctx(s, capi.FP64), gpu_ctx(s, capi.MIXED)
    for c in (a, b):
        c.run(s.integrators[0].name, s.integrators[0].dt, 0, 0, True)
    return a, b


def test_noalloc_matches_force_rmse_on_shrinking_inputs():
    big = configs.graphene_on_cu(cx=16, cy=9, gx=16, gy=9, layers=4, seed=91)
    small = configs.fcc_cu_rgl(nx=4, seed=92)
    a1, b1 = pair(big)
    a2, b2 = pair(small)
    g_big = big.all_moving
    print(f"[noalloc] {big.n} atoms, then {small.n}", flush=True)
    r1 = b1.force_rmse_noalloc(a1, g_big)
    r2 = b2.force_rmse_noalloc(a2, 1)
    r3 = b1.force_rmse_noalloc(a1, 1)  # graphene group of the first system
    assert r1 == b1.force_rmse(a1, g_big)
    assert r2 == b2.force_rmse(a2, 1)
    assert r3 == b1.force_rmse(a1, 1)
    assert 0.0 < r1 < 1e-4 and 0.0 < r2 < 1e-4
    print(f"  rmse {r1:.3e} {r2:.3e} {r3:.3e}", flush=True)


def test_noalloc_takes_no_device_memory():
    s = configs.fcc_cu_rgl(nx=30, seed=93)  # 108,000 atoms: pfm_force_rmse would stage 6.9 MB
    a, b = pair(s)
    warm_a, warm_b = pair(configs.fcc_cu_rgl(nx=3, seed=94))
    warm_b.force_rmse_noalloc(warm_a, 1)  # the first call on the device looks up the array
    before = free_device_bytes()
    r = b.force_rmse_noalloc(a, 1)
    after = free_device_bytes()
    print(f"[noalloc] free device memory {before} -> {after}", flush=True)
    assert before - after < (1 << 20)
    assert r == b.force_rmse(a, 1)


def test_noalloc_refuses_more_atoms_than_its_array():
    n = (1 << 20) + 1
    box = np.array([400.0, 400.0, 400.0])
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 400.0, (n, 3))
    vel = np.zeros((n, 3))
    mass = np.full(n, 63.546)
    a = capi.Context.gpu(box, pos, vel, mass)
    b = capi.Context.gpu(box, pos, vel, mass)
    for c in (a, b):
        c.set_group(1, np.arange(1, n + 1, dtype=np.int32))
    with pytest.raises(capi.PfmError) as e:
        b.force_rmse_noalloc(a, 1)
    assert e.value.code == capi.E_INVALID and "1048576 atoms" in e.value.msg
