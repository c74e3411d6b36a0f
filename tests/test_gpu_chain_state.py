"""Nose-Hoover chain states across handles (pfm_get_chain / pfm_set_chain)
and pfm_set_state's re-sort.

A run continued in a second handle, from the first one's positions,
velocities and chain states, follows the run that stayed in one handle: the
second handle re-sorts and rebuilds its lists at its first step, so the two
agree to rounding (positions within 1e-10 relative after 40 nvt steps, chain
positions within 1e-9), where a chain restarted from rest drifts apart at
once.  Chain states set before the first pfm_run are refused.
"""
import numpy as np
import pytest

from pfmds_amd import capi, configs
from tests.harness import gpu_ctx, rel_inf

pytestmark = pytest.mark.gpu


def test_run_continues_in_another_handle():
    s = configs.graphene_on_cu(cx=10, cy=6, gx=10, gy=6, layers=4, seed=95)
    a = gpu_ctx(s, capi.FP64)
    a.run("nvt", 0.5, 0, 40, True)
    p, v, _ = a.state()
    M = s.nhcs[0].M
    x0, v0 = a.get_chain(1, M)
    assert np.abs(v0).max() > 0
    b = gpu_ctx(s, capi.FP64)
    with pytest.raises(capi.PfmError):
        b.set_chain(1, x0, v0)
    b.run("nvt", 0.5, 0, 0, True)
    b.set_state(p, v)
    b.set_chain(1, x0, v0)
    c = gpu_ctx(s, capi.FP64)  # the same without the chain state
    c.run("nvt", 0.5, 0, 0, True)
    c.set_state(p, v)
    for h in (a, b, c):
        h.run("nvt", 0.5, 41, 80, True)
    pa, pb, pc = a.state()[0], b.state()[0], c.state()[0]
    xa, _ = a.get_chain(1, M)
    xb, _ = b.get_chain(1, M)
    print(f"  continued: {rel_inf(pb, pa):.2e}, chain from rest: {rel_inf(pc, pa):.2e}", flush=True)
    assert rel_inf(pb, pa) < 1e-10 and rel_inf(xb, xa) < 1e-9
    assert rel_inf(pc, pa) > 100 * rel_inf(pb, pa)
    for h in (a, b, c):
        h.close()
