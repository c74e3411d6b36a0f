"""Shared test helpers: contexts on the HIP library and on the CPU oracle
(oracle/liborc.so), built from the same System description."""
import os

import numpy as np

from pfmds_amd import capi, system

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liborc.so")


def oracle_lib():
    return capi.load_library(ORACLE_LIB, "orc")


def oracle_ctx(s):
    lib = oracle_lib()
    return system.build_context(s, lambda b, p, v, m: capi.Context(lib, "orc", b, p, v, m))


def gpu_ctx(s, precision=capi.FP64):
    return system.build_context(s, lambda b, p, v, m: capi.Context.gpu(b, p, v, m, precision))


def list_sets(ctx, inter, k):
    nnum, nl, less = ctx.get_list(inter, k)
    return [tuple(sorted(nl[r, : nnum[r]].tolist())) for r in range(len(nnum))], less


def rel_inf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
