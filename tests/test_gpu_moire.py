"""The moire fitting driver (pfmds_amd/moire.py) with md() on the GPU.

One sweep of the coordinate search (twelve evaluations of two 30-step nvms
relaxations each, from the start xyz) and one evaluation that continues from
the previous evaluation's final xyz, run with one md() at a time and with
eight at once: fit_out.txt and every final file byte-identical.  The first
evaluation's binding energy, height difference and corrugation of both
systems against the CPU oracle running md()'s schedule on the same inputs,
to 1e-6 (fit_out.txt prints six decimals).
"""
import io
import os

import numpy as np
import pytest

from pfmds_amd import capi, configs, moire, system
from tests.harness import oracle_ctx

pytestmark = pytest.mark.gpu


def fit_once(root, jobs):
    path = configs.moire_fit_case(os.path.join(root, "in"), steps=30, spread=0.1)
    fs = moire.read_fitting_parameters(path)
    log = io.StringIO()
    fit = moire.Fitter(fs, out_period=1000, jobs=jobs, out=log)
    params, error = fit.fit(1e9, 1e9, max_sweeps=1)
    fit.calc_error(params, False)
    fit.close()
    return fit, fs


def test_fit_sweep_concurrent_equals_sequential_and_oracle(tmp_path):
    print("[moire] jobs=8", flush=True)
    f8, fs8 = fit_once(str(tmp_path / "p"), 8)
    print("[moire] jobs=1", flush=True)
    f1, fs1 = fit_once(str(tmp_path / "s"), 1)
    assert len(f8.points) == 13 and not f8.points[-1].from_init
    out8, out1 = fs8.out_path, fs1.out_path
    a = open(os.path.join(out8, "fitfit_out.txt")).read()
    b = open(os.path.join(out1, "fitfit_out.txt")).read()
    assert a == b
    for sim in (1, 7, 13):
        for name in ("final_1.txt", "final_2.txt", "final_gr_cu_1.xyz", "final_gr_cu_2.xyz"):
            fa = open(os.path.join(out8, f"fit{sim:06d}_{name}")).read()
            fb = open(os.path.join(out1, f"fit{sim:06d}_{name}")).read()
            if name.endswith(".txt"):
                # the all_out line starts with the output prefix, which holds the folder
                fa, fb = fa[32:], fb[32:]
            assert fa == fb, (sim, name)

    # evaluation 1 against the oracle: md()'s schedule on the same inputs
    p = f8.points[0]
    for i in (0, 1):
        chk = str(tmp_path / f"chk{i}")
        os.makedirs(chk)
        for name in os.listdir(fs8.input_path):
            src = os.path.join(fs8.input_path, name)
            if os.path.isfile(src):
                os.symlink(src, os.path.join(chk, name))
        os.symlink(os.path.join(fs8.input_path, fs8.start_xyz[i]), os.path.join(chk, fs8.xyz_file[i]))
        s = system.read_files(chk, fs8.settings[i])
        for it in s.interactions:
            if it.name == "morsec":
                it.params[:4] = [p.params[2], p.params[0], p.params[3], p.params[1]]
        o = oracle_ctx(s)
        _, last = system.run_md(s, o, out_period=1000)
        e = o.energies()
        be = (e["pe"] + e["ke"] - fs8.zero_energy[i]) / fs8.c_num[i]
        pos = o.state()[0]
        sp = np.array(s.species)
        zc = [pos[j, 2] for nm in moire.C_NAMES for j in np.nonzero(sp == nm)[0]]
        zm = [pos[j, 2] for nm in moire.METAL_NAMES for j in np.nonzero(sp == nm)[0]]
        a1 = moire.position_analysis(zc, fs8.z, 1000.0)
        a2 = moire.position_analysis(zm, fs8.z, 1000.0)
        print(f"  system {i + 1}: be {p.be[i]:.9f} oracle {be:.9f}, grd {p.grd[i]:.9f} oracle {a1[2] - a1[1]:.9f}",
              flush=True)
        assert p.be[i] == pytest.approx(be, abs=1e-6)
        assert p.bd[i] == pytest.approx(a1[0] - a2[0], abs=1e-6)
        assert p.grd[i] == pytest.approx(a1[2] - a1[1], abs=1e-6)
