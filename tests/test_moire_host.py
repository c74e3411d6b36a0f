"""Host side of the moire fitting driver (pfmds_amd/moire.py) on the CPU.

The position analysis is pinned by closed forms on a hand-made xyz file; the
file reader by a hand-written fitting file; the search and its file chain by
a full fit in which tests/fake_md.py stands in for the md() binary with a
separable quadratic error whose minimum is known: the golden-section loop of
run_gr_moire_fitting.f90:47-101 must land within its final bracket of it,
the bracket must shrink by the golden ratio at every step, every
golden step must continue from the previous evaluation's final xyz, and the
results must not depend on how many md() runs go at once.
"""
import io
import os
import stat
import sys

import numpy as np
import pytest

from pfmds_amd import moire, system

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import fake_md  # noqa: E402


def write_xyz(path, rows, box=(10.0, 10.0, 30.0)):
    with open(path, "w") as f:
        f.write(f"{len(rows)}\n")
        f.write(f'Lattice=" {box[0]} 0 0 0 {box[1]} 0 0 0 {box[2]} " Properties=pos:R:3:vel:R:3:mass:R:1:species:S:1\n')
        for (x, y, z, sp) in rows:
            f.write(f"{x} {y} {z} 0 0 0 12.0 {sp}\n")


def test_position_analysis_closed_forms(tmp_path):
    rows = [(0, 0, 5.0, "CU"), (1, 0, 7.0, "C"), (2, 0, 2.0, "C"), (3, 0, 6.5, "CU_fixed"), (4, 0, 9.0, "C_a"),
            (5, 0, 3.0, "X"), (6, 0, 1000.0, "C"), (7, 0, 4.0, "C_b"), (8, 0, 3.0, "C")]
    p = tmp_path / "f.xyz"
    write_xyz(p, rows)
    a1, a2 = moire.gr_on_cu_analysis(str(p), 3.0)
    # carbon above z=3 and below 1000 (strict): 7, 9, 4 -> mean 20/3, min 4, max 9
    np.testing.assert_allclose(a1, [20.0 / 3.0, 4.0, 9.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(a2, [5.75, 5.0, 6.5], rtol=0, atol=1e-15)
    av, mi, ma = moire.position_analysis([], 0.0, 1.0)
    assert np.isnan(av) and mi == 1.0 and ma == 0.0


def make_inputs(d, interaction="ljc"):
    """Two settings files (their 3rd line names the xyz, an ljc/morsec line
    the parameter file), start xyz files and the bounds files."""
    os.makedirs(d, exist_ok=True)
    for k, name in enumerate(("sys1", "sys2")):
        with open(os.path.join(d, f"{name}.txt"), "w") as f:
            f.write(f"md_step_limit 10\nlogfilename log.txt\ninit_xyz_filename {name}.xyz\n"
                    f"some other line\n{interaction} pfile.txt\n1 2 10 9.0 10\n")
        rows = [(0.5 * i, 0, 2.0 + 0.01 * i, "CU") for i in range(4)]
        rows += [(0.5 * i, 1, 6.0 + 0.05 * i + k, "C") for i in range(fake_md.C_NUM[k])]
        write_xyz(os.path.join(d, f"start_{name}.xyz"), rows)
    if interaction == "ljc":
        lo, hi = "0.001 2.9 0.1\n1.0 9.5\nF\n", "0.03 3.8 0.9\n"
    else:
        lo, hi = "0.01 3.0 1.1 0.2\n2.0 8.0\nF\n", "0.09 3.9 1.9 0.8\n"
    open(os.path.join(d, "pmin.txt"), "w").write(lo)
    open(os.path.join(d, "pmax.txt"), "w").write(hi)
    open(os.path.join(d, "pfile.txt"), "w").write("original\n")


def write_fit_file(path, inp, out, interaction="ljc"):
    vals = [interaction, "fit", "sys1.txt", "start_sys1.xyz", fake_md.C_NUM[0], fake_md.ZERO[0], "final1.txt",
            "sys2.txt", "start_sys2.xyz", fake_md.C_NUM[1], fake_md.ZERO[1], "final2.txt", "pmin.txt", "pmax.txt",
            4.0, fake_md.BE0, fake_md.G1, fake_md.G2]
    names = ["interaction", "prefix", "settings1", "start1", "c_num1", "zero1", "final1", "settings2", "start2",
             "c_num2", "zero2", "final2", "min", "max", "z", "be0", "grd1", "grd2"]
    with open(path, "w") as f:
        f.write(f"{'input_path:':<16}{inp}\n{'out_path:':<16}{out}\n")
        for n, v in zip(names, vals):
            f.write(f"{n} {v}\n")


def test_read_fitting_parameters(tmp_path):
    for inter in ("ljc", "morsec"):
        inp = str(tmp_path / f"in_{inter}") + "/"
        make_inputs(inp, inter)
        write_fit_file(tmp_path / f"fit_{inter}.txt", inp, str(tmp_path) + "/", inter)
        fs = moire.read_fitting_parameters(str(tmp_path / f"fit_{inter}.txt"))
        assert fs.input_path == inp and fs.interaction == inter and fs.param_file == "pfile.txt"
        assert fs.xyz_file == ("sys1.xyz", "sys2.xyz") and fs.c_num == fake_md.C_NUM
        assert fs.rcut == ((1.0, 9.5) if inter == "ljc" else (2.0, 8.0))
        if inter == "ljc":
            # file order eps sig delt; fitted vector (sig, delt, eps, 0)
            np.testing.assert_array_equal(fs.min_params, [2.9, 0.1, 0.001, 0.0])
            np.testing.assert_array_equal(fs.max_params, [3.8, 0.9, 0.03, 0.0])
        else:
            # file order d r a delt; fitted vector (r, delt, d, a)
            np.testing.assert_array_equal(fs.min_params, [3.0, 0.2, 0.01, 1.1])
            np.testing.assert_array_equal(fs.max_params, [3.9, 0.8, 0.09, 1.9])
        # the candidate file reads back through the settings reader's parameter order
        p = fs.min_params + 0.25
        txt = moire.param_file_text(inter, p, fs.rcut)
        vals = [float(x) for x in txt.split()[:-1]]
        want = [p[2], p[0], p[1]] if inter == "ljc" else [p[2], p[0], p[3], p[1]]
        assert vals == want + list(fs.rcut) and txt.split()[-1] == "F"
        n = sum(system.PARAM_LINES[inter])
        assert len(want) + 2 + 1 == n


def fake_exe(tmp_path):
    exe = tmp_path / "fake_md"
    exe.write_text(f"#!/bin/sh\nexec {sys.executable} {os.path.join(HERE, 'fake_md.py')} \"$@\"\n")
    exe.chmod(exe.stat().st_mode | stat.S_IEXEC)
    return str(exe)


def run_fit(tmp_path, jobs, tag, delta_gold=1e-5):
    inp = str(tmp_path / f"in_{tag}") + "/"
    out = str(tmp_path / f"out_{tag}") + "/"
    os.makedirs(out)
    make_inputs(inp)
    write_fit_file(tmp_path / f"fit_{tag}.txt", inp, out)
    log = io.StringIO()
    fit = moire.Fitter(moire.read_fitting_parameters(str(tmp_path / f"fit_{tag}.txt")), jobs=jobs,
                       exe=fake_exe(tmp_path), out=log)
    params, error = fit.fit(delta_gold, 1e9, max_sweeps=1)
    fit.close()
    return fit, params, error, log.getvalue(), inp, out


def test_fit_finds_the_minimum_and_chains_its_runs(tmp_path):
    fit, params, error, log, inp, out = run_fit(tmp_path, 4, "a")
    lo, hi = fit.fs.min_params, fit.fs.max_params
    target = np.array([fake_md.SIG, fake_md.DELT, fake_md.EPS])
    # every coordinate search: 4 starting points, then golden steps until the
    # four errors agree to delta_error_gold or 21 steps
    pts = fit.points
    assert [p.sim for p in pts] == list(range(1, len(pts) + 1))
    k0 = 0
    for k in range(3):
        start = pts[k0:k0 + 4]
        assert all(p.from_init for p in start)
        np.testing.assert_allclose([p.params[k] for p in start],
                                   [lo[k], hi[k], lo[k] + (hi[k] - lo[k]) * moire.GOLD,
                                    hi[k] - (hi[k] - lo[k]) * moire.GOLD], rtol=1e-15)
        j = k0 + 4
        while j < len(pts) and not pts[j].from_init:
            j += 1
        steps = pts[k0 + 4:j]
        assert 1 <= len(steps) <= 21
        # the bracket after n golden steps is (1 - gold)^n of the start (the
        # start term of the energy moves the minimum by about 1e-7)
        width = (hi[k] - lo[k]) * (1 - moire.GOLD) ** len(steps)
        assert abs(params[k] - target[k]) <= width + 1e-4, (k, params[k], target[k], width)
        k0 = j
    assert len(pts) == k0
    # golden steps continue from the previous evaluation's final xyz
    starts = [ln.split()[1] for ln in log.splitlines() if ln.startswith("start: ")]
    assert len(starts) == 2 * len(pts)
    for p in pts:
        for i, name in enumerate(("sys1", "sys2")):
            got = starts[2 * (p.sim - 1) + i]
            want = os.path.realpath(inp + f"start_{name}.xyz") if p.from_init else \
                os.path.realpath(out + f"fit{p.sim - 1:06d}_final_{name}.xyz")
            assert got == want, (p.sim, i)
    # files: one fit_out line per evaluation (i6 + 13 f21.6 fields for ljc),
    # a parameter copy per evaluation, finals with the analysis appended
    lines = open(out + "fitfit_out.txt").read().splitlines()
    assert len(lines) == len(pts) and all(len(ln) == 6 + 21 * 13 for ln in lines)
    for p, ln in zip(pts, lines):
        assert int(ln[:6]) == p.sim
        assert float(ln[-21:]) == pytest.approx(p.error, abs=5e-7)
        assert os.path.exists(inp + f"{p.sim:06d}pfile.txt")
    fin = open(out + "fit000001_final1.txt").read().splitlines()[0]
    assert len(fin) == 32 + 9 + 5 * 20 + 3 * 16
    assert 0 <= error < min(p.error for p in pts[:4])


def test_fit_does_not_depend_on_concurrency(tmp_path):
    a = run_fit(tmp_path, 1, "s", delta_gold=1e-3)
    b = run_fit(tmp_path, 8, "p", delta_gold=1e-3)
    la = open(a[5] + "fitfit_out.txt").read()
    lb = open(b[5] + "fitfit_out.txt").read()
    assert la == lb
    np.testing.assert_array_equal(a[1], b[1])
