"""Moire parameter fitting of graphene on a metal: golden-section search over
the ljc or morsec parameters (run_gr_moire_fitting.f90:1-111,
fit_gr_moire.f90:1-184), with every md() run on the GPU.

The search, its error function and its files are the reference's:
  - `read_fitting_parameters` reads the fitting-parameters file and the two
    parameter bounds files (set_fitting_parameters, fit_gr_moire.f90:106-184);
  - one error evaluation (`Fitter.evaluate`, calc_error, fit_gr_moire.f90:16-104)
    writes the candidate parameter file, runs md() on the two systems, each
    from its start xyz or from the previous evaluation's final xyz, analyses
    the final positions (`gr_on_cu_analysis`) and scores the binding energy
    per carbon and the graphene corrugation against their targets;
  - `Fitter.fit` is the coordinate-wise golden-section loop
    (run_gr_moire_fitting.f90:47-101).

Every md() run is one process of the Fortran front end
(pfmds_amd/bin/run_md_simulation, one HIP stream on the GPU).  The reference
chains its runs through renamed files in input_path; here each run gets a
private input folder of links instead, so runs that do not depend on each
other (the four starting points of every coordinate's search, eight md()
runs) execute at the same time.  Results do not depend on how many run
together: each run is the same single-system computation.
"""
import argparse
import dataclasses
import math
import os
import shutil
import subprocess
import sys
from typing import List, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "bin", "run_md_simulation")
GOLD = (math.sqrt(5.0) - 1.0) / (math.sqrt(5.0) + 1.0)
LINE = "_" * 80
C_NAMES = ("C", "C_a", "C_b")      # graphene_on_surface_analysis.f90:17
METAL_NAMES = ("CU", "CU_fixed")   # graphene_on_surface_analysis.f90:20


@dataclasses.dataclass
class FitSettings:
    input_path: str
    out_path: str
    interaction: str            # 'ljc' or 'morsec'
    output_prefix: str
    settings: Tuple[str, str]   # md() settings file of each system
    start_xyz: Tuple[str, str]  # starting configuration of each system
    c_num: Tuple[int, int]      # carbon atoms of each system
    zero_energy: Tuple[float, float]
    final_file: Tuple[str, str]
    z: float                    # lower bound of the analysed heights
    be0: float                  # target binding energy per carbon
    grd0: Tuple[float, float]   # target corrugations
    min_params: np.ndarray      # [4]: ljc (sig, delt, eps, 0); morsec (r, delt, d, a)
    max_params: np.ndarray
    rcut: Tuple[float, float]
    param_file: str             # the interaction's parameter file named in settings 1
    xyz_file: Tuple[str, str]   # init_xyz of each settings file


def _tok(line, k):
    return line.split()[k]


def _read_bounds(path, interaction):
    """A parameter bounds file in read_*_parameters order; returns the four
    fitted parameters (fit_gr_moire.f90:156-174) and the two cutoffs."""
    with open(path) as f:
        a = [float(x) for x in f.readline().replace(",", " ").split()]
        line2 = f.readline().replace(",", " ").split()
    r = tuple(float(x) for x in line2[:2]) if len(line2) >= 2 else (0.0, 0.0)
    if interaction == "ljc":
        p = np.array([a[1], a[2], a[0], 0.0])       # eps sig delt -> (sig, delt, eps)
    else:
        p = np.array([a[1], a[3], a[0], a[2]])      # d r a delt -> (r, delt, d, a)
    return p, r


def read_fitting_parameters(path):
    """set_fitting_parameters (fit_gr_moire.f90:106-184): the fitting file, the
    bounds files, the interaction's parameter file (the first settings line
    of system 1 that starts with the interaction name) and each system's
    init_xyz name (third settings line)."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    input_path = lines[0][16:].strip()
    out_path = lines[1][16:].strip()
    v = [ln.split()[1] if len(ln.split()) > 1 else "" for ln in lines[2:20]]
    interaction = v[0]
    if interaction not in ("ljc", "morsec"):
        raise ValueError(f"fitting needs interaction ljc or morsec, got {interaction!r}")
    prefix = v[1]
    s1, x1, c1, z1, f1 = v[2], v[3], int(v[4]), float(v[5]), v[6]
    s2, x2, c2, z2, f2 = v[7], v[8], int(v[9]), float(v[10]), v[11]
    min_file, max_file = v[12], v[13]
    z, be0, g1, g2 = float(v[14]), float(v[15]), float(v[16]), float(v[17])
    pmin, rcut = _read_bounds(os.path.join(input_path, min_file), interaction)
    pmax, _ = _read_bounds(os.path.join(input_path, max_file), interaction)
    if interaction == "ljc":
        pmin[3] = pmax[3] = 0.0
    param_file, xyz = None, []
    for k, sf in enumerate((s1, s2)):
        with open(os.path.join(input_path, sf)) as f:
            sl = [ln.rstrip("\n") for ln in f]
        xyz.append(_tok(sl[2], 1))
        if k == 0:
            n = len(interaction)
            for ln in sl:
                if ln[:n] == interaction:
                    param_file = ln[n:].split()[0]
                    break
    if param_file is None:
        raise ValueError(f"no {interaction} line in {s1}")
    return FitSettings(input_path, out_path, interaction, prefix, (s1, s2), (x1, x2), (c1, c2), (z1, z2), (f1, f2),
                       z, be0, (g1, g2), pmin, pmax, rcut, param_file, (xyz[0], xyz[1]))


def param_file_text(interaction, params, rcut):
    """The parameter file calc_error writes for a candidate (fit_gr_moire.f90:27-45):
    the fitted values in read_*_parameters order, the cutoffs, simplified = F."""
    p = [float(x) for x in params]
    first = [p[2], p[0], p[1]] if interaction == "ljc" else [p[2], p[0], p[3], p[1]]
    fmt = lambda xs: " ".join(repr(float(x)) for x in xs)  # noqa: E731
    return f" {fmt(first)}\n {fmt(rcut)}\n F\n"


def position_analysis(values, minimum, maximum):
    """position_analysis (md_general.f90:400-421): mean, min and max of the
    values strictly inside (minimum, maximum), summed in the given order."""
    av, k, mi, ma = 0.0, 0, maximum, minimum
    for x in values:
        if minimum < x < maximum:
            av += x
            k += 1
            ma = x if x > ma else ma
            mi = x if x < mi else mi
    return (av / k if k else float("nan")), mi, ma


def read_xyz_columns(path):
    with open(path) as f:
        n = int(f.readline().split()[0])
        f.readline()
        rows = [f.readline().split() for _ in range(n)]
    z = np.array([float(r[2]) for r in rows])
    return z, [r[7] if len(r) > 7 else "" for r in rows]


def gr_on_cu_analysis(path, z):
    """gr_on_cu_analysis (graphene_on_surface_analysis.f90:8-30): (mean, min,
    max) height above z of the graphene atoms and of the metal atoms, each
    group ordered by species name, then index."""
    zz, sp = read_xyz_columns(path)
    sp = np.array(sp)
    out = []
    for names in (C_NAMES, METAL_NAMES):
        idx = np.concatenate([np.nonzero(sp == nm)[0] for nm in names])
        out.append(np.array(position_analysis(zz[idx], z, 1000.0)))
    return out[0], out[1]


def _f21(x):
    return f"{'NaN':>21}" if math.isnan(x) else f"{x:21.6f}"


@dataclasses.dataclass
class Point:
    sim: int
    params: np.ndarray
    from_init: bool
    error: float = float("nan")
    be: Tuple[float, float] = (0.0, 0.0)
    bd: Tuple[float, float] = (0.0, 0.0)
    grd: Tuple[float, float] = (0.0, 0.0)
    terms: Tuple[float, float, float] = (0.0, 0.0, 0.0)


class Fitter:
    """Error evaluations and the golden-section search of one fitting run."""

    def __init__(self, fs: FitSettings, out_period=1, precision="fp64", jobs=8, exe=EXE, out=sys.stdout):
        if not os.path.exists(exe):
            raise FileNotFoundError(f"{exe} is not built (run __graft_entry__.build())")
        self.fs, self.out_period, self.precision, self.jobs, self.exe, self.out = fs, out_period, precision, jobs, \
            exe, out
        self.sim_num = 0
        self.work = os.path.join(fs.out_path or ".", ".pfm_fit")
        self.oid = open(os.path.join(fs.out_path, fs.output_prefix + "fit_out.txt"), "w")
        self.points: List[Point] = []

    def close(self):
        self.oid.close()
        shutil.rmtree(self.work, ignore_errors=True)

    def _op(self, sim):
        return f"{self.fs.output_prefix}{sim:06d}_"

    # -- one md() run ---------------------------------------------------------
    def _prepare(self, pt: Point, i: int):
        """Private input folder of run i of a point: links to every file of
        input_path, the candidate parameter file, and the starting xyz under
        the settings' init_xyz name."""
        fs = self.fs
        d = os.path.join(self.work, f"{pt.sim:06d}_{i}")
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
        src_dir = os.path.abspath(fs.input_path or ".")
        for name in os.listdir(src_dir):
            p = os.path.join(src_dir, name)
            if os.path.isfile(p) and name not in (fs.param_file, fs.xyz_file[i]):
                os.symlink(p, os.path.join(d, name))
        with open(os.path.join(d, fs.param_file), "w") as f:
            f.write(param_file_text(fs.interaction, pt.params, fs.rcut))
        if pt.from_init:
            start = os.path.join(src_dir, fs.start_xyz[i])
        else:
            start = os.path.abspath(os.path.join(fs.out_path, self._op(pt.sim - 1) + "final_" + fs.xyz_file[i]))
        os.symlink(start, os.path.join(d, fs.xyz_file[i]))
        with open(os.path.join(d, "pfm_fit_list.txt"), "w") as f:
            f.write(f"1\n{fs.settings[i]} {self._op(pt.sim)}\n")
        return d

    def _launch(self, pt: Point, i: int):
        d = self._prepare(pt, i)
        log = open(os.path.join(d, "stdout.txt"), "w")
        cmd = [self.exe, "-ilist", "pfm_fit_list.txt", "-ipath", d + "/", "-opath", self.fs.out_path,
               "-ofile", self._op(pt.sim) + self.fs.final_file[i], "-op", str(self.out_period), "-prec",
               self.precision]
        return subprocess.Popen(cmd, stdout=log, stderr=subprocess.STDOUT), log, d

    def _finish(self, pt: Point, i: int, d: str):
        """Analysis of run i's final xyz, appended to its final file, and the
        binding energy per carbon read back from it (fit_gr_moire.f90:63-90)."""
        fs = self.fs
        op = os.path.join(fs.out_path, self._op(pt.sim))
        with open(os.path.join(d, "stdout.txt")) as f:
            self.out.write(f"{LINE}\n{f.read()}{LINE}\n")
        arr1, arr2 = gr_on_cu_analysis(op + "final_" + fs.xyz_file[i], fs.z)
        self.out.write(f" gr_on_cu_analysis: {' '.join(repr(float(x)) for x in (*arr1, *arr2))}\n")
        fin = op + fs.final_file[i]
        with open(fin) as f:
            first, rest = f.readline().rstrip(), f.read()
        d3 = arr1 - arr2[0]
        first = first + "".join(f"{x:16.6f}" for x in d3)
        with open(fin, "w") as f:
            f.write(first + "\n" + rest)
        be = float(first[61:81])
        be = (be - fs.zero_energy[i]) / fs.c_num[i]
        return be, arr1[0] - arr2[0], arr1[2] - arr1[1]

    # -- error evaluations ------------------------------------------------------
    def evaluate(self, cands: List[Tuple[np.ndarray, bool]]) -> List[float]:
        """calc_error for consecutive candidates (params, from_init_xyz); the
        md() runs of candidates that start from their start xyz run together,
        a candidate that continues from the previous one's final xyz waits for
        it.  Writes fit_out.txt lines and renamed parameter copies in order."""
        fs = self.fs
        pts = []
        for params, from_init in cands:
            self.sim_num += 1
            pts.append(Point(self.sim_num, np.array(params, dtype=float), bool(from_init)))
        groups, cur = [], []
        for pt in pts:
            if not pt.from_init and cur:
                groups.append(cur)
                cur = []
            cur.append(pt)
        groups.append(cur)
        for grp in groups:
            runs = [(pt, i) for pt in grp for i in (0, 1)]
            done = {}
            for k in range(0, len(runs), max(1, self.jobs)):
                batch = [(pt, i, *self._launch(pt, i)) for pt, i in runs[k:k + self.jobs]]
                for pt, i, proc, log, d in batch:
                    rc = proc.wait()
                    log.close()
                    if rc != 0:
                        with open(os.path.join(d, "stdout.txt")) as f:
                            tail = f.read()[-3000:]
                        raise RuntimeError(f"md() of evaluation {pt.sim}, system {i + 1} failed ({rc}):\n{tail}")
                    done[(pt.sim, i)] = d
            for pt in grp:
                self._write_point(pt, [done[(pt.sim, 0)], done[(pt.sim, 1)]])
        self.points += pts
        return [pt.error for pt in pts]

    def calc_error(self, params, from_init):
        return self.evaluate([(params, from_init)])[0]

    def _write_point(self, pt: Point, dirs):
        fs = self.fs
        p = pt.params
        self.out.write(f"\n {os.path.join(fs.input_path, fs.param_file)}\n")
        vals = [p[2], p[0], p[1]] if fs.interaction == "ljc" else [p[2], p[0], p[3], p[1]]
        self.out.write(f" {pt.sim} " + " ".join(repr(float(x)) for x in vals) + "\n")
        line = f"{pt.sim:6d}" + "".join(_f21(x) for x in vals)
        be, bd, grd = [], [], []
        for i in (0, 1):
            b, d, g = self._finish(pt, i, dirs[i])
            be.append(b)
            bd.append(d)
            grd.append(g)
            line += _f21(b) + _f21(d) + _f21(g)
            if i == 0:
                t1, t2 = (b / fs.be0 - 1.0) ** 2, (g / fs.grd0[0] - 1.0) ** 2
                error = t1 + t2
                line += _f21(t1) + _f21(t2)
            else:
                t3 = (g / fs.grd0[1] - 1.0) ** 2
                error = error + t3
                line += _f21(t3)
        self.oid.write(line + _f21(error) + "\n")
        self.oid.flush()
        with open(os.path.join(fs.input_path, f"{pt.sim:06d}{fs.param_file}"), "w") as f:
            f.write(param_file_text(fs.interaction, p, fs.rcut))
        pt.error, pt.be, pt.bd, pt.grd, pt.terms = error, tuple(be), tuple(bd), tuple(grd), (t1, t2, t3)
        for d in dirs:
            shutil.rmtree(d, ignore_errors=True)

    # -- the search -------------------------------------------------------------
    def fit(self, delta_error_gold, delta_error_fit, max_sweeps: Optional[int] = None):
        """The coordinate-wise golden-section loop of run_gr_moire_fitting.f90:47-101:
        per sweep, for each of the first three parameters, the bracket ends and
        both golden points from the start xyz, then up to 21 golden steps, each
        continuing from the previous evaluation's final xyz, while the spread of
        the four errors exceeds delta_error_gold; sweeps repeat until the error
        changes by at most delta_error_fit over one coordinate's search."""
        fs = self.fs
        lo0, hi0 = fs.min_params.copy(), fs.max_params.copy()
        lo, hi = lo0.copy(), hi0.copy()
        params = (lo0 + hi0) / 2
        error = prev = 0.0
        sweeps = 0
        self.out.write(LINE + "\n")
        while abs(prev - error) > delta_error_fit or self.sim_num == 0:
            for k in range(3):
                prev = 1e6 if self.sim_num == 0 else error
                lo[k], hi[k] = lo0[k], hi0[k]
                cand = []
                for x in (lo[k], hi[k], lo[k] + (hi[k] - lo[k]) * GOLD, hi[k] - (hi[k] - lo[k]) * GOLD):
                    q = params.copy()
                    q[k] = x
                    cand.append((q, True))
                e = self.evaluate(cand)
                arr = [e[0], e[2], e[3], e[1]]
                params[k] = cand[3][0][k]
                de = max(arr) - min(arr)
                self.out.write(f"  delta_error: {de!r}\n")
                gi = 0
                while de > delta_error_gold and gi <= 20:
                    gi += 1
                    if arr[1] < arr[2]:
                        arr[3], arr[2] = arr[2], arr[1]
                        hi[k] = hi[k] - (hi[k] - lo[k]) * GOLD
                        params[k] = lo[k] + (hi[k] - lo[k]) * GOLD
                        arr[1] = self.calc_error(params, False)
                    else:
                        arr[0], arr[1] = arr[1], arr[2]
                        lo[k] = lo[k] + (hi[k] - lo[k]) * GOLD
                        params[k] = hi[k] - (hi[k] - lo[k]) * GOLD
                        arr[2] = self.calc_error(params, False)
                    de = max(arr) - min(arr)
                    self.out.write(f"  delta_error: {de!r}\n")
                error = min(arr)
                # minloc: the first of equal errors
                at = {0: lo[k], 1: lo[k] + (hi[k] - lo[k]) * GOLD, 2: hi[k] - (hi[k] - lo[k]) * GOLD, 3: hi[k]}
                params[k] = at[int(np.argmin(arr))]
                self.out.write(f"  parameters: {' '.join(repr(float(x)) for x in params)}\n")
            sweeps += 1
            if max_sweeps is not None and sweeps >= max_sweeps:
                break
        self.out.write(LINE + "\n")
        return params, error


def main(argv=None):
    """The flags of run_gr_moire_fitting.f90:14-38 (-op -omp_n -fpfn
    -delta_error_gold -delta_error_fit), plus -jobs (md() runs at once) and
    -prec fp64|mixed."""
    ap = argparse.ArgumentParser(prog="python -m pfmds_amd.moire")
    ap.add_argument("-op", "--out_period", type=int, default=1)
    ap.add_argument("-omp_n", "--openmp_threads_num", type=int, default=1)
    ap.add_argument("-fpfn", required=True)
    ap.add_argument("-delta_error_gold", type=float, required=True)
    ap.add_argument("-delta_error_fit", type=float, required=True)
    ap.add_argument("-jobs", type=int, default=8)
    ap.add_argument("-prec", choices=("fp64", "mixed"), default="fp64")
    a = ap.parse_args(argv)
    print(f"out_period: {a.out_period}\nfitting_parameters_file_name: {a.fpfn}\n"
          f"delta_error_gold: {a.delta_error_gold}\ndelta_error_fit: {a.delta_error_fit}")
    fit = Fitter(read_fitting_parameters(a.fpfn), a.out_period, a.prec, a.jobs)
    try:
        fit.fit(a.delta_error_gold, a.delta_error_fit)
    finally:
        fit.close()


if __name__ == "__main__":
    main()
