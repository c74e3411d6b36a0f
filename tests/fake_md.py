#!/usr/bin/env python3
"""Stand-in for pfmds_amd/bin/run_md_simulation in the CPU tests of the
fitting driver.  It takes the same -ilist/-ipath/-opath/-ofile flags, reads
the settings, start xyz and the candidate parameter file as md() would, and
writes md()'s final xyz and all_out line.  Its "relaxation" is synthetic:
with the candidate (eps, sig, delt) of an ljc parameter file, every carbon
ends at height top + 3 + amplitude * (+-1/2, alternating), the amplitude of
system 1 being G1 * (1 + sig - SIG), of system 2 G2 * (1 + delt - DELT), and
the total energy zero + c_num * BE0 * (1 + eps - EPS) + mean start carbon
height * 1e-7 (so the start xyz matters).  The fitting error is then
(eps - EPS)^2 + (sig - SIG)^2 + (delt - DELT)^2 up to the start term.
"""
import os
import sys

EPS, SIG, DELT = 0.0123, 3.31, 0.42
BE0, G1, G2 = -0.05, 0.3, 0.2
ZERO = (-100.0, -200.0)
C_NUM = (8, 6)


def main():
    a = sys.argv[1:]
    opt = {a[i]: a[i + 1] for i in range(0, len(a) - 1, 2)}
    ipath, opath = opt["-ipath"], opt["-opath"]
    with open(os.path.join(ipath, opt["-ilist"])) as f:
        f.readline()
        settings, prefix = f.readline().split()[:2]
    system = 0 if "sys1" in settings else 1
    with open(os.path.join(ipath, settings)) as f:
        sl = f.read().splitlines()
    xyz = sl[2].split()[1]
    pfile = [ln.split()[1] for ln in sl if ln.startswith("ljc")][0]
    with open(os.path.join(ipath, pfile)) as f:
        eps, sig, delt = [float(x) for x in f.readline().split()[:3]]
    start = os.path.realpath(os.path.join(ipath, xyz))
    print(f"start: {start}")
    with open(start) as f:
        n = int(f.readline().split()[0])
        head = f.readline()
        rows = [f.readline().split() for _ in range(n)]
    metal = [float(r[2]) for r in rows if r[7] in ("CU", "CU_fixed")]
    carbon = [float(r[2]) for r in rows if r[7] == "C"]
    top = max(metal)
    amp = G1 * (1 + sig - SIG) if system == 0 else G2 * (1 + delt - DELT)
    k = 0
    for r in rows:
        if r[7] == "C":
            r[2] = repr(top + 3.0 + amp * (0.5 if k % 2 == 0 else -0.5))
            k += 1
    with open(os.path.join(opath, prefix + "final_" + xyz), "w") as f:
        f.write(f"{n}\n{head}")
        for r in rows:
            f.write(" ".join(r) + "\n")
    total = ZERO[system] + C_NUM[system] * BE0 * (1 + eps - EPS) + 1e-7 * sum(carbon) / len(carbon)
    with open(os.path.join(opath, opt["-ofile"]), "w") as f:
        f.write(f"{prefix[:32]:<32}{7:9d}{3.5:20.9f}{total:20.9f}{total:20.9f}{0.0:20.9f}{0.0:20.9f}\n")


if __name__ == "__main__":
    main()
