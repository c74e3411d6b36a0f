"""Build the native parts in-tree (no JIT cache, so the built files travel with
the source tree):

  pfmds_amd/lib/libpfmds.so          HIP library, gfx950 only (hipcc)
  pfmds_amd/bin/run_md_simulation    Fortran front end (amdflang), links libpfmds.so
  oracle/liborc.so                   CPU oracle used by the tests (gcc)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "bin", "amdflang")
HIP_SOURCES = ["pfm_context.hip", "pfm_neigh.hip", "pfm_forces.hip", "pfm_surrogate.hip", "pfm_rebo.hip"]
# packed FP32 instructions are disabled on purpose: the compiler forms them
# from plain float code and some forms are unreliable on this hardware
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip():
    out = os.path.join(HERE, "lib", "libpfmds.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(HERE, "lib", src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(HERE, "csrc", src), "-o", obj])
        objs.append(obj)
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out, *objs])
    return out


def build_fortran():
    src = os.path.join(HERE, "fortran")
    if not os.path.isdir(src):
        return None
    out_dir = os.path.join(HERE, "bin")
    mod_dir = os.path.join(HERE, "lib", "mod")
    os.makedirs(out_dir, exist_ok=True)
    os.makedirs(mod_dir, exist_ok=True)
    files = ["pfm_binding.f90", "pfm_md.f90", "run_md_simulation.f90"]
    files = [os.path.join(src, f) for f in files if os.path.exists(os.path.join(src, f))]
    if not files:
        return None
    out = os.path.join(out_dir, "run_md_simulation")
    _run([FLANG, "-O2", "-fdefault-real-8", "-fdefault-double-8", "-fopenmp", "-module-dir", mod_dir, *files,
          "-L" + os.path.join(HERE, "lib"), "-lpfmds", "-Wl,-rpath,$ORIGIN/../lib", "-o", out])
    return out


def build_oracle():
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liborc.so")


def build_all():
    build_oracle()
    build_hip()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "fortran": build_fortran, "oracle": build_oracle}[w]()
