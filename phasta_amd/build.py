"""In-tree build of the native libraries.

* ``phasta_amd/libphmesh.so``  host-only box-mesh generator (g++/OpenMP)
  and the phastaIO file format (include/phastaio.h)
* ``phasta_amd/libphgpu.so``   the HIP library behind ``include/phgpu.h``,
  gfx950 only, packed-FP32 instructions disabled
* ``phasta_amd/fortran/phg_boxdrv``  Fortran host (amdflang, iso_c_binding
  module ``phgpu``) that drives the library in the shape of the
  reference's time loop
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "phasta_amd")
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")

HIP_SOURCES = ["phg_api.hip", "k_assembly.hip", "k_solver.hip", "k_mpnn.hip"]
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-fopenmp", f"-I{INC}"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_mesh():
    _run(["g++", "-O2", "-fopenmp", "-fPIC", "-shared", f"-I{INC}",
          os.path.join(CSRC, "boxmesh.cpp"), os.path.join(CSRC, "phastaio.cpp"),
          "-o", os.path.join(PKG, "libphmesh.so")])


def build_gpu():
    """Each source compiled to an object in parallel (no device code
    crosses files: every kernel is launched from its own file), then one
    link."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    odir = os.path.join(PKG, "_build")
    os.makedirs(odir, exist_ok=True)
    objs, procs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(odir, src.replace(".hip", ".o"))
        cmd = [hipcc] + [f for f in HIPCC_FLAGS if f != "-shared"] + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
    if any(p.wait() != 0 for p in procs):
        raise RuntimeError("hipcc failed")
    out = os.path.join(PKG, "libphgpu.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fopenmp"] + objs + ["-o", tmp])
    os.replace(tmp, out)


def build_fortran():
    fdir = os.path.join(PKG, "fortran")
    mdir = os.path.join(fdir, "_build")
    os.makedirs(mdir, exist_ok=True)
    flang = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    out = os.path.join(fdir, "phg_boxdrv")
    _run([flang, "-O2", "-ffp-contract=off", "-module-dir", mdir,
          os.path.join(fdir, "phgpu_mod.f90"), os.path.join(fdir, "phg_boxdrv.f90"),
          f"-L{PKG}", "-lphgpu", "-lphmesh", "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)


def build_all():
    build_mesh()
    build_gpu()
    build_fortran()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "mesh": build_mesh, "gpu": build_gpu,
         "fortran": build_fortran}[w]()
