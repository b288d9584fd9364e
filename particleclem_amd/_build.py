"""Build the gfx950 hydro library in-tree (particleclem_amd/libpc_hydro.so).

hipcc cross-compiles for gfx950 without a GPU.  Flags:
  -ffp-contract=off   products and sums round separately, as the reference's
                      expressions read (keeps the kernels reproducible
                      against the CPU restatement of the reference);
  -target-feature -packed-fp32-ops   no packed FP32 instructions.
"""
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libpc_hydro.so")
SOURCES = ["pc_hydro.hip", "pc_plotfile.cpp"]
HEADERS = ["pc_phys.h", "pc_kernels.h", "pc_level_kernels.h", "pc_mol_kernels.h",
           "pc_derive_kernels.h", "pc_plotfile.h", "pc_fluxreg.h", "pc_shell.h",
           "pc_trace_lds.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops"]


def build(verbose=False):
    tmp = LIB + ".tmp%d" % os.getpid()
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    if verbose and r.stderr.strip():
        print(r.stderr[-2000:])
    os.replace(tmp, LIB)
    return LIB


def is_stale():
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    for f in SOURCES + HEADERS:
        if os.path.getmtime(os.path.join(CSRC, f)) > t:
            return True
    inc = os.path.join(os.path.dirname(PKG), "include", "pc_hydro.h")
    return os.path.getmtime(inc) > t
