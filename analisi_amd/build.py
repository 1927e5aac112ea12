"""Builds the in-tree native libraries of analisi_amd.

  analisi_amd/libanalisi_hip.so   HIP kernels + C ABI (include/analisi_hip.h),
                                  gfx950 code objects, hipcc
  analisi_amd/_pyanalisi*.so      pybind11 module: the pyanalisi classes over
                                  the C++ operator classes (csrc/host/*.h),
                                  linked against libanalisi_hip.so (rpath $ORIGIN)

hipcc cross-compiles gfx950 without a GPU, so this runs in any container with
ROCm.  Usage: python -m analisi_amd.build
"""
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HIP_SOURCES = ["runtime.hip", "gofrt.hip", "msd.hip", "msd_noalloc.hip", "gk.hip", "gk_mfma.hip", "blocks.hip",
               "sph.hip"]
LIB = os.path.join(PKG, "libanalisi_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no v_pk_*_f32 (they are not used by these kernels, and
# some forms have given wrong results on these machines).
# -ffp-contract=off: FP64 paths follow the reference's unfused x86-64 order.
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# per-file extra flags: the g(r,t) kernels take MFMA results in VGPRs (the
# cluster screen compares them right away; AGPR results would need a copy)
FILE_FLAGS = {"gofrt.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def build_hip():
    objs = []
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + FILE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)


def build_pybind():
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "_pyanalisi" + ext)
    inc = [pybind11.get_include(), sysconfig.get_paths()["include"], os.path.join(ROOT, "include"),
           os.path.join(CSRC, "host")]
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           os.path.join(CSRC, "pyanalisi.cpp"), "-o", out + ".tmp"]
    cmd += ["-I" + i for i in inc]
    cmd += ["-L" + PKG, "-l:libanalisi_hip.so", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(out + ".tmp", out)


def build():
    build_hip()
    build_pybind()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "hip" in what:
        build_hip()
    elif "pybind" in what:
        build_pybind()
    else:
        build()
