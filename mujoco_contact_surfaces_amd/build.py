"""Builds libhcs.so (the HIP engine, gfx950) in-tree and the CPU oracle.

    python -m mujoco_contact_surfaces_amd.build

Every HIP source is compiled with hipcc for gfx950 without floating-point
contraction (the engine's arithmetic contract) and with packed FP32 ops
disabled; the host generator is compiled with g++.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libhcs.so")

HIP_SOURCES = ["hcs_engine.hip", "hcs_batch.hip", "hcs_taxels.hip", "hcs_taxel_sensor.hip", "hcs_adjoint.hip",
               "hcs_noalloc.hip", "hcs_prim.hip"]
CXX_SOURCES = ["hcs_gen.cpp", "hcs_gen_geom.cpp", "hcs_mujoco.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_engine(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    cmds = []
    objs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s.replace(".cpp", ".o"))
        cmds.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_glue_demo():
    """The compiled driver of the MuJoCo host glue used by the GPU tests
    (tests/glue/glue_demo.cpp), linked against the in-tree libhcs.so."""
    src = os.path.join(ROOT, "tests", "glue", "glue_demo.cpp")
    out = os.path.join(ROOT, "tests", "glue", "hcs_glue_demo")
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), src, "-o", out + ".tmp", "-L" + PKG,
          "-lhcs", "-Wl,-rpath,$ORIGIN/../../mujoco_contact_surfaces_amd"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-B", "liboracle.so"])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def main():
    print(build_engine())
    print(build_glue_demo())
    print(build_oracle())


if __name__ == "__main__":
    sys.exit(main())
