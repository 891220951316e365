"""Build recipe of the native libraries (runs on a CPU host; hipcc cross-compiles gfx950).

    python -m fases_amd.build            # product library + oracle (+ reference checker)

Products
    fases_amd/libfases_core.so           HIP kernels + host C++ (formats, geometry, solver)
    fases_amd/fases3D                    CLI drop-in for the reference solver executable
    oracle/_build/libfases_oracle.so     FP64 CPU oracle (test infrastructure)
    oracle/_ref/libref_geom.so           reference geometry compiled from its own sources
                                         (only where the reference checkout exists)
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "fases_amd"
CSRC = PKG / "csrc"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FASES_ARCH", "gfx950")

# Host C++ is compiled by g++ without FMA contraction so the geometry classification
# rounds exactly like the reference's g++ host build.
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result", "-Wno-cuda-compat"]

HOST_SOURCES = ["host_formats.cpp", "host_geometry.cpp", "host_ga.cpp"]
HIP_SOURCES = ["fvm_context.hip", "fvm_batch.hip", "fvm_adjoint.hip", "fvm_geometry.hip", "fvm_ga.hip"]
# The geometry and geometrical-acoustics kernels reproduce the host's float decisions bit
# for bit: no FMA contraction.
HIP_EXTRA = {"fvm_geometry.hip": ["-ffp-contract=off"], "fvm_ga.hip": ["-ffp-contract=off"]}


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_core() -> Path:
    OBJ.mkdir(exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        o = OBJ / (Path(src).stem + ".o")
        _run(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(str(o))
    for src in HIP_SOURCES:
        o = OBJ / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, *HIP_EXTRA.get(src, []), "-c", str(CSRC / src), "-o", str(o)])
        objs.append(str(o))
    out = PKG / "libfases_core.so"
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", str(tmp), "-fopenmp"])
    os.replace(tmp, out)
    return out


def build_cli() -> Path:
    out = PKG / "fases3D"
    _run(["g++", "-std=c++17", "-O2", "-Wall", str(CSRC / "fases3d_main.cpp"), "-o", str(out),
          "-L" + str(PKG), "-lfases_core", "-Wl,-rpath,$ORIGIN"])
    return out


def build_oracle() -> Path:
    d = ROOT / "oracle" / "_build"
    d.mkdir(exist_ok=True)
    out = d / "libfases_oracle.so"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          str(ROOT / "oracle" / "fvm_oracle.cpp"), "-o", str(out)])
    return out


def build_reference_checker() -> None:
    _run(["bash", str(ROOT / "oracle" / "build_ref.sh")])


def build_all() -> None:
    build_core()
    build_cli()
    build_oracle()
    build_reference_checker()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
