"""Build libaxs.so (HIP for gfx950 + C++ host) and the CPU oracle libraries.

Used by __graft_entry__.build().  Everything is built in-tree:

  axisim-software_amd/lib/libaxs.so   the product (hipcc, --offload-arch=gfx950)
  oracle/liboracle.so                 the FP64 CPU restatement (g++), test infrastructure
  oracle/_ref/                        reference-compiled checkers (oracle/Makefile), only
                                      where the reference sources are present
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libaxs.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
# FP64 as written (no contraction into fused multiply-adds), OpenMP loops
ORACLE_FLAGS = ["-std=c++17", "-O3", "-ffp-contract=off", "-fopenmp"]

SOURCES = [
    "csrc/hip/kernels.hip",
    "csrc/hip/kernels_mixed.hip",
    "csrc/hip/mpnn.hip",
    "csrc/hip/noalloc.hip",
    "csrc/hip/quality.hip",
    "csrc/host/solver.cpp",
    "csrc/host/net.cpp",
    "csrc/host/capi.cpp",
    "csrc/host/topology.cpp",
    "csrc/host/casegen.cpp",
    "csrc/host/export.cpp",
    "csrc/host/cdt.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # keep every multiply-add rounded as written (FP64 parity with the oracle)
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{PKG / 'csrc' / 'host'}",
]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in res.stdout.splitlines() if "packed-fp32-ops" not in line and line.strip()
    )
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{out}")


def build_product(jobs: int = 8) -> Path:
    """Compile every source with hipcc for gfx950 and link libaxs.so."""
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    srcs = [s for s in SOURCES if (PKG / s).exists()]

    def compile_one(src: str) -> Path:
        obj = obj_dir / (Path(src).name + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", str(PKG / src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_compat() -> Path:
    """The C++ drop-ins (compat/: the Solver class and the AxiMesh mesher
    interface) as a static archive plus the demo that drives the solver like
    the GUI does; plain g++ against include/axs.h + libaxs.so."""
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    inc = [f"-I{ROOT / 'include'}", f"-I{PKG / 'compat'}"]
    obj = obj_dir / "compat_solver.o"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", *inc, "-c", str(PKG / "compat" / "solver.cpp"), "-o", str(obj)])
    mesher = obj_dir / "compat_aximesh.o"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-Wall", *inc, "-c", str(PKG / "compat" / "aximesh.cpp"), "-o",
          str(mesher)])
    archive = LIB_DIR / "libaxs_compat.a"
    if archive.exists():
        archive.unlink()
    _run(["ar", "rcs", str(archive), str(obj), str(mesher)])
    # the demo plays the GUI: its console and residual plot are the stand-ins
    # under tests/compat_stub, and it compiles the GUI-side hooks against them
    demo = LIB_DIR / "compat_demo"
    stub = [f"-I{ROOT / 'tests' / 'compat_stub'}"]
    _run(["g++", "-std=c++17", "-O2", "-Wall", *inc, *stub, str(PKG / "compat" / "compat_demo.cpp"),
          str(PKG / "compat" / "gui_hooks.cpp"), str(archive), f"-L{LIB_DIR}", "-laxs", f"-Wl,-rpath,$ORIGIN",
          "-pthread", "-o", str(demo)])
    return demo


def build_oracle() -> Path:
    """g++ build of the CPU oracle (FP64, no contraction; OpenMP for the
    per-cell loops of the multi-thread CPU baseline)."""
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([
        "g++", *ORACLE_FLAGS, "-fPIC", "-shared",
        f"-I{ROOT / 'include'}", str(ORACLE_DIR / "oracle.cpp"), "-o", str(tmp),
    ])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build_reference_checkers() -> bool:
    """oracle/_ref: compile the reference's header-only multiblock packer into a
    golden-vector writer and its mesher into a ctypes library.  Only where the
    reference sources are present."""
    if not Path("/root/reference/structs/multiblock.h").exists():
        return False
    if shutil.which("make") is None:
        return False
    _run(["make", "-C", str(ORACLE_DIR), "ref"])
    # the reference's mesher harness, once on the reference and once on the
    # product's drop-in (needs libaxs.so and libaxs_compat.a built)
    _run(["make", "-C", str(ORACLE_DIR), "harness"])
    return True


def build_all(jobs: int = 8) -> None:
    build_product(jobs)
    build_compat()
    build_oracle()
    build_reference_checkers()


if __name__ == "__main__":
    build_all(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
    print(f"built {LIB}")
