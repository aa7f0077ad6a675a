"""Build recipe for libpyro_cmc.so (HIP, gfx950) and the C++ host tools.

Run ``python -m pyro_cmc_amd.build`` from the repository root.  The library
is written next to this file (in-tree) so it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "libpyro_cmc.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PYRO_OFFLOAD_ARCH", "gfx950")
HDF5_ROOT = Path(os.environ.get("PYRO_HDF5_ROOT", "/opt/conda"))

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]

SOURCES = [
    "kernels.hip",
    "kernels_pyro_colour_f32.hip",
    "kernels_pyro_colour_f64.hip",
    "kernels_pyro_resident_f32.hip",
    "kernels_pyro_resident_f64.hip",
    "kernels_pyro_tiled_f32.hip",
    "kernels_pyro_tiled_f64.hip",
    "kernels_pyro_xcd_f32.hip",
    "kernels_pyro_xcd_f64.hip",
    "kernels_pyro_batch_f32.hip",
    "kernels_pyro_batch_f64.hip",
    "lt.hip",
    "capi.cpp",
    "topology.cpp",
]

# per-source flags: the XCD-resident and batch kernels use the
# threadgroup-split memory model, under which a workgroup-scope release
# waits for the workgroup's stores to complete
EXTRA_FLAGS = {
    "kernels_pyro_xcd_f32.hip": ["-mtgsplit"],
    "kernels_pyro_xcd_f64.hip": ["-mtgsplit"],
    "kernels_pyro_batch_f32.hip": ["-mtgsplit"],
    "kernels_pyro_batch_f64.hip": ["-mtgsplit"],
}

LINK_LIBS: list[str] = []


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_library() -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = BUILD / (src + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        cmds.append([HIPCC, *HIP_FLAGS, *EXTRA_FLAGS.get(src, []), *lang, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), *LINK_LIBS])
    os.replace(tmp, LIB)
    return LIB


def build_host_tools() -> list[Path]:
    """C++ mirror of the reference host API and the anneal/temper CLIs."""
    if not HOST.exists() or not (HDF5_ROOT / "include" / "hdf5.h").exists():
        return []
    out = []
    common = [
        "g++", "-O2", "-std=c++20", f"-I{ROOT / 'include'}", f"-I{HOST}",
        f"-I{HDF5_ROOT / 'include'}",
    ]
    # libhdf5 by full path, never -L on its directory: that directory also
    # holds an older libstdc++ which must not win at link or load time, so
    # the system library directory precedes it in the run path.
    hdf5 = sorted((HDF5_ROOT / "lib").glob("libhdf5.so.*"))[0]
    link = [
        f"-L{PKG}", "-lpyro_cmc", str(hdf5),
        f"-Wl,-rpath,$ORIGIN:/usr/lib/x86_64-linux-gnu:{HDF5_ROOT / 'lib'}",
    ]
    tools = sorted(p for p in HOST.glob("*.cpp") if p.name not in ("host_lib.cpp",))
    lib_src = HOST / "host_lib.cpp"
    extra = [str(lib_src)] if lib_src.exists() else []
    cmds = []
    for src in tools:
        exe = PKG / src.stem
        cmds.append([*common, str(src), *extra, "-o", str(exe), *link])
        out.append(exe)
    with ThreadPoolExecutor(max_workers=max(1, min(8, len(cmds)))) as ex:
        list(ex.map(_run, cmds))
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


def build_all() -> None:
    build_library()
    build_host_tools()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print(f"built {LIB}")
