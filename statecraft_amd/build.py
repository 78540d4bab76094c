"""Build recipe for libstatecraft_hip.so (gfx950) -- plain hipcc, in-tree.

Usage: python -m statecraft_amd.build [--jobs N]

Every .hip / .cpp file under csrc/ is compiled separately for gfx950 and the
objects are linked into statecraft_amd/lib/libstatecraft_hip.so.  Packed FP32
instructions are disabled for all device code.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
OBJDIR = PKG / "lib" / "obj"
LIB = LIBDIR / "libstatecraft_hip.so"

ARCH = os.environ.get("STATECRAFT_ARCH", "gfx950")
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libstatecraft_hip.so")


# Per-source extra flags.
EXTRA: dict = {
    # the exact sequential sums round every product and sum on its own
    "sc_seqsum.hip": ["-ffp-contract=off"],
}


def _compile(src: Path) -> Path:
    obj = OBJDIR / (src.name + ".o")
    cmd = [hipcc(), *COMMON, *EXTRA.get(src.name, []), "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [hipcc(), *COMMON, "-x", "hip", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def _embed_ops_header() -> Path:
    """The runtime-compiled pass kernels include sc_ops.hpp; its text is compiled into the library."""
    text = (CSRC / "sc_ops.hpp").read_text()
    delim = "SCOPS"
    if f"){delim}\"" in text:
        raise RuntimeError("sc_ops.hpp contains the raw-string delimiter")
    src = OBJDIR / "sc_ops_text.cpp"
    body = f'namespace sc {{\nextern const char kOpsHeaderText[] = R"{delim}({text}){delim}";\n}}\n'
    if not src.exists() or src.read_text() != body:
        src.write_text(body)
    obj = OBJDIR / "sc_ops_text.o"
    cmd = [os.environ.get("CXX", "g++"), "-O1", "-fPIC", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    OBJDIR.mkdir(parents=True, exist_ok=True)
    sources = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    objs.append(_embed_ops_header())
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lhiprtc"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_engine()
    if verbose:
        print(f"built {LIB} from {len(sources)} sources, {ENGINE_LIB.name} and {ENGINE_TEST.name}", flush=True)
    return LIB


ENGINE_SRC = PKG / "engine" / "statecraft.cpp"
ENGINE_LIB = LIBDIR / "libstatecraft_engine.so"
ENGINE_TEST_SRC = PKG.parent / "tests" / "cpp" / "engine_test.cpp"
ENGINE_TEST = LIBDIR / "engine_test"


def build_engine() -> None:
    """The C++ engine mirror (host code only) over the C-ABI, and its test program."""
    cxx = os.environ.get("CXX", "g++")
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", f"-I{PKG.parent}"]
    rpath = "-Wl,-rpath,$ORIGIN"
    cmd = [cxx, *flags, "-shared", "-o", str(ENGINE_LIB), str(ENGINE_SRC), f"-L{LIBDIR}", "-lstatecraft_hip", rpath]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"engine build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    cmd = [cxx, *flags, "-o", str(ENGINE_TEST), str(ENGINE_TEST_SRC), f"-L{LIBDIR}", "-lstatecraft_engine",
           "-lstatecraft_hip", rpath]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"engine test build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args(argv)
    build(args.jobs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
