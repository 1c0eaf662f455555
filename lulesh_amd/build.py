"""Build the native libraries in-tree.

  lulesh_amd/lib/liblulesh_gpu.so    device path + C ABI (include/lulesh_gpu.h)
  lulesh_amd/lib/liblulesh_host.so   C++ host Domain and driver loop
                                     (include/lulesh_host.h)
  lulesh_amd/lib/lulesh2.0-mi355x    command-line driver, reference flags
  lulesh_amd/lib_asan/               (--sanitize) the host library with
                                     AddressSanitizer + UBSan, for tests/sanitize.sh

Every HIP source is compiled for gfx950 only.  The FP64 step kernels and the
context code are built without FMA contraction so they round as the
reference's x86-64 build; the mixed-precision kernels and the FP64_FMA
instantiation allow contraction.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIBDIR = PKG / "lib"
OBJDIR = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
HIP_SOURCES = {
    "lgpu_api.hip": ["-ffp-contract=off"],
    "lgpu_step_fp64.hip": ["-ffp-contract=off"],
    "lgpu_step_mixed.hip": ["-ffp-contract=fast"],
    "lgpu_step_fp64fma.hip": ["-ffp-contract=fast"],
    "lgpu_ensemble.hip": ["-ffp-contract=off"],
    "lgpu_setup.hip": ["-ffp-contract=off"],
    "lgpu_noalloc.hip": ["-ffp-contract=off"],
}
HOST_SOURCES = ["host/domain.cc", "host/glibc_rand.cc", "host/viz.cc"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(target: Path, deps: list[Path]) -> bool:
    if not target.exists():
        return True
    t = target.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def _headers() -> list[Path]:
    return sorted(CSRC.glob("*.h")) + sorted(CSRC.glob("*.inc")) + sorted(INCLUDE.glob("*.h")) + \
        sorted((CSRC / "host").glob("*.h"))


def build(force: bool = True, verbose: bool = True) -> None:
    LIBDIR.mkdir(exist_ok=True)
    OBJDIR.mkdir(exist_ok=True)
    hdrs = _headers()

    objs = []
    for src, extra in HIP_SOURCES.items():
        o = OBJDIR / (Path(src).stem + ".o")
        if force or _stale(o, [CSRC / src] + hdrs):
            _run([HIPCC, *HIP_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(o)
    gpu_so = LIBDIR / "liblulesh_gpu.so"
    if force or _stale(gpu_so, objs):
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(gpu_so),
              *map(str, objs)])

    host_objs = []
    for src in HOST_SOURCES:
        o = OBJDIR / (Path(src).stem + ".o")
        if force or _stale(o, [CSRC / src] + hdrs):
            _run([CXX, "-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
                  f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / src), "-o", str(o)])
        host_objs.append(o)
    host_so = LIBDIR / "liblulesh_host.so"
    if force or _stale(host_so, host_objs + [gpu_so]):
        _run([CXX, "-shared", "-fPIC", "-o", str(host_so), *map(str, host_objs),
              f"-L{LIBDIR}", "-llulesh_gpu", "-Wl,-rpath,$ORIGIN"])
    exe = LIBDIR / "lulesh2.0-mi355x"
    main_src = CSRC / "host" / "main.cc"
    if force or _stale(exe, [main_src, host_so] + hdrs):
        _run([CXX, "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", f"-I{INCLUDE}",
              f"-I{CSRC}", "-o", str(exe), str(main_src), f"-L{LIBDIR}", "-llulesh_host",
              "-llulesh_gpu", "-Wl,-rpath,$ORIGIN"])


SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined"]


def build_sanitized() -> Path:
    """The host library (C++ Domain, set-up, VTK writer) with AddressSanitizer
    and UndefinedBehaviorSanitizer, in lulesh_amd/lib_asan next to a link to
    the ordinary device library (the HIP code is not instrumented).  Run the
    CPU tests against it with tests/sanitize.sh."""
    out = PKG / "lib_asan"
    out.mkdir(exist_ok=True)
    gpu = out / "liblulesh_gpu.so"
    if not (LIBDIR / "liblulesh_gpu.so").exists():
        build(force=False)
    if gpu.is_symlink() or gpu.exists():
        gpu.unlink()
    gpu.symlink_to(LIBDIR / "liblulesh_gpu.so")
    objs = []
    for src in HOST_SOURCES:
        o = OBJDIR / (Path(src).stem + ".asan.o")
        _run([CXX, "-std=c++17", *SAN_FLAGS, "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
              f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / src), "-o", str(o)])
        objs.append(str(o))
    _run([CXX, "-shared", "-fPIC", *SAN_FLAGS, "-o", str(out / "liblulesh_host.so"), *objs,
          f"-L{out}", "-llulesh_gpu", "-Wl,-rpath,$ORIGIN"])
    return out


def clean() -> None:
    shutil.rmtree(OBJDIR, ignore_errors=True)
    shutil.rmtree(LIBDIR, ignore_errors=True)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    if "--sanitize" in sys.argv:
        OBJDIR.mkdir(exist_ok=True)
        build_sanitized()
    else:
        build(force="--incremental" not in sys.argv)
