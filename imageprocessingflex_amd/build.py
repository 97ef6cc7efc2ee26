"""Builds libipfg.so (the gfx950 HIP kernels + C ABI) in-tree with hipcc.

    python -m imageprocessingflex_amd.build        # or build() from Python

Every source under csrc/ is compiled to an object for gfx950 only, then
linked into imageprocessingflex_amd/libipfg.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libipfg.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: no implicit fused multiply-add anywhere (the exact paths
# depend on separate multiply and add; fast paths call fmaf explicitly).
# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def sources() -> list[Path]:
    return sorted(p for p in CSRC.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")
    return obj


_INCLUDE = re.compile(r'^\s*#\s*include\s*"([^"]+)"', re.M)


def _deps(src: Path, seen: set[Path] | None = None) -> set[Path]:
    """The source and every header it includes with quotes, recursively."""
    seen = set() if seen is None else seen
    if src in seen or not src.exists():
        return seen
    seen.add(src)
    for name in _INCLUDE.findall(src.read_text(errors="replace")):
        _deps((src.parent / name).resolve(), seen)
    return seen


def _stale(src: Path) -> bool:
    obj = BUILD / (src.stem + ".o")
    if not obj.exists():
        return True
    newest = max(p.stat().st_mtime for p in _deps(src.resolve()))
    return obj.stat().st_mtime < newest


def build(jobs: int | None = None, verbose: bool = True, incremental: bool = False) -> Path:
    """Compiles every csrc/*.hip for gfx950 and links libipfg.so. Rebuilds
    everything unless `incremental` (then only sources older than their
    object or than any header are recompiled)."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    todo = [s for s in srcs if not incremental or _stale(s)]
    # largest translation units first: they bound the wall time
    todo.sort(key=lambda p: -p.stat().st_size if "_k" not in p.stem else -10**9)
    jobs = jobs or min(max(1, len(todo)), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [BUILD / (s.stem + ".o") for s in srcs]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs),
           "-Wl,-rpath,/opt/rocm/lib", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


ASAN_EXE = BUILD / "abi_check_asan"
SANITIZE = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]


def build_abi_check_asan() -> Path:
    """tests/cpp/abi_check.cpp linked with the library's objects, the runtime's
    host code compiled under AddressSanitizer and UndefinedBehaviorSanitizer
    (device code is built as usual). Needs the objects of build()."""
    objs = [BUILD / (s.stem + ".o") for s in sources() if s.stem != "runtime"]
    missing = [o for o in objs if not o.exists()]
    if missing:
        build(incremental=True, verbose=False)
    rt = BUILD / "runtime_asan.o"
    host_san = [x for f in SANITIZE for x in ("-Xarch_host", f)]
    cmd = [HIPCC, *[f for f in CFLAGS if f != "-O3"], "-O1", "-g", *host_san, "-c", str(CSRC / "runtime.hip"),
           "-o", str(rt)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for runtime.hip (sanitized):\n{r.stderr}")
    clang = Path(HIPCC).resolve().parent.parent / "llvm" / "bin" / "clang++"
    drv = BUILD / "abi_check.o"
    r = subprocess.run([str(clang), "-std=c++17", "-O1", "-g", *SANITIZE, "-I", str(PKG.parent / "include"), "-c",
                        str(PKG.parent / "tests" / "cpp" / "abi_check.cpp"), "-o", str(drv)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"clang++ failed for abi_check.cpp:\n{r.stderr}")
    r = subprocess.run([HIPCC, f"--offload-arch={ARCH}", *host_san[:4], str(drv), str(rt), *map(str, objs),
                        "-Wl,-rpath,/opt/rocm/lib", "-pthread", "-o", str(ASAN_EXE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed (sanitized ABI check):\n{r.stderr}")
    return ASAN_EXE


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
