"""Build libbem3qd2.so (HIP kernels for gfx950 + host C) in-tree.

    python -m bem3_emf_qd2_amd.build        # or: python bem3_emf_qd2_amd/build.py

Objects go to bem3_emf_qd2_amd/_build/, the library to bem3_emf_qd2_amd/lib/.  Every call
compiles from scratch (no timestamps are trusted), in parallel.
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
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libbem3qd2.so"
BINDIR = PKG / "bin"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=fast-honor-pragmas",
    # packed FP32 instructions are kept out of all generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-variable",
           "-Wno-unused-but-set-variable", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
LINK_LIBS = ["-lrocblas", "-lm", "-lpthread"]
# Per-source additions.  The assembly kernels' polynomial and constant-heavy code: with
# machine loop-invariant code motion the compiler hoists every double constant of the lattice
# sums into registers for the whole kernel (spilling scalars to vector lanes and vectors to
# scratch) and copies them into the destinations of v_fmac_f64; without it each constant is
# built next to its use and holds no register across the sums.
HIP_FILE_FLAGS = {"bq_assemble.hip": ["-mllvm", "--disable-machine-licm"]}


def _run(cmd: list[str]) -> tuple[list[str], int, str]:
    p = subprocess.run(cmd, capture_output=True, text=True)
    return cmd, p.returncode, p.stdout + p.stderr


CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-pthread", f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def sources() -> tuple[list[Path], list[Path]]:
    hip = sorted((CSRC / "dev").glob("*.hip"))
    c = sorted((CSRC / "host").glob("*.c")) + sorted((CSRC / "host").glob("*.cpp"))
    return hip, c


def build(verbose: bool = False) -> Path:
    if BUILD.exists():
        shutil.rmtree(BUILD)
    BUILD.mkdir(parents=True)
    LIBDIR.mkdir(exist_ok=True)
    hip, c = sources()
    jobs = []
    for s in hip:
        jobs.append([HIPCC, *HIP_FLAGS, *HIP_FILE_FLAGS.get(s.name, []), "-c", str(s), "-o",
                     str(BUILD / (s.stem + ".hip.o"))])
    for s in c:
        if s.suffix == ".cpp":
            jobs.append(["g++", *CXX_FLAGS, "-c", str(s), "-o", str(BUILD / (s.stem + ".cpp.o"))])
        else:
            jobs.append(["gcc", *C_FLAGS, "-c", str(s), "-o", str(BUILD / (s.stem + ".c.o"))])
    workers = min(8, os.cpu_count() or 4)
    failed = []
    with cf.ThreadPoolExecutor(workers) as ex:
        for cmd, rc, out in ex.map(_run, jobs):
            if rc != 0:
                failed.append((cmd, out))
            elif verbose and out.strip():
                print(out)
    if failed:
        for cmd, out in failed:
            print(" ".join(cmd), file=sys.stderr)
            print(out, file=sys.stderr)
        raise RuntimeError(f"{len(failed)} compile step(s) failed")
    objs = sorted(str(p) for p in BUILD.glob("*.o"))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs,
           f"-L{ROCM}/lib", f"-Wl,-rpath,{ROCM}/lib", *LINK_LIBS]
    _, rc, out = _run(cmd)
    if rc != 0:
        print(out, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    _build_tools()
    return LIB


def _build_tools() -> None:
    """Command-line drop-ins (d3qd2_bv_solver, example1) linked against the library."""
    tools = sorted((CSRC / "tools").glob("*.c"))
    if not tools:
        return
    BINDIR.mkdir(exist_ok=True)
    for t in tools:
        out = BINDIR / t.stem
        cmd = ["gcc", "-O2", "-std=gnu11", f"-I{ROOT / 'include'}", str(t), "-o", str(out),
               f"-L{LIBDIR}", "-lbem3qd2", "-Wl,-rpath,$ORIGIN/../lib", "-lm"]
        _, rc, o = _run(cmd)
        if rc != 0:
            print(o, file=sys.stderr)
            raise RuntimeError(f"tool {t.name} failed to build")


def build_oracle() -> None:
    """CPU checkers under oracle/ (and oracle/_ref when /root/reference is present)."""
    p = subprocess.run(["make", "-C", str(ROOT / "oracle"), "-j8", "all"], capture_output=True, text=True)
    if p.returncode != 0:
        print(p.stdout + p.stderr, file=sys.stderr)
        raise RuntimeError("oracle build failed")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    build_oracle()
