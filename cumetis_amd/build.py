"""Build recipe for the gfx950 engine library and the command-line tool.

``python -m cumetis_amd.build`` compiles every HIP/C++ source under
``cumetis_amd/csrc`` with hipcc into ``cumetis_amd/libcumetis_amd.so`` and
links the ``cuMetis`` executable next to it.  Objects are compiled in parallel
and cached by content hash under ``cumetis_amd/_build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libcumetis_amd.so"
CLI = PKG / "cuMetis"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CM_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          f"-I{ROOT / 'include'}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _sources():
    hip = sorted(CSRC.glob("*.hip"))
    cpp = sorted((CSRC / "host").glob("*.cpp")) if (CSRC / "host").is_dir() else []
    return hip, cpp


def _digest(src: Path, flags) -> str:
    h = hashlib.sha256()
    h.update(" ".join(flags).encode())
    h.update(src.read_bytes())
    for hdr in sorted(list(CSRC.glob("*.hpp")) + list((ROOT / "include").glob("*.h"))
                      + list((CSRC / "host").glob("*.hpp"))):
        h.update(hdr.read_bytes())
    return h.hexdigest()[:16]


def _compile(src: Path, flags, force: bool = False) -> Path:
    obj = BUILD / f"{src.stem}-{_digest(src, flags)}.o"
    if obj.exists() and not force:
        return obj
    cmd = [HIPCC, *flags, "-c", str(src), "-o", str(obj) + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(str(obj) + ".tmp", obj)
    return obj


def build(verbose: bool = False, force: bool = False) -> Path:
    """Compile and link; force recompiles every object (otherwise objects
    whose source, headers and flags are unchanged are reused)."""
    BUILD.mkdir(exist_ok=True)
    hip, cpp = _sources()
    jobs = [(s, COMMON + DEVICE + ["-x", "hip"]) for s in hip]
    jobs += [(s, COMMON + ["-x", "c++"]) for s in cpp if s.name != "cm_cli.cpp"]
    workers = min(8, max(1, len(jobs)))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(lambda j: _compile(*j, force=force), jobs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cli_src = CSRC / "host" / "cm_cli.cpp"
    if cli_src.exists():
        cli_obj = _compile(cli_src, COMMON + ["-x", "c++"], force=force)
        cmd = [HIPCC, "-o", str(CLI), str(cli_obj), f"-L{PKG}", "-lcumetis_amd",
               f"-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    r = subprocess.run(["make", "-s", "-B", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    return ROOT / "oracle" / "liboracle.so"


if __name__ == "__main__":
    build(verbose=True)
    if "--no-oracle" not in sys.argv:
        build_oracle()
