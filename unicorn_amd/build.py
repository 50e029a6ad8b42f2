"""Build the native libraries of unicorn_amd in-tree.

Outputs (unicorn_amd/lib/):
  libunicorn_hip.so   HIP kernels + C-ABI (hipcc, gfx950)
  libpmlib.so         the pm:: runtime (g++, links libunicorn_hip.so)
  libunicorn_apps.so  the four benchmark apps and their C-ABI drivers (g++)

Objects are compiled in parallel and cached by content hash under unicorn_amd/lib/obj/; every
call relinks. Run as `python -m unicorn_amd.build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INC = ROOT / "include"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = LIB / "obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("UNICORN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wno-unused-result",
    "-Wno-unused-value",
    # packed FP32 instructions are disabled for all device code of this project
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INC}", f"-I{CSRC / 'hip'}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-unused-function",
             f"-I{INC}", f"-I{CSRC / 'runtime'}", f"-I{CSRC / 'apps'}"]


def _hash(paths, flags):
    h = hashlib.sha256()
    for p in paths:
        h.update(Path(p).read_bytes())
    h.update(" ".join(flags).encode())
    return h.hexdigest()[:16]


def _headers(*dirs):
    out = []
    for d in dirs:
        out += sorted(Path(d).glob("*.h"))
    return out


def _compile(src: Path, compiler: str, flags, deps):
    key = _hash([src] + deps, flags + [compiler])
    obj = OBJ / f"{src.stem}-{key}.o"
    if not obj.exists():
        cmd = [compiler] + flags + ["-c", str(src), "-o", str(obj) + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(str(obj) + ".tmp", obj)
    return obj


def _link(out: Path, linker: str, objs, extra):
    tmp = str(out) + ".tmp"
    cmd = [linker, "-shared", "-o", tmp] + [str(o) for o in objs] + extra
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)


def build_all(verbose: bool = True) -> None:
    LIB.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    hip_srcs = sorted((CSRC / "hip").glob("*.hip"))
    rt_srcs = sorted((CSRC / "runtime").glob("*.cpp"))
    app_srcs = sorted((CSRC / "apps").glob("*.cpp"))
    hip_deps = _headers(INC, CSRC / "hip")
    cxx_deps = _headers(INC, CSRC / "runtime", CSRC / "apps", CSRC / "hip")
    jobs = max(1, min(16, os.cpu_count() or 4))
    with cf.ThreadPoolExecutor(jobs) as ex:
        hip_f = [ex.submit(_compile, s, HIPCC, HIP_FLAGS, hip_deps) for s in hip_srcs]
        rt_f = [ex.submit(_compile, s, CXX, CXX_FLAGS, cxx_deps) for s in rt_srcs]
        app_f = [ex.submit(_compile, s, CXX, CXX_FLAGS, cxx_deps) for s in app_srcs]
        hip_o = [f.result() for f in hip_f]
        rt_o = [f.result() for f in rt_f]
        app_o = [f.result() for f in app_f]
    _link(LIB / "libunicorn_hip.so", HIPCC, hip_o, [f"--offload-arch={ARCH}", "-fPIC", "-Wl,--no-undefined"])
    rpath = ["-Wl,-rpath,$ORIGIN", f"-L{LIB}", "-Wl,--no-undefined"]
    _link(LIB / "libpmlib.so", CXX, rt_o, rpath + ["-lunicorn_hip", "-pthread"])
    if app_o:
        _link(LIB / "libunicorn_apps.so", CXX, app_o, rpath + ["-lpmlib", "-lunicorn_hip", "-pthread"])
    # drop stale objects (test-support objects are named after their own sources)
    keep = {p.name for p in hip_o + rt_o + app_o}
    test_stems = ({p.stem for p in TEST_NATIVE.glob("*.hip")} | {p.stem for p in TEST_NATIVE.glob("*.cpp")}
                  if TEST_NATIVE.exists() else set())
    for p in OBJ.glob("*.o"):
        if p.name not in keep and p.name.rsplit("-", 1)[0] not in test_stems:
            p.unlink()
    if verbose:
        print(f"built {LIB}/libunicorn_hip.so libpmlib.so libunicorn_apps.so", flush=True)


TEST_NATIVE = ROOT / "tests" / "native"


def build_test_support(verbose: bool = True) -> None:
    """Test-support libraries of tests/native (sources written against the pm:: API: HIP ones such
    as the kernel-kind contract probe, C++ ones such as the task-ordering probe), one shared
    library per source: tests/native/lib<stem>.so."""
    srcs = sorted(TEST_NATIVE.glob("*.hip")) + sorted(TEST_NATIVE.glob("*.cpp"))
    if not srcs:
        return
    OBJ.mkdir(parents=True, exist_ok=True)
    deps = _headers(INC)
    rpath = ["-Wl,-rpath,$ORIGIN/../../unicorn_amd/lib", f"-L{LIB}", "-Wl,--no-undefined"]
    for src in srcs:
        if src.suffix == ".hip":
            obj = _compile(src, HIPCC, HIP_FLAGS, deps)
            _link(TEST_NATIVE / f"lib{src.stem}.so", HIPCC, [obj],
                  [f"--offload-arch={ARCH}", "-fPIC"] + rpath + ["-lpmlib", "-lunicorn_hip"])
        else:
            obj = _compile(src, CXX, CXX_FLAGS, deps)
            _link(TEST_NATIVE / f"lib{src.stem}.so", CXX, [obj], rpath + ["-lpmlib", "-lunicorn_hip", "-pthread"])
    if verbose:
        print(f"built test support in {TEST_NATIVE}", flush=True)


def clean() -> None:
    shutil.rmtree(LIB, ignore_errors=True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "clean":
        clean()
    else:
        build_all()
        build_test_support()
