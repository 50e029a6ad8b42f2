"""Source compatibility of the app-facing boundary: the reference's own testSuite sources compile
unchanged against include/ (g++ -fsyntax-only; nothing is built or run, and nothing of the
reference is copied). Runs only where the reference tree is present (this container).

- imageFiltering.cpp, pageRank.cpp, matrixMultiply.cpp (the plain kernel-kind app) and
  matrixTranspose.cpp compile as they are.
- The harness testSuite/common/source/code/src/commonAPI.cpp names the CUDA kernel-kind callback
  twice (the type pmSubtaskCallback_GPU_CUDA at :61 and the member subtask_gpu_cuda at :91); it
  compiles with exactly those two names mapped to their HIP counterparts and fails without them.
- matrixMultiplyBlas.cpp and luDecomposition.cpp include <cblas.h>, which this image lacks; the
  other reference apps need the reference's own build (prefixSum.cpp does not compile against
  the reference's own headers either). They are not checked here.
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
TS = REF / "testSuite"

pytestmark = pytest.mark.skipif(not (TS / "common").exists(), reason="reference sources not present")

RENAMES = ["-Dsubtask_gpu_cuda=subtask_gpu_hip", "-DpmSubtaskCallback_GPU_CUDA=pmSubtaskCallback_GPU_HIP"]


def _compile(src: Path, extra=()):
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-w", f"-I{ROOT / 'include'}",
           f"-I{TS / 'common' / 'source' / 'code' / 'inc'}", f"-I{src.parents[1] / 'inc'}", *extra, str(src)]
    return subprocess.run(cmd, capture_output=True, text=True)


@pytest.mark.parametrize("app", ["imageFiltering", "pageRank", "matrixMultiply", "matrixTranspose"])
def test_reference_app_compiles_unchanged(app):
    r = _compile(TS / app / "source" / "code" / "src" / f"{app}.cpp")
    assert r.returncode == 0, r.stderr[-2000:]


def test_reference_harness_binds_with_the_hip_names():
    src = TS / "common" / "source" / "code" / "src" / "commonAPI.cpp"
    r = _compile(src, ["-DSAMPLE_NAME=probe", *RENAMES])
    assert r.returncode == 0, r.stderr[-2000:]
    bare = _compile(src, ["-DSAMPLE_NAME=probe"])
    assert bare.returncode != 0
    errors = [l for l in bare.stderr.splitlines() if "error" in l]
    assert errors and all("subtask_gpu_cuda" in l or "pmSubtaskCallback_GPU_CUDA" in l for l in errors), errors
