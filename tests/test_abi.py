"""The built libraries export every entry point the public headers declare.

Runs without a GPU: the libraries are loaded (or their dynamic symbol tables read) but nothing
is called.
  - include/unicorn_hip.h  -> extern "C" uh_*  in libunicorn_hip.so
  - include/unicorn_apps.h -> extern "C" ua_*  in libunicorn_apps.so
  - include/pmPublicDefinitions.h, pmPublicUtilities.h, pmExtensions.h -> C++ pm::* in libpmlib.so
    (source-compatible with the reference API, pmPublicDefinitions.h / pmPublicUtilities.h of
    source/code/inc)
"""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
INC = ROOT / "include"
LIB = ROOT / "unicorn_amd" / "lib"

DECL = re.compile(r"^\s*(?:extern\s+\"C\"\s+)?(?:const\s+)?[A-Za-z_][\w:]*[\s\*]+\**(?P<name>{prefix}\w*)\s*\(")


def declared(header: str, prefix: str) -> list[str]:
    rx = re.compile(DECL.pattern.format(prefix=prefix))
    names = []
    for line in (INC / header).read_text().splitlines():
        s = line.strip()
        if s.startswith(("//", "/*", "*", "#", "typedef", "return")) or "inline" in s or "{" in s:
            continue
        m = rx.match(line)
        if m:
            names.append(m.group("name"))
    return sorted(set(names))


def test_headers_parse():
    assert len(declared("unicorn_hip.h", "uh_")) >= 40
    assert len(declared("unicorn_apps.h", "ua_")) >= 30
    assert "pmRegisterCallbacks" in declared("pmPublicDefinitions.h", "pm[A-Z]")


@pytest.mark.parametrize("header,lib,prefix", [
    ("unicorn_hip.h", "libunicorn_hip.so", "uh_"),
    ("unicorn_apps.h", "libunicorn_apps.so", "ua_"),
])
def test_c_abi_exports(header, lib, prefix):
    path = LIB / lib
    assert path.exists(), f"{path} not built"
    handle = ctypes.CDLL(str(path), mode=ctypes.RTLD_GLOBAL)
    missing = [n for n in declared(header, prefix) if not hasattr(handle, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_pm_api_exports():
    path = LIB / "libpmlib.so"
    assert path.exists(), f"{path} not built"
    out = subprocess.run(["nm", "-DC", "--defined-only", str(path)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bpm::(pm[A-Z]\w*)\(", out))
    names = set()
    for h in ("pmPublicDefinitions.h", "pmPublicUtilities.h", "pmExtensions.h"):
        names.update(declared(h, "pm[A-Z]"))
    missing = sorted(names - exported)
    assert not missing, f"libpmlib.so lacks pm::{missing}"
