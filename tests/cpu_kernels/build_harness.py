"""Builds the CPU kernel-check harnesses under tests/cpu_kernels/_build/.

A library source (ed_batch.hip or ed_core.hip, with ed_internal.h,
ed_colouring.h and include/cmpy_amd.h) is copied with its LDS address-space
qualifiers removed (the host has one address space) and compiled by the host
clang++ against the device emulation in hip/hip_runtime.h, with
AddressSanitizer and UBSan.  Nothing else in the copy changes.  ``build()`` returns the executable's path and rebuilds only
when a source changed.
"""

from __future__ import annotations

import hashlib
import re
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
BUILD = HERE / "_build"
CLANG = shutil.which("clang++", path="/opt/rocm/llvm/bin") or "/opt/rocm/llvm/bin/clang++"
_AS = re.compile(r"__attribute__\(\(address_space\(\d+\)\)\)\s*")
COMMON = {
    "cmpy_amd/csrc/ed_internal.h": REPO / "cmpy_amd/csrc/ed_internal.h",
    "cmpy_amd/csrc/ed_colouring.h": REPO / "cmpy_amd/csrc/ed_colouring.h",
    "include/cmpy_amd.h": REPO / "include/cmpy_amd.h",
}
# harness -> (driver, library source compiled into it)
TARGETS = {
    "batch": (HERE / "batch_check.cpp", "ed_batch"),
    "core": (HERE / "core_check.cpp", "ed_core"),
}
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined", "-Wno-unused-function", "-Wno-unused-variable"]


def _sources(target):
    driver, lib = TARGETS[target]
    return driver, {**COMMON, f"cmpy_amd/csrc/{lib}.cpp": REPO / f"cmpy_amd/csrc/{lib}.hip"}


def _digest(target, mutate) -> str:
    driver, src = _sources(target)
    h = hashlib.sha256()
    for p in list(src.values()) + [HERE / "hip/hip_runtime.h", driver, Path(__file__)]:
        h.update(p.read_bytes())
    h.update(" ".join([CLANG] + FLAGS).encode())
    h.update(repr(mutate).encode())
    return h.hexdigest()


def build(mutate=None, name: str = "batch_check", target: str = "batch") -> Path:
    """Builds tests/cpu_kernels/_build/<name> from harness ``target`` ("batch":
    k_batch_solve / k_batch_backward; "core": operator construction).
    ``mutate`` = (old, new): a text replacement applied to the copy of the
    library source (exactly one match), for planting a known mistake."""
    driver, sources = _sources(target)
    exe = BUILD / name
    stamp = BUILD / f"{name}.sha256"
    dig = _digest(target, mutate)
    if exe.exists() and stamp.exists() and stamp.read_text() == dig:
        return exe
    src_dir = BUILD / f"{name}_src"
    for rel, src in sources.items():
        dst = src_dir / rel
        dst.parent.mkdir(parents=True, exist_ok=True)
        text = _AS.sub("", src.read_text())
        if mutate is not None and rel.endswith(".cpp"):
            if text.count(mutate[0]) != 1:
                raise ValueError(f"mutation must match once, matched {text.count(mutate[0])} times")
            text = text.replace(mutate[0], mutate[1])
        dst.write_text(text)
    cmd = [CLANG, *FLAGS, f"-I{HERE}", f"-I{src_dir}", str(driver), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("harness build failed:\n" + r.stderr[-6000:])
    stamp.write_text(dig)
    return exe


if __name__ == "__main__":
    print(build())
    print(build(name="core_check", target="core"))
