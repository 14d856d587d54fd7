"""Builds the CPU kernel-check harnesses under tests/cpu_kernels/_build/.

The library's ed_batch.hip (with ed_internal.h and include/cmpy_amd.h) is
copied with its LDS address-space qualifiers removed (the host has one
address space) and compiled by the host clang++ against the device emulation
in hip/hip_runtime.h, with AddressSanitizer and UBSan.  Nothing else in the
copy changes.  ``build()`` returns the executable's path and rebuilds only
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
SOURCES = {
    "cmpy_amd/csrc/ed_batch.cpp": REPO / "cmpy_amd/csrc/ed_batch.hip",
    "cmpy_amd/csrc/ed_internal.h": REPO / "cmpy_amd/csrc/ed_internal.h",
    "include/cmpy_amd.h": REPO / "include/cmpy_amd.h",
}
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined", "-Wno-unused-function", "-Wno-unused-variable"]


def _digest(mutate) -> str:
    h = hashlib.sha256()
    for p in list(SOURCES.values()) + [HERE / "hip/hip_runtime.h", HERE / "batch_check.cpp", Path(__file__)]:
        h.update(p.read_bytes())
    h.update(" ".join([CLANG] + FLAGS).encode())
    h.update(repr(mutate).encode())
    return h.hexdigest()


def build(mutate=None, name: str = "batch_check") -> Path:
    """Builds tests/cpu_kernels/_build/<name>.  ``mutate`` = (old, new): a
    text replacement applied to the copy of ed_batch.hip (exactly one match),
    for planting a known mistake."""
    exe = BUILD / name
    stamp = BUILD / f"{name}.sha256"
    dig = _digest(mutate)
    if exe.exists() and stamp.exists() and stamp.read_text() == dig:
        return exe
    src_dir = BUILD / f"{name}_src"
    for rel, src in SOURCES.items():
        dst = src_dir / rel
        dst.parent.mkdir(parents=True, exist_ok=True)
        text = _AS.sub("", src.read_text())
        if mutate is not None and rel.endswith("ed_batch.cpp"):
            if text.count(mutate[0]) != 1:
                raise ValueError(f"mutation must match once, matched {text.count(mutate[0])} times")
            text = text.replace(mutate[0], mutate[1])
        dst.write_text(text)
    cmd = [CLANG, *FLAGS, f"-I{HERE}", f"-I{src_dir}", str(HERE / "batch_check.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("harness build failed:\n" + r.stderr[-6000:])
    stamp.write_text(dig)
    return exe


if __name__ == "__main__":
    print(build())
