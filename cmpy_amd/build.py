"""Build libcmpy_amd.so in-tree with hipcc for gfx950.

Usage: ``python -m cmpy_amd.build`` (or ``cmpy_amd.build.build()``).  Each
``csrc/*.hip`` file is compiled to an object under ``csrc/_build/`` and the
objects are linked into ``cmpy_amd/libcmpy_amd.so`` next to this file.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libcmpy_amd.so"
ARCH = os.environ.get("CMPY_AMD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    # no packed FP32 instructions in any device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
LDFLAGS = ["-shared", f"-L{ROCM}/lib", "-lrocsolver", "-lrocblas"]

# Largest private (scratch) size per lane a kernel may compile to.  The
# persistent batch kernels (1024-thread workgroups, one per CU slot) run at
# 240-304 bytes/lane; two builds that spilled to ~1.37 KB/lane (164-166
# spilled VGPRs) faulted on MI355X with a memory-aperture violation, while the
# same code run on the CPU under AddressSanitizer (tests/cpu_kernels) stays in
# bounds.  A build past this limit is refused rather than shipped.
SCRATCH_LIMIT = 512
_REMARK = re.compile(r"remark:\s+(Function Name: (\S+)|ScratchSize \[bytes/lane\]: (\d+))")


def kernel_scratch(remarks: str) -> dict:
    """Per-kernel ScratchSize (bytes/lane) from -Rpass-analysis=kernel-resource-usage output."""
    out, name = {}, None
    for m in _REMARK.finditer(remarks):
        if m.group(2):
            name = m.group(2)
        elif name is not None:
            out[name] = int(m.group(3))
            name = None
    return out


def _compile(src: Path, obj: Path, verbose: bool) -> None:
    cmd = [HIPCC, *CFLAGS, "-Rpass-analysis=kernel-resource-usage", "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = r.stdout + r.stderr
    err = "\n".join(
        line for line in out.splitlines()
        if "packed-fp32-ops" not in line and "kernel-resource-usage" not in line
    )
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{err}")
    over = {k: v for k, v in kernel_scratch(out).items() if v > SCRATCH_LIMIT}
    if over:
        obj.unlink(missing_ok=True)
        raise RuntimeError(f"{src.name}: kernels over the {SCRATCH_LIMIT} bytes/lane scratch limit "
                           f"(register spills): {over}")
    if verbose and err.strip():
        print(err, flush=True)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    bdir = CSRC / "_build"
    bdir.mkdir(exist_ok=True)
    objs = [bdir / (s.stem + ".o") for s in sources]
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, o, verbose) for s, o in zip(sources, objs)]
        for f in futs:
            f.result()
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), *LDFLAGS]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
