"""Build the HIP shared library ``bldfm_amd/_bldfm_hip.so`` for gfx950.

Run ``python -m bldfm_amd.build``.  Compiles every ``csrc/*.hip`` with hipcc
(packed-FP32 instructions disabled) and links them into one C-ABI shared
object that ``bldfm_amd._lib`` loads with ctypes.  The library links the HIP
runtime by soname (libamdhip64.so.7), so inside a process that imported torch
it binds to the runtime torch already loaded.
"""

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
LIB = PKG / "_bldfm_hip.so"
ARCH = os.environ.get("BLDFM_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"  # host pass does not know the (device) feature
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
    if err.strip():
        print(err, file=sys.stderr)


def build(verbose: bool = True) -> Path:
    """Compile every csrc/*.hip (in parallel: they are independent units) and link."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = PKG / "_obj"
    objdir.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [str(objdir / (src.stem + ".o")) for src in srcs]
    if verbose:
        print(f"[bldfm build] hipcc {' '.join(s.name for s in srcs)}", flush=True)
    with ThreadPoolExecutor(max_workers=len(srcs)) as pool:
        for f in [pool.submit(_run, [hipcc(), *HIPCC_FLAGS, "-c", str(s), "-o", o]) for s, o in zip(srcs, objs)]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"[bldfm build] wrote {LIB}", flush=True)
    return LIB


if __name__ == "__main__":
    build()
