"""In-tree build of libstntt.so for gfx950 (hipcc, no CMake).

    python -m stntt_amd.build

The device code is split over translation units that compile in parallel:
the host library (stntt.hip), the persistent batch kernel
(batch_kernel.hip) and the pass kernels once per field and direction
(launch_pass.hip with STNTT_TU_FIELD / STNTT_TU_DIR).  Objects go to
stntt_amd/_build/, then one hipcc link makes the shared library.
"""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
OBJ_DIR = os.path.join(PKG_DIR, "_build")
OUT = os.path.join(PKG_DIR, "libstntt.so")

# -packed-fp32-ops off: the library does no float math, and packed fp32 is
# kept out of every code object built here.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

FIELDS = {"gl": "FieldGL", "u32": "FieldU32", "u32l": "FieldU32L", "m64": "FieldM64"}


def units() -> list[tuple[str, str, list[str]]]:
    """(object name, source, extra flags) of every translation unit."""
    out = [("stntt.o", "stntt.hip", []), ("batch_kernel.o", "batch_kernel.hip", [])]
    for tag, field in FIELDS.items():
        for d in (0, 1):
            out.append((f"launch_{tag}_{d}.o", "launch_pass.hip",
                        [f"-DSTNTT_TU_FIELD={field}", f"-DSTNTT_TU_DIR={d}"]))
    return out


def build(verbose: bool = False, jobs: int | None = None) -> str:
    from .tuning.gen_tuned import generate
    generate()  # csrc/tuned_gfx950.hpp from tuning/gfx950.json
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(len(units()), max(1, min(os.cpu_count() or 1, 16)))

    def compile_one(u):
        obj, src, extra = u
        cmd = [hipcc, *HIPCC_FLAGS, *extra, "-c", "-o", os.path.join(OBJ_DIR, obj), os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src} {extra} ({r.returncode}):\n{r.stderr[-4000:]}")
        return os.path.join(OBJ_DIR, obj)

    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, units()))
    tmp = OUT + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed ({r.returncode}):\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
