"""Builds libdel3d_amd.so (gfx950) in-tree with hipcc.

Device code is compiled with -ffp-contract=off (the reference's FP policy,
CMakeLists.txt:81: no product may be fused into an FMA, since the filter's
error bound assumes separately rounded operations) and without packed FP32.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libdel3d_amd.so"
SOURCES = ["capi.hip", "predicates.hip", "predicates2.hip", "circumcenter.hip", "batch.hip", "server.hip",
           "spatial_sort.hip", "bridge.hip", "host_scalar.cpp", "dropin.cpp"]
ARCH = os.environ.get("DEL3D_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-result", "-Wno-bitwise-instead-of-logical",
]
HIP_ONLY = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and Path(c).exists():
            return c
    return "hipcc"


def _compile(src: str, outdir: Path, verbose: bool) -> str:
    obj = outdir / (src + ".o")
    cmd = [hipcc()] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip", f"--offload-arch={ARCH}"] + HIP_ONLY
    cmd += ["-c", str(CSRC / src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if err.strip():
        print(err, file=sys.stderr, flush=True)
    if r.returncode != 0:
        raise subprocess.CalledProcessError(r.returncode, cmd)
    return str(obj)


def build(verbose: bool = False) -> Path:
    from concurrent.futures import ThreadPoolExecutor

    outdir = HERE / "_obj"
    outdir.mkdir(exist_ok=True)
    # the slowest file first so the pool overlaps the rest with it
    order = sorted(SOURCES, key=lambda s: s != "predicates.hip")
    with ThreadPoolExecutor(max_workers=len(order)) as ex:
        objs = list(ex.map(lambda s: _compile(s, outdir, verbose), order))
    # measurement helpers, separate from the product library: a copy kernel of
    # known size (counter calibration) and the benchmarks' query generator
    for src, lib in (("calib.hip", "libdel3d_calib.so"), ("workload_gen.hip", "libdel3d_workloads.so")):
        obj = _compile(src, outdir, verbose)
        subprocess.run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o",
                        str(HERE / lib), obj], check=True)
    tmp = LIB.with_suffix(".so.tmp")
    link = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp)] + objs
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
