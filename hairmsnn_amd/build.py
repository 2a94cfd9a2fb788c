"""Builds the HIP library (gfx950) and the headless CLIs.

    python -m hairmsnn_amd.build            # library + CLIs
    python -m hairmsnn_amd.build --jobs 8

Outputs (in-tree, git-ignored): hairmsnn_amd/libhmsnn.so and
hairmsnn_amd/bin/{render_path_tracing,render_hair_msnn,render_nrc}.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build" / "hmsnn"
LIB = PKG / "libhmsnn.so"
BIN = PKG / "bin"
ARCH = os.environ.get("HMSNN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{ROOT / 'include'}", f"-I{CSRC}"]



def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP library)")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}):\n{' '.join(cmd)}\n{r.stdout}")
    if r.stdout.strip():
        # keep warnings visible but short
        lines = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            print("\n".join(lines[:40]), file=sys.stderr)


# Sources whose floating-point code must not depend on how the compiler fuses
# multiplies and adds across statements: every shading kernel (per-pass,
# streaming, NRC) and the adjoint's replay share the same inlined routines,
# and schedules that trace the same paths must produce the same bits whatever
# the surrounding kernel looks like. (Explicit fmaf calls are kept.)
NO_CONTRACT = {"hm_pt.hip", "hm_adjoint.hip"}

# The MLP kernels keep matrix-core accumulators in ordinary VGPRs, so the
# ReLU/convert tails read them without copies out of the accumulator file.
MLP_FLAGS = {"hm_mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        extra = (["-ffp-contract=off"] if src.name in NO_CONTRACT else []) + MLP_FLAGS.get(src.name, [])
        cmd = [hipcc(), *HIP_FLAGS, *extra, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [hipcc(), *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    srcs = sorted(p for p in CSRC.iterdir() if p.suffix in (".hip", ".cpp") and not p.name.startswith("cli_"))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs), "-lz"])
    os.replace(tmp, LIB)
    BIN.mkdir(exist_ok=True)
    clis = sorted(CSRC.glob("cli_*.cpp"))
    def _cli(src: Path) -> None:
        name = src.stem[len("cli_"):]
        _run(["g++", *HOST_FLAGS, str(src), "-o", str(BIN / name), f"-L{PKG}", "-lhmsnn",
              "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"])
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_cli, clis))
    if verbose:
        print(f"built {LIB} ({len(srcs)} sources) and {len(clis)} CLIs")
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs)


if __name__ == "__main__":
    main()
