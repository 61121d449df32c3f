"""Build libspme_hip.so for gfx950 with hipcc (in tree, next to this file)."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspme_hip.so")
SOURCES = ["spme_group.hip", "spme_kernels.hip", "spme_real.hip", "spme_engine.hip",
           "spme_mlp.hip", "spme_dense.hip", "spme_block.hip", "spme_stream.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPME_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _obj(src: str) -> str:
    return os.path.join(CSRC, "_build", os.path.splitext(src)[0] + ".o")


def _compile(src: str) -> str:
    obj = _obj(src)
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


PME = os.path.join(HERE, "pme")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_cli() -> str:
    """The pme driver (cli/pme.c), C against include/spme.h and the library."""
    tmp = PME + ".tmp"
    cmd = ["gcc", "-O2", "-std=gnu11", "-Wall", f"-I{INCLUDE}", os.path.join(HERE, "cli", "pme.c"),
           f"-L{HERE}", "-lspme_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib",
           "-lm", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"pme driver failed to build:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, PME)
    return PME


def build(verbose: bool = True) -> str:
    """Compile every HIP source, link the shared library, build the driver."""
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    if verbose:
        print(f"built {LIB} and {PME}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
