"""Build recipe for libclmi.so (HIP, gfx950).

The library is compiled in-tree so that the built .so travels with the source tree.  Every HIP
translation unit is compiled with FP contraction off (the FP64 kernels then evaluate the
reference's expressions in the reference's operation order, without fused multiply-adds) and
with packed-FP32 math disabled.
"""
from __future__ import annotations

import concurrent.futures as _cf
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libclmi.so"
OBJDIR = PKG / "_objs"

HIP_SOURCES = [
    "scalar_pass.hip",
    "scalar_tiled.hip",
    "scalar_rows.hip",
    "scalar_rows_f32.hip",
    "scalar_rows_f32_1f.hip",
    "scalar_rows_f64.hip",
    "scalar_rows_f64_1f.hip",
    "scalar_rows_batch.hip",
    "gauge_pass.hip",
    "gauge_tiled.hip",
    "background.hip",
    "fft_kernels.hip",
    "adjoint_kernels.hip",
    "clmi_api.hip",
    "noalloc.hip",
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc() -> str:
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(h).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return h


def _compile(src: Path, obj: Path) -> None:
    cmd = [_hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")


def build_library(verbose: bool = True) -> Path:
    """Compile every HIP source for gfx950 and link libclmi.so (always rebuilds)."""
    OBJDIR.mkdir(exist_ok=True)
    objs = []
    jobs = min(len(HIP_SOURCES), max(1, (os.cpu_count() or 2)))
    with _cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for name in HIP_SOURCES:
            src = CSRC / name
            obj = OBJDIR / (name + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile, src, obj))
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
           "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


HOST = PKG / "host"
BIN = HOST / "_bin"
# Model main programs (the reference builds one executable per model header).
HOST_MODELS = {"lphi4": "models/lphi4.h", "lphi4_1f": "models/lphi4_1f.h", "lphi4U1": "models/lphi4U1.h"}


def _gxx() -> str:
    g = shutil.which("g++")
    if not g:
        raise RuntimeError("g++ not found; the host programs cannot be built")
    return g


def host_compile_cmd(src: Path, out: Path, model_header: str, model_class: str) -> list:
    """g++ command for a C++ program over the CosmoInterface mirror and libclmi.so.  MODELLABEL is the
    executable's name, as the reference's CMakeLists.txt:139 defines it (model.name, used in the
    backup and output file names)."""
    return [_gxx(), "-std=c++20", "-O2", "-Wall", "-Wextra", f"-I{HOST}", f"-I{INCLUDE}",
            f'-DCLMI_MODEL_HEADER="{model_header}"', f"-DCLMI_MODEL_CLASS={model_class}",
            f"-DMODELLABEL={out.name}", str(src),
            "-o", str(out), f"-L{PKG}", "-lclmi", f"-Wl,-rpath,{PKG}"]


def build_host(verbose: bool = True) -> Path:
    """Compile the model main programs (host/cosmolattice_main.cpp) into host/_bin/<model>."""
    BIN.mkdir(exist_ok=True)
    with _cf.ThreadPoolExecutor(max_workers=len(HOST_MODELS)) as ex:
        futs = {}
        for name, header in HOST_MODELS.items():
            cmd = host_compile_cmd(HOST / "cosmolattice_main.cpp", BIN / name, header, name)
            futs[name] = ex.submit(subprocess.run, cmd, capture_output=True, text=True)
        for name, f in futs.items():
            r = f.result()
            if r.returncode != 0:
                raise RuntimeError(f"host build failed for {name}:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {', '.join(str(BIN / n) for n in HOST_MODELS)}")
    return BIN


if __name__ == "__main__":
    build_library()
    build_host()
