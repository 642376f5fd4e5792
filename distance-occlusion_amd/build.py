"""Build libdolhip.so (the gfx950 kernels and the C ABI of include/dol_hip.h)
in-tree, next to this file.

    python distance-occlusion_amd/build.py [-v]

Every source is compiled with hipcc for gfx950 only.  Floating-point flags
keep the reference's float semantics: no FMA contraction, correctly rounded
fp32 divide and sqrt.  Packed-fp32 instruction selection is disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdolhip.so")

SOURCES = [
    "dol_init.hip",
    "dol_jfa.hip",
    "dol_edt.hip",
    "dol_edt_tile.hip",
    "dol_encode.hip",
    "dol_batch.hip",
    "dol_seeds.hip",
    "dol_voxel.hip",
    "dol_sweep.hip",
    "dol_occlusion.hip",
    "dol_brute.hip",
    "dol_raster.hip",
    "dol_capi.cpp",
    "dol_scenes.cpp",
    "dol_shim.cpp",
]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def _compile(src, verbose):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC, "-x", "hip", *CFLAGS, "-c", path, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False, jobs=None):
    """Compile every source and link libdolhip.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, verbose), srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


CLI_SRC = os.path.join(HERE, "cli", "distocc_hip.cpp")
CLI_BIN = os.path.join(HERE, "cli", "distocc_hip")


def build_cli(verbose=False):
    """The distocc_hip command (DOL:: interface and C ABI over libdolhip.so)."""
    cxx = shutil.which("g++") or "c++"
    cmd = [cxx, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), CLI_SRC,
           "-L" + HERE, "-ldolhip", "-Wl,-rpath,$ORIGIN/..", "-o", CLI_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"distocc_hip build failed:\n{r.stdout}\n{r.stderr}")
    return CLI_BIN



TORCH_EXT_DIR = os.path.join(HERE, "torch_ext")


def build_torch_ext(verbose=False):
    """The libtorch autograd Function (csrc/dol_torch.cpp), linked to
    libdolhip.so, built in-tree into torch_ext/dol_torch*.so."""
    import torch.utils.cpp_extension as ce

    os.makedirs(TORCH_EXT_DIR, exist_ok=True)
    os.environ.setdefault("PYTORCH_ROCM_ARCH", ARCH)
    return ce.load(
        name="dol_torch",
        sources=[os.path.join(CSRC, "dol_torch.cpp")],
        extra_include_paths=[os.path.join(ROOT, "include")],
        extra_cflags=["-O2"],
        extra_ldflags=[f"-L{HERE}", "-ldolhip", f"-Wl,-rpath,{HERE}"],
        build_directory=TORCH_EXT_DIR,
        with_cuda=True,
        verbose=verbose,
    )


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
