"""Builds libeqfluid.so (HIP, gfx950) in-tree with hipcc.

Flags that matter for parity: -ffp-contract=off (the reference's rustc f32
code never fuses a*b+c into an FMA; hipcc would by default) and no fast-math
(IEEE division, denormals kept).  -packed-fp32-ops is disabled for all device
code.
"""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
SOURCES = ["eqf_runtime.hip", "eqf_wavefront.hip", "eqf_redblack.hip", "eqf_server.hip", "eqf_setup.hip",
           "eqf_render.hip"]
OUT = os.path.join(PKG_DIR, "libeqfluid.so")
ARCH = os.environ.get("EQF_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO_DIR, "include"),
]


def build(verbose=False):
    """Compiles every source to an object in parallel (each carries its own
    gfx950 code object; no relocatable device code), then links the shared
    library."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(PKG_DIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc] + cflags + ["-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        return cmd, obj, r

    jobs = int(os.environ.get("EQF_BUILD_JOBS", min(len(SOURCES), os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        results = list(ex.map(compile_one, SOURCES))
    objs = []
    for cmd, obj, r in results:
        if verbose:
            print(" ".join(cmd), flush=True)
        err = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed:\n" + err)
        if err.strip() and verbose:
            print(err, file=sys.stderr)
        objs.append(obj)
    cmd = [hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


BWPROBE_OUT = os.path.join(PKG_DIR, "libeqf_bwprobe.so")


def build_bwprobe(verbose=False):
    """libeqf_bwprobe.so: the streaming kernels that measure the practical
    HBM bandwidth ceiling (loaded by bench.py only)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + FLAGS + [os.path.join(CSRC, "eqf_bwprobe.hip"), "-o", BWPROBE_OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(BWPROBE_OUT + ".tmp", BWPROBE_OUT)
    return BWPROBE_OUT


def build_cpp_tests(verbose=False):
    """g++ build of tests/cpp/test_fluid_api (host code only) against
    libeqfluid.so and the oracle library."""
    src = os.path.join(REPO_DIR, "tests", "cpp", "test_fluid_api.cpp")
    outdir = os.path.join(REPO_DIR, "tests", "cpp", "_build")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "test_fluid_api")
    cmd = ["g++", "-std=c++17", "-O1", "-ffp-contract=off", src, "-o", out,
           "-L" + PKG_DIR, "-leqfluid", "-L" + os.path.join(REPO_DIR, "oracle"), "-l:liboracle.so",
           "-Wl,-rpath,$ORIGIN/../../../equilibrium_amd:$ORIGIN/../../../oracle:/opt/rocm/lib",
           "-L/opt/rocm/lib", "-lamdhip64"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stdout + r.stderr)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
