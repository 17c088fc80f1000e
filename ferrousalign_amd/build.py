"""Build recipe for the native parts of ferrousalign_amd (gfx950 only).

Produces, in-tree:
  ferrousalign_amd/libferrousalign_amd.so  HIP kernels + C ABI (hipcc)
  ferrousalign_amd/libfa_synth.so          seeded workload generator (g++)
and, as test infrastructure, oracle/liboracle.so (make -C oracle).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    # packed fp32 VALU forms are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HIP_SOURCES = ["extend_kernel.hip", "extend16_kernel.hip", "predictor.hip", "mate16_kernel.hip", "mate16_sym.hip", "mate8_kernel.hip", "valu_probe.hip",
               "ref_kernel.hip", "swtb_kernel.hip", "chain_kernel.hip", "noalloc.hip", "fa_api.cpp",
               "fa_pipeline.cpp", "chain_api.cpp", "pack_host.cc"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host (x86) pass of every hipcc line reports the GPU-only feature as
    # unknown; check_device_features() makes sure the device pass knows it
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), out))
    return out


def check_device_features():
    """The device pass alone must accept every -target-feature of HIP_FLAGS:
    an unknown one would be ignored silently, and packed fp32 VALU forms
    could enter the code objects."""
    os.makedirs(OBJ, exist_ok=True)
    src = os.path.join(OBJ, "feature_probe.hip")
    with open(src, "w") as f:
        f.write("#include <hip/hip_runtime.h>\n__global__ void probe(float *x) { x[threadIdx.x] *= 2.0f; }\n")
    cmd = [HIPCC] + HIP_FLAGS + ["--cuda-device-only", "-c", src, "-o", os.path.join(OBJ, "feature_probe.o")]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0 or "not a recognized feature" in r.stdout:
        raise RuntimeError("the gfx950 device pass does not accept the build's target features:\n" + r.stdout)


def build_hip_library():
    check_device_features()
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        objs.append(obj)
        if src.endswith(".cc"):  # plain host C++ (x86 intrinsics), no device pass
            cmds.append(["g++", "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj])
            continue
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    lib = os.path.join(PKG, "libferrousalign_amd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)
    return lib


def build_synth_library():
    lib = os.path.join(PKG, "libfa_synth.so")
    tmp = lib + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-pthread", "-o", tmp,
          os.path.join(CSRC, "synth.cpp")])
    os.replace(tmp, lib)
    return lib


def build_oracle():
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all(verbose=True):
    libs = [build_synth_library(), build_oracle(), build_hip_library()]
    if verbose:
        for l in libs:
            print("built", os.path.relpath(l, ROOT))
    return libs


if __name__ == "__main__":
    build_all()
    sys.exit(0)
