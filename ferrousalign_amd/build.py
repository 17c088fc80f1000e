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
HIP_SOURCES = ["extend_kernel.hip", "extend16_kernel.hip", "predictor.hip", "mate16_kernel.hip", "mate16_sym.hip", "mate8_kernel.hip", "valu_probe.hip", "extend_long_kernel.hip",
               "ref_kernel.hip", "swtb_kernel.hip", "chain_kernel.hip", "noalloc.hip", "plan_kernel.hip", "fa_api.cpp",
               "fa_pipeline.cpp", "chain_api.cpp", "aggregator.cpp", "pack_host.cc"]


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


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer for the sanitized
# build (host code only: each -fsanitize= goes to the host pass alone).
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined",
            "-Xarch_host", "-shared-libasan"]
SAN_LIB = os.path.join(OBJ, "san", "libferrousalign_amd.so")


def build_hip_library(sanitize=False):
    check_device_features()
    obj_dir = os.path.join(OBJ, "san") if sanitize else OBJ
    os.makedirs(obj_dir, exist_ok=True)
    san = SAN_HOST if sanitize else []
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        objs.append(obj)
        if src.endswith(".cc"):  # plain host C++ (x86 intrinsics), no device pass
            cxx = ["/opt/rocm/llvm/bin/clang++"] + san[1::2] if sanitize else ["g++"]
            cmds.append(cxx + ["-O3", "-std=c++17", "-fPIC", "-Wall", "-c", os.path.join(CSRC, src), "-o", obj])
            continue
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC] + HIP_FLAGS + san + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    lib = SAN_LIB if sanitize else os.path.join(PKG, "libferrousalign_amd.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + san + objs)
    os.replace(tmp, lib)
    return lib


def asan_runtime():
    """clang's shared ASan runtime (it carries the UBSan handlers too)."""
    r = subprocess.run(["/opt/rocm/llvm/bin/clang++", "-print-file-name=libclang_rt.asan-x86_64.so"],
                       stdout=subprocess.PIPE, text=True)
    path = r.stdout.strip()
    if not os.path.isabs(path) or not os.path.exists(path):
        import glob
        hits = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
        path = hits[0] if hits else ""
    return path


def run_sanitized_cpu_tests(pytest_args=()):
    """Builds the sanitized library and runs the CPU test suite against it
    (FA_LIBRARY points the ctypes loader there; the runtime is preloaded
    because the interpreter itself is not instrumented).  Host only: the
    sanitized library never goes to a GPU."""
    lib = build_hip_library(sanitize=True)
    env = dict(os.environ)
    env["FA_LIBRARY"] = lib
    env["LD_PRELOAD"] = asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-q", "-x", "-m", "not gpu",
           "-p", "no:cacheprovider"] + list(pytest_args)
    return subprocess.run(cmd, env=env, cwd=ROOT).returncode


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
    if len(sys.argv) > 1 and sys.argv[1] == "--sanitize":
        sys.exit(run_sanitized_cpu_tests(sys.argv[2:]))
    build_all()
    sys.exit(0)
