"""In-tree build of the native libraries (hipcc for gfx950, g++ for host-only code)."""
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Device code: IEEE float semantics (no contraction, no fast math), correctly
# rounded fp32 division, and no packed-fp32 instructions.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument", "-Wno-unused-result", "-Wno-unused-value",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", "-fno-fast-math"]

HIP_SOURCES = ["mh_index.hip", "mh_classify.hip", "mh_api.hip", "mh_serve.hip",
               "mh_autotune.hip", "mh_mhm.hip", "mh_fasta.hip"]
HOST_SOURCES = ["matrix_hip.cpp", "matrix_hip_check.cpp", "mh_results.cpp"]


def datagen_lib_path():
    return os.path.join(LIBDIR, "libmh_datagen.so")


def hip_lib_path():
    return os.path.join(LIBDIR, "libmirage_hip.so")


def _run(cmd):
    subprocess.run(cmd, check=True)


def build_datagen():
    os.makedirs(LIBDIR, exist_ok=True)
    out = datagen_lib_path()
    tmp = out + ".tmp"
    _run(["g++", *HOST_FLAGS, "-shared", os.path.join(CSRC, "mh_datagen.cpp"), "-o", tmp,
          "-lpthread"])
    os.replace(tmp, out)


def build_hip(verbose=False):
    """Compile every HIP and host source of the classification library into
    mirage_amd/_lib/libmirage_hip.so (gfx950 code objects)."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(ROOT, "_build", "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", os.path.join(CSRC, src),
               "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for src in HOST_SOURCES:
        if not os.path.exists(os.path.join(CSRC, src)):
            continue
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC, *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}", f"-I{CSRC}", "-c",
               os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [c for c, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("compile failed: " + " ".join(failed[0]))
    out = hip_lib_path()
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread"])
    os.replace(tmp, out)
    return out


def shim_driver_path():
    return os.path.join(LIBDIR, "shim_driver")


def build_shim_driver():
    """Test driver of the C++ call surface (tests/cpp/shim_driver.cpp)."""
    src = os.path.join(ROOT, "tests", "cpp", "shim_driver.cpp")
    out = shim_driver_path()
    _run([HIPCC, *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}", f"-I{CSRC}", src,
          "-o", out, f"-L{LIBDIR}", "-lmirage_hip", f"-Wl,-rpath,{LIBDIR}",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_all():
    build_datagen()
    out = build_hip()
    build_shim_driver()
    return out


def clean():
    shutil.rmtree(LIBDIR, ignore_errors=True)
