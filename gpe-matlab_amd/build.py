"""Build the gfx950 HIP library libgpehip.so in-tree.

    python gpe-matlab_amd/build.py [-j JOBS]

Every pencil length is compiled as its own object (sweep_kernels.hip with
-DGPE_N=<n>) so the build parallelises; the objects are linked with hipcc into
gpe-matlab_amd/gpe_hip/libgpehip.so, which the ctypes layer loads.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "gpe_hip", "libgpehip.so")
LENGTHS = (8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512)
SOURCES = ("state_kernels.hip", "tune.hip", "gpe_capi.hip", "spinor.hip", "adjoint.hip", "rabi_adjoint.hip", "itp2.hip", "vtd.hip", "sgpe.hip", "diag.hip", "batch.hip", "applyham.hip")
ARCH = os.environ.get("GPE_OFFLOAD_ARCH", "gfx950")


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found (ROCm toolchain required to build the HIP library)")
    return p


def flags():
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wno-unused-command-line-argument",
            # packed FP32 instructions are disabled for all device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(HERE, "..", "include"), "-I", CSRC]


DEFINES = []

# Host-side AddressSanitizer + UndefinedBehaviorSanitizer for the CPU tests of the C ABI
# (the device code is built as usual; each sanitizer flag applies to the host pass only).
SANITIZE_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer",
                  "-Xarch_host", "-fsanitize=address",
                  "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-fno-sanitize-recover=undefined"]
SANITIZE = False


def compile_one(src, obj, extra=()):
    cmd = [hipcc(), *flags(), *DEFINES, *(SANITIZE_FLAGS if SANITIZE else ()), *extra,
           "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=None, verbose=True, out=None, build_dir=None, sanitize=None):
    global BUILD, OUT, SANITIZE
    if sanitize is not None:
        SANITIZE = bool(sanitize)
    if out:
        OUT = out
    if build_dir:
        BUILD = build_dir
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    tasks = []
    for n in LENGTHS:
        tasks.append(("sweep_kernels.hip", os.path.join(BUILD, f"sweep_{n}.o"), (f"-DGPE_N={n}",)))
        tasks.append(("spinor_kernels.hip", os.path.join(BUILD, f"spinor_{n}.o"), (f"-DGPE_N={n}",)))
        tasks.append(("adjoint_kernels.hip", os.path.join(BUILD, f"adjoint_{n}.o"), (f"-DGPE_N={n}",)))
        tasks.append(("batch_kernels.hip", os.path.join(BUILD, f"batch_{n}.o"), (f"-DGPE_N={n}",)))
    for n in (64, 128, 256):   # x lengths of the y-split 3D step
        tasks.append(("ysplit_kernels.hip", os.path.join(BUILD, f"ysplit_{n}.o"), (f"-DGPE_N={n}",)))
    for s in SOURCES:
        if os.path.exists(os.path.join(CSRC, s)):
            tasks.append((s, os.path.join(BUILD, s.replace(".hip", ".o")), ()))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(compile_one, *t) for t in tasks]
        for f in futs:
            objs.append(f.result())
    tmp = OUT + ".tmp"
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libasan"] if SANITIZE else []
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *san, "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} ({len(objs)} objects)")
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=None)
    ap.add_argument("-D", "--define", action="append", default=[], help="extra -D for every source")
    ap.add_argument("--out", default=None, help="output library path (default gpe_hip/libgpehip.so)")
    ap.add_argument("--build-dir", default=None)
    ap.add_argument("--sanitize", action="store_true",
                    help="host-side ASan + UBSan build (into _build/asan/libgpehip_asan.so unless --out)")
    a = ap.parse_args()
    DEFINES.extend(f"-D{d}" for d in a.define)
    if a.sanitize:
        SANITIZE = True
        a.out = a.out or os.path.join(HERE, "_build", "asan", "libgpehip_asan.so")
        a.build_dir = a.build_dir or os.path.join(HERE, "_build", "asan")
    try:
        build(a.jobs, out=a.out, build_dir=a.build_dir)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
