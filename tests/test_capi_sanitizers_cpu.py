"""The host side of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU, no GPU): a host-sanitized build of libgpehip (build.py --sanitize; device code
unchanged) runs tests/capi_host_paths.py -- null handles and arguments on every entry
point, bad and valid grid and batch descriptions without a device, malformed plan
wisdom -- with the ASan runtime preloaded.  Any report aborts the run."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "gpe-matlab_amd")
LIB = os.path.join(PKG, "_build", "asan", "libgpehip_asan.so")


def _asan_runtime():
    clang = shutil.which("clang") or "/opt/rocm/llvm/bin/clang"
    if not os.path.exists(clang):
        return None
    r = subprocess.run([clang, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if r.returncode == 0 and os.path.isabs(p) and os.path.exists(p) else None


def _stale():
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    srcs = glob.glob(os.path.join(PKG, "csrc", "*")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    return any(os.path.getmtime(s) > t for s in srcs)


@pytest.mark.timeout(900)
def test_host_paths_clean_under_asan_and_ubsan():
    rt = _asan_runtime()
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if rt is None or not os.path.exists(hipcc):
        pytest.skip("clang ASan runtime or hipcc not available")
    if _stale():
        r = subprocess.run([sys.executable, os.path.join(PKG, "build.py"), "--sanitize", "-j", "8"],
                           capture_output=True, text=True, timeout=900)
        assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "capi_host_paths.py"), LIB],
                       capture_output=True, text=True, env=env, timeout=300, cwd="/tmp")
    report = r.stdout + r.stderr
    assert "AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert r.returncode == 0 and "OK" in r.stdout, report[-4000:]
