"""Host-side paths of the C ABI that need no GPU, run by test_capi_sanitizers_cpu.py
inside a process with the ASan/UBSan runtime preloaded (python capi_host_paths.py LIB).

Every declared entry point is called with null handles and null/zero arguments (each
must return an error code), the grid and batch constructors with bad and valid
descriptions (valid ones fail cleanly without a device), and the host-only plan wisdom
with well-formed, malformed, truncated and oversized input.  Prints OK at the end."""
import ctypes as C
import os
import re
import sys

LIB = sys.argv[1]
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [ROOT, os.path.join(ROOT, "gpe-matlab_amd")]
os.environ["GPE_HIP_LIB"] = LIB

lib = C.CDLL(LIB)
import gpe_hip._lib as L  # noqa: E402  (signature table only; nothing is loaded from LIBPATH here)

lib.gpe_last_error.restype = C.c_char_p
lib.gpe_last_error.argtypes = [C.c_void_p]
lib.gpe_batch_last_error.restype = C.c_char_p
lib.gpe_batch_last_error.argtypes = [C.c_void_p]

# 1. every entry point with null / zero arguments returns a negative code (or 0 for
#    the few that accept nothing to do) and never dereferences a null pointer
skip = {"gpe_last_error", "gpe_batch_last_error", "gpe_wisdom_forget", "gpe_destroy", "gpe_batch_destroy"}
for name, argtypes in sorted(L.SIGNATURES.items()):
    if name in skip:
        continue
    fn = getattr(lib, name)
    fn.argtypes = argtypes
    fn.restype = C.c_int
    args = [0 if t in (C.c_int, C.c_double, C.c_longlong, C.c_ulonglong) else None for t in argtypes]
    rc = fn(*args)
    assert rc <= 0, (name, rc)
    if name not in ("gpe_wisdom_get",):
        assert rc < 0 or name in ("gpe_wisdom_export",), (name, rc)
assert lib.gpe_destroy(None) == 0 and lib.gpe_batch_destroy(None) == 0
assert lib.gpe_last_error(None) == b"null handle"

# 2. grid descriptions: rejected before any device call, or failing cleanly without one
h = C.c_void_p()
for nd, n, xm in [(4, (8, 8, 8), (1, 1, 1)), (2, (100, 64, 1), (1, 1, 1)), (2, (64, 64, 1), (1, -1, 1)),
                  (3, (512, 512, 1024), (1, 1, 1)), (2, (0, 64, 1), (1, 1, 1)), (3, (64, 64, 64), (1, 1, 1))]:
    d = L.GridDesc()
    d.ndims = nd
    for i in range(3):
        d.n[i], d.xmax[i] = n[i], xm[i]
    rc = lib.gpe_create(C.byref(d), C.byref(h))
    assert rc < 0 and not h.value, (nd, n, rc)

b = C.c_void_p()
descs = (L.GridDesc * 3)()
for i, (nx, ny) in enumerate([(64, 64), (96, 128), (12, 8)]):
    descs[i].ndims = 2
    descs[i].n[0], descs[i].n[1] = nx, ny
    descs[i].xmax[0] = descs[i].xmax[1] = 5.0
assert lib.gpe_batch_create(0, descs, C.byref(b)) < 0
assert lib.gpe_batch_create(3, descs, C.byref(b)) < 0 and not b.value
buf = C.create_string_buffer(256)
assert lib.gpe_device_info(buf, 0) < 0
lib.gpe_device_info(buf, 256)
assert len(buf.value) < 256

# 3. plan wisdom: host-only state
lib.gpe_wisdom_export.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int)]
lib.gpe_wisdom_import.argtypes = [C.c_char_p]
assert lib.gpe_wisdom_forget() == 0
for ax, n in [(0, 64), (1, 128), (2, 256), (0, 384), (1, 512)]:
    assert lib.gpe_wisdom_set(3, ax, n, 1) == 0
assert lib.gpe_wisdom_set(3, 0, 100, 1) < 0 and lib.gpe_wisdom_set(5, 0, 64, 0) < 0
need = C.c_int()
assert lib.gpe_wisdom_export(None, 0, C.byref(need)) == 0
small = C.create_string_buffer(4)
assert lib.gpe_wisdom_export(small, 4, C.byref(need)) < 0       # too small: refused, nothing written past it
full = C.create_string_buffer(need.value)
assert lib.gpe_wisdom_export(full, len(full), C.byref(need)) == 0
text = full.value
assert lib.gpe_wisdom_forget() == 0
for bad in [b"3 0 64", b"3 0 64 1 7\n", b"x y z w\n", b"3 0 64 9\n", b"\n\n\n", b"3 0 64 1\n" * 10000 + b"junk",
            b"9" * 5000 + b" 0 64 1\n"]:
    rc = lib.gpe_wisdom_import(bad)
    assert rc <= 0
assert lib.gpe_wisdom_forget() == 0
assert lib.gpe_wisdom_import(text) == 0
assert lib.gpe_wisdom_get(3, 2, 256) == 1
assert lib.gpe_wisdom_forget() == 0
print("OK", flush=True)
