"""CPU checks of the C-ABI library: it loads, exports every declared entry point,
and rejects bad arguments without touching a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gpe_hip.h")
LIB = os.path.join(ROOT, "gpe-matlab_amd", "gpe_hip", "libgpehip.so")


def declared_symbols():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gpe_[a-z_0-9]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} missing: run `python gpe-matlab_amd/build.py` (or __graft_entry__.build())")
    return C.CDLL(LIB)


def test_header_declares_core_entry_points():
    syms = declared_symbols()
    for s in ["gpe_create", "gpe_destroy", "gpe_set_params", "gpe_set_potential", "gpe_set_state",
              "gpe_get_state", "gpe_split_begin", "gpe_split_outer", "gpe_split_block", "gpe_itp",
              "gpe_measure", "gpe_last_error"]:
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"


def test_ctypes_signature_table_matches_header():
    import gpe_hip._lib as L
    syms = set(declared_symbols()) - {"gpe_last_error", "gpe_batch_last_error"}
    assert syms == set(L.SIGNATURES), (syms ^ set(L.SIGNATURES))


def test_create_rejects_bad_grids(lib):
    import gpe_hip._lib as L
    lib.gpe_create.argtypes = [C.POINTER(L.GridDesc), C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    d = L.GridDesc()
    d.ndims = 4
    assert lib.gpe_create(C.byref(d), C.byref(h)) == -1
    d.ndims = 2
    d.n[0], d.n[1] = 100, 64          # 100 has a factor 5: unsupported length
    d.xmax[0], d.xmax[1] = 1.0, 1.0
    assert lib.gpe_create(C.byref(d), C.byref(h)) == -1
    d.n[0] = 64
    d.xmax[1] = -1.0
    assert lib.gpe_create(C.byref(d), C.byref(h)) == -1
    lib.gpe_last_error.restype = C.c_char_p
    lib.gpe_last_error.argtypes = [C.c_void_p]
    assert lib.gpe_last_error(None) == b"null handle"


def test_package_fails_loudly_without_library(monkeypatch, tmp_path):
    import importlib
    import gpe_hip._lib as L
    monkeypatch.setattr(L, "LIBPATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(L.GPELibraryError):
        L.load()
    importlib.reload(L)


def test_profile_names_cover_every_sweep_sequence():
    # gpe_profile_read rejects a buffer shorter than 3 * S_NSEQ entries
    src = open(os.path.join(ROOT, "gpe-matlab_amd", "csrc", "sweep.h")).read()
    nseq = int(re.search(r"S_NSEQ\s*=\s*(\d+)", src).group(1))
    from gpe_hip import _lib
    assert sorted(_lib.SEQ_NAMES) == list(range(nseq))
    import bench
    assert set(_lib.SEQ_NAMES.values()) <= set(bench.SEQ_BYTES)


def test_plan_wisdom_set_export_import_roundtrip(lib):
    """The plan wisdom is host-side state: set, export, forget and import need no GPU."""
    lib.gpe_wisdom_get.restype = C.c_int
    lib.gpe_wisdom_export.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int)]
    lib.gpe_wisdom_import.argtypes = [C.c_char_p]
    assert lib.gpe_wisdom_forget() == 0
    assert lib.gpe_wisdom_get(3, 0, 256) == -1
    assert lib.gpe_wisdom_set(3, 0, 256, 1) == 0
    assert lib.gpe_wisdom_set(2, 1, 384, 1) == 0
    assert lib.gpe_wisdom_set(2, 1, 192, 1) != 0          # 192 has one plan
    assert lib.gpe_wisdom_set(2, 2, 128, 0) != 0          # axis out of range for 2D
    need = C.c_int()
    assert lib.gpe_wisdom_export(None, 0, C.byref(need)) == 0
    buf = C.create_string_buffer(need.value)
    assert lib.gpe_wisdom_export(buf, len(buf), C.byref(need)) == 0
    text = buf.value.decode()
    assert sorted(text.split("\n")[:-1]) == ["2 1 384 1", "3 0 256 1"]
    assert lib.gpe_wisdom_forget() == 0
    assert lib.gpe_wisdom_get(3, 0, 256) == -1
    assert lib.gpe_wisdom_import(b"3 0 256 1\n2 1 384 x\n") != 0   # malformed: nothing recorded
    assert lib.gpe_wisdom_get(3, 0, 256) == -1
    assert lib.gpe_wisdom_import(text.encode()) == 0
    assert lib.gpe_wisdom_get(3, 0, 256) == 1 and lib.gpe_wisdom_get(2, 1, 384) == 1
    assert lib.gpe_wisdom_forget() == 0


def test_host_order_picks_entry_point_without_copies():
    """Device.set_state / set_potential / set_coupling send a column-major array to the
    column-major entry points as it is, and any other array in C order to the
    entry points that reorder on the device, copying only when the array is neither."""
    from gpe_hip import _lib
    shape = (6, 4, 8)
    a = (np.arange(np.prod(shape)).reshape(shape) * (1 + 0.5j))
    v, c = _lib.host_order(a, shape, np.complex128)
    assert c and np.shares_memory(v, a) and np.array_equal(v.reshape(shape), a)
    f = np.asfortranarray(a)
    v, c = _lib.host_order(f, shape, np.complex128)
    assert not c and np.shares_memory(v, f) and np.array_equal(v, _lib.to_colmajor_c128(a))
    s = a[:, ::-1]
    v, c = _lib.host_order(s, shape, np.complex128)
    assert c and v.flags.c_contiguous and np.array_equal(v.reshape(shape), s)
    v, c = _lib.host_order(2.5, shape, np.float64)          # scalars broadcast to the grid
    assert c and v.shape == (np.prod(shape),) and np.all(v == 2.5)
    v, c = _lib.host_order(a.real, shape, np.complex128)    # real input converted once
    assert c and v.dtype == np.complex128 and np.array_equal(v.reshape(shape), a.real)
