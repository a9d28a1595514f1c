"""CPU checks of the C ABI: the library loads, exports exactly what
include/rocblas_amd.h declares, and its host-only paths (null handle, status
strings) behave like the reference's (library/src/rocblas_auxiliary.cpp,
library/src/handle.cpp, library/src/blas3/rocblas_trsm_imp.hpp:68-69)."""
import ctypes
import os
import re
import subprocess

import pytest

import rocblas_amd as rb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rocblas_amd.h")
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared_symbols():
    names = set()
    for hdr in HEADERS:
        text = open(hdr).read()
        names |= set(re.findall(r"ROCBLAS_AMD_EXPORT[^;(]*?\b(rocblas_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_trsm_path():
    syms = declared_symbols()
    for name in ["rocblas_dtrsm", "rocblas_strsm", "rocblas_dtrsm_batched",
                 "rocblas_strsm_batched", "rocblas_dtrsm_strided_batched",
                 "rocblas_strsm_strided_batched", "rocblas_create_handle",
                 "rocblas_destroy_handle", "rocblas_set_stream", "rocblas_get_stream",
                 "rocblas_set_pointer_mode", "rocblas_get_pointer_mode"]:
        assert name in syms


def test_library_exports_every_declared_symbol():
    lib = rb.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    # and the python binding table covers the header exactly
    assert sorted(n for n, _, _ in rb.SYMBOLS) == declared_symbols()
    out = subprocess.run(["nm", "-D", "--defined-only", rb.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = sorted(l.split()[-1] for l in out.splitlines()
                      if " T " in l and not l.split()[-1].startswith("_Z"))
    assert exported == declared_symbols()


INTERNAL = [
    "rocblas_internal_trsm_workspace_size<{T}>(rocblas_side_, rocblas_operation_, int, int, int, "
    "int, unsigned long*, unsigned long*, unsigned long*, unsigned long*, unsigned long*)",
    "rocblas_internal_trsm_batched_workspace_size<{T}>(rocblas_side_, rocblas_operation_, int, "
    "int, int, int, unsigned long*, unsigned long*, unsigned long*, unsigned long*, "
    "unsigned long*)",
    "rocblas_internal_trsm_template<{T}>(_rocblas_handle*, rocblas_side_, rocblas_fill_, "
    "rocblas_operation_, rocblas_diagonal_, int, int, {T} const*, {T} const*, long, int, long, "
    "{T}*, long, int, long, int, bool, void*, void*, void*, void*, {T} const*, int, long, long)",
    "rocblas_internal_trsm_batched_template<{T}>(_rocblas_handle*, rocblas_side_, rocblas_fill_, "
    "rocblas_operation_, rocblas_diagonal_, int, int, {T} const*, {T} const* const*, long, int, "
    "long, {T}* const*, long, int, long, int, bool, void*, void*, void*, void*, "
    "{T} const* const*, int, long, long)",
    "rocblas_internal_trtri_template<{T}>(_rocblas_handle*, rocblas_fill_, rocblas_diagonal_, "
    "int, {T} const*, long, int, long, long, {T}*, long, int, long, long, int, int, {T}*)",
    "rocblas_internal_trtri_batched_template<{T}>(_rocblas_handle*, rocblas_fill_, "
    "rocblas_diagonal_, int, {T} const* const*, long, int, long, long, {T}* const*, long, int, "
    "long, long, int, int, {T}* const*)",
    "rocblas_internal_trsm_workspace_size_64<{T}>(rocblas_side_, rocblas_operation_, long, long, "
    "long, long, long, long, unsigned long*, unsigned long*, unsigned long*, unsigned long*, "
    "unsigned long*)",
    "rocblas_internal_trsm_batched_workspace_size_64<{T}>(rocblas_side_, rocblas_operation_, "
    "long, long, long, long, long, long, unsigned long*, unsigned long*, unsigned long*, "
    "unsigned long*, unsigned long*)",
    "rocblas_internal_trsm_workspace_max_size_64<{T}>(rocblas_side_, long, long, long, "
    "unsigned long*, unsigned long*, unsigned long*)",
    "rocblas_internal_trsm_batched_workspace_max_size_64<{T}>(rocblas_side_, long, long, long, "
    "unsigned long*, unsigned long*, unsigned long*, unsigned long*, unsigned long*)",
    "rocblas_internal_trsm_template_64<{T}>(_rocblas_handle*, rocblas_side_, rocblas_fill_, "
    "rocblas_operation_, rocblas_diagonal_, long, long, {T} const*, {T} const*, long, long, long, "
    "{T}*, long, long, long, long, bool, void*, void*, void*, void*, {T} const*, long, long, long)",
    "rocblas_internal_trsm_batched_template_64<{T}>(_rocblas_handle*, rocblas_side_, "
    "rocblas_fill_, rocblas_operation_, rocblas_diagonal_, long, long, {T} const*, "
    "{T} const* const*, long, long, long, {T}* const*, long, long, long, long, bool, void*, void*, "
    "void*, void*, {T} const* const*, long, long, long)",
]


def test_internal_cxx_exports():
    """The rocSOLVER-facing templates (include/rocblas_amd_internal.hpp) are
    exported for float and double with the reference's signatures
    (library/src/blas3/rocblas_trsm.hpp:173-257, rocblas_trtri.hpp:1098-1150),
    so their mangled names are the reference's."""
    out = subprocess.run(["nm", "-DC", "--defined-only", rb.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    demangled = {l.split(" ", 2)[2] for l in out.splitlines() if "rocblas_internal" in l}
    want = {"rocblas_status_ " + sig.format(T=t) for sig in INTERNAL for t in ("float", "double")}
    want.add("rocblas_internal_trtri_temp_elements(int, int)")
    assert demangled == want, sorted(demangled ^ want)


def _internal(prefix, restype, argtypes):
    out = subprocess.run(["nm", "-D", "--defined-only", rb.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    dem = subprocess.run(["nm", "-DC", "--defined-only", rb.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for raw, d in zip(out.splitlines(), dem.splitlines()):
        name = d.split(" ", 2)[2]
        if name.startswith(prefix) or name.startswith("rocblas_status_ " + prefix):
            fn = getattr(rb.lib(), raw.split()[-1])
            fn.restype, fn.argtypes = restype, argtypes
            return fn
    raise KeyError(prefix)


@pytest.mark.parametrize("T", ["float", "double"])
def test_internal_workspace_max_size_covers_every_smaller_size(T):
    """The reference's testing_trsm_internal_interfaces
    (clients/include/blas3/testing_trsm.hpp:188-240, _batched.hpp:290-350):
    for every m <= M, n <= N the per-call workspace query (transpose, no
    supplied invA) asks for no more than the max-size query reported."""
    sz = ctypes.POINTER(ctypes.c_size_t)
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    mx = _internal(f"rocblas_internal_trsm_workspace_max_size_64<{T}>", ctypes.c_int32,
                   [i32, i64, i64, i64, sz, sz, sz])
    mxb = _internal(f"rocblas_internal_trsm_batched_workspace_max_size_64<{T}>", ctypes.c_int32,
                    [i32, i64, i64, i64] + [sz] * 5)
    ws = _internal(f"rocblas_internal_trsm_workspace_size<{T}>", ctypes.c_int32,
                   [i32, i32, i32, i32, i32, i32] + [sz] * 5)
    wsb = _internal(f"rocblas_internal_trsm_batched_workspace_size<{T}>", ctypes.c_int32,
                    [i32, i32, i32, i32, i32, i32] + [sz] * 5)
    ws64 = _internal(f"rocblas_internal_trsm_workspace_size_64<{T}>", ctypes.c_int32,
                     [i32, i32, i64, i64, i64, i64, i64, i64] + [sz] * 5)
    S = rb.rocblas_status
    for side in (141, 142):
        for (M, N, batch) in [(130, 70, 3), (20, 33, 1)]:
            top = [ctypes.c_size_t(99) for _ in range(5)]
            assert mx(side, M, N, batch, *[ctypes.byref(top[i]) for i in (0, 2, 4)]) == S.success
            assert mxb(side, M, N, batch, *[ctypes.byref(x) for x in top]) == S.success
            assert all(x.value == 0 for x in top)
            for m in range(M, 0, -7):
                for n in range(N, 0, -5):
                    got = [ctypes.c_size_t(7) for _ in range(5)]
                    refs = [ctypes.byref(x) for x in got]
                    for st in (ws(side, 112, m, n, batch, 0, *refs),
                               wsb(side, 112, m, n, batch, 0, *refs),
                               ws64(side, 112, m, n, max(m, n), m, batch, 0, *refs)):
                        assert st in (S.success, S.continue_)
                        assert got[0].value <= top[0].value and got[2].value <= top[2].value
                        assert got[4].value <= top[4].value
    got = [ctypes.c_size_t(0) for _ in range(3)]
    assert mx(141, 8, 8, 1, None, ctypes.byref(got[1]), ctypes.byref(got[2])) == S.invalid_pointer


def test_internal_trtri_temp_elements():
    out = subprocess.run(["nm", "-D", "--defined-only", rb.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    name = [l.split()[-1] for l in out.splitlines() if "rocblas_internal_trtri_temp_elements"
            in l][0]
    fn = getattr(rb.lib(), name)
    fn.restype = ctypes.c_size_t
    fn.argtypes = [ctypes.c_int, ctypes.c_int]
    assert fn(100, 3) == 100 * 100 * 3
    assert fn(0, 3) == 0 and fn(5, 0) == 0


def test_header_compiles_as_c():
    src = ('#include "rocblas_amd.h"\n#include "rocblas_amd_ext.h"\n'
           'int main(void){ return (int)rocblas_status_success; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER),
                        "-x", "c", "-", "-o", "/dev/null"], input=src, text=True,
                       capture_output=True)
    assert r.returncode == 0, r.stderr


def test_status_strings():
    for s in rb.rocblas_status:
        assert rb.status_to_string(s) == "rocblas_status_" + s.name.rstrip("_")
    assert rb.status_to_string(99) == "<undefined rocblas_status value>"


def test_null_handle_everywhere():
    L = rb.lib()
    IH = rb.rocblas_status.invalid_handle
    assert L.rocblas_create_handle(None) == IH
    assert L.rocblas_destroy_handle(None) == IH
    assert L.rocblas_set_stream(None, None) == IH
    assert L.rocblas_get_stream(None, None) == IH
    assert L.rocblas_set_pointer_mode(None, 0) == IH
    assert L.rocblas_get_pointer_mode(None, None) == IH
    assert L.rocblas_start_device_memory_size_query(None) == IH
    assert L.rocblas_stop_device_memory_size_query(None, None) == IH
    assert L.rocblas_get_device_memory_size(None, None) == IH
    assert L.rocblas_set_device_memory_size(None, 0) == IH
    assert L.rocblas_set_workspace(None, None, 0) == IH
    assert not L.rocblas_is_managing_device_memory(None)
    assert not L.rocblas_is_user_managing_device_memory(None)
    assert not L.rocblas_is_device_memory_size_query(None)
    one = ctypes.c_double(1.0)
    onef = ctypes.c_float(1.0)
    args = (141, 121, 111, 131, 4, 4)
    assert L.rocblas_dtrsm(None, *args, ctypes.addressof(one), 1, 4, 1, 4) == IH
    assert L.rocblas_strsm(None, *args, ctypes.addressof(onef), 1, 4, 1, 4) == IH
    assert L.rocblas_dtrsm_batched(None, *args, ctypes.addressof(one), 1, 4, 1, 4, 1) == IH
    assert L.rocblas_dtrsm_strided_batched(None, *args, ctypes.addressof(one), 1, 4, 16, 1, 4,
                                           16, 1) == IH


def test_create_handle_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    st = rb.lib().rocblas_create_handle(ctypes.byref(h))
    assert st != rb.rocblas_status.success
    assert h.value is None
