"""The ctypes mirrors of the C-ABI structs match include/rblas_hip.h: size and
every field offset, as a C compiler lays them out (gcc on this host)."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from rblas_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rblas_hip.h")

STRUCTS = {
    "rblas_error_t": _ffi.RblasError,
    "rblas_complex32_t": _ffi.Complex32,
    "rblas_complex64_t": _ffi.Complex64,
    "rblas_sgemm_problem_t": _ffi.SgemmProblem,
    "rblas_dgemm_problem_t": _ffi.DgemmProblem,
    "rblas_strsm_problem_t": _ffi.StrsmProblem,
    "rblas_dtrsm_problem_t": _ffi.DtrsmProblem,
    "rblas_grouped_info_t": _ffi.GroupedInfo,
    "rblas_serve_config_t": _ffi.ServeConfig,
    "rblas_serve_stats_t": _ffi.ServeStats,
}


def test_struct_layouts_match_header(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler on this host")
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        name, field, value = line.split()
        got[(name, field)] = int(value)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)


def test_result_records_match_header(tmp_path):
    """The numpy record types of the serving loop match the C structs."""
    from rblas_amd import serve
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler on this host")
    pairs = {"rblas_serve_request_t": serve.REQUEST_DTYPE,
             "rblas_serve_result_t": serve.RESULT_DTYPE}
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, dt in pairs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field in dt.names:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "rec.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "rec"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {(a, b): int(c) for a, b, c in (line.split() for line in out.splitlines())}
    for cname, dt in pairs.items():
        assert got[(cname, "size")] == dt.itemsize, cname
        for field in dt.names:
            assert got[(cname, field)] == dt.fields[field][1], (cname, field)
