"""The Rust shim crate (rust/rblas-hip) against the C header, without a Rust
toolchain (none is installed here, so the crate is checked as text):

  - its `extern "C"` block declares every entry point of include/*.h with
    the same parameter names, count, order and mapped types, and the same
    return type;
  - its #[repr(C)] structs have the header structs' fields, in order, with
    mapped types;
  - the wrappers of rblas's routed functions keep rblas's parameter lists
    (names and types, in order) and return types;
  - the per-type trait maps f32 to the s / c entry points and f64 to d / z,
    one method per entry point, in the macro's parameter order.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_RS = os.path.join(ROOT, "rust", "rblas-hip", "src", "lib.rs")

# C type -> Rust FFI type (by value / pointee).
C_TO_RUST = {
    "int32_t": "i32", "size_t": "usize", "float": "f32", "double": "f64", "void": "c_void",
    "char": "c_char", "uint64_t": "u64",
    "rblas_complex32_t": "C32", "rblas_complex64_t": "C64", "rblas_error_t": "RblasError",
    "rblas_grouped_t": "Grouped", "rblas_sgemm_problem_t": "SgemmProblem",
    "rblas_dgemm_problem_t": "DgemmProblem", "rblas_strsm_problem_t": "StrsmProblem",
    "rblas_dtrsm_problem_t": "DtrsmProblem", "rblas_grouped_info_t": "GroupedInfo",
    "rblas_serve_request_t": "ServeRequest", "rblas_serve_config_t": "ServeConfig",
    "rblas_serve_result_t": "ServeResult", "rblas_serve_stats_t": "ServeStats",
}


def rust_type(ctype: str) -> str:
    """`const T *` -> `*const T'`, `T *` -> `*mut T'`, `T **` -> `*mut *mut
    T'`, `T *const *` -> `*const *mut T'`."""
    c = " ".join(ctype.split())
    if c.endswith("*const *"):  # pointer to const pointer
        return "*const *mut " + C_TO_RUST[c[:-len("*const *")].strip()]
    const = c.startswith("const ")
    if const:
        c = c[len("const "):]
    stars = c.count("*")
    base = C_TO_RUST[c.replace("*", "").strip()]
    if stars == 0:
        return base
    if stars == 1:
        return ("*const " if const else "*mut ") + base
    assert stars == 2 and not const, ctype
    return "*mut *mut " + base


def _strip_c_comments(text):
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def header_text():
    out = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            out.append(_strip_c_comments(open(os.path.join(ROOT, "include", fn)).read()))
    return "\n".join(out)


def header_prototypes():
    protos = {}
    for ret, name, args in re.findall(r"([A-Za-z_][\w \*]*?)\b(rblas_\w+)\s*\(([^;{]*?)\)\s*;",
                                      header_text(), flags=re.S):
        args = " ".join(args.split())
        params = []
        if args != "void":
            for a in args.split(","):
                m = re.match(r"(.*?)(\w+)$", a.strip())
                params.append((m.group(2), rust_type(m.group(1))))
        protos[name] = (params, rust_type(ret.strip()))
    return protos


def header_structs():
    structs = {}
    text = header_text()
    for body, name in re.findall(r"typedef struct\s*\w*\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"((?:const )?\w+(?: \*const \*| \*+)?)\s*(.*)$", decl)
            ctype, names = m.group(1), m.group(2)
            for nm in names.split(","):
                nm = nm.strip()
                stars = len(nm) - len(nm.lstrip("*"))
                t = ctype + " " + "*" * stars if stars else ctype
                fields.append((nm.lstrip("*").strip(), rust_type(t)))
        structs[C_TO_RUST[name]] = fields
    return structs


def rust_text():
    text = open(LIB_RS).read()
    return re.sub(r"//[^\n]*", "", text)


def rust_extern_block():
    text = rust_text()
    m = re.search(r'extern "C" \{(.*?)\n    \}', text, flags=re.S)
    assert m, "no extern \"C\" block"
    decls = {}
    for name, args, ret in re.findall(r"pub fn (rblas_\w+)\((.*?)\)\s*->\s*([^;]+);",
                                      m.group(1), flags=re.S):
        args = " ".join(args.split())
        params = []
        if args:
            for a in args.split(","):
                pn, pt = a.split(":", 1)
                params.append((pn.strip(), " ".join(pt.split())))
        assert name not in decls, f"{name} declared twice"
        decls[name] = (params, " ".join(ret.split()))
    return decls


def rust_structs():
    out = {}
    for name, body in re.findall(r"#\[repr\(C\)\][^{]*?pub struct (\w+) \{(.*?)\n    \}",
                                 rust_text(), flags=re.S):
        fields = []
        for f in body.split(","):
            f = " ".join(f.split())
            if not f:
                continue
            nm, ty = f.split(":", 1)
            fields.append((nm.replace("pub ", "").strip(), ty.strip()))
        out[name] = fields
    return out


def test_extern_block_matches_header():
    hdr = header_prototypes()
    rs = rust_extern_block()
    assert len(hdr) >= 90
    missing = sorted(set(hdr) - set(rs))
    extra = sorted(set(rs) - set(hdr))
    assert not missing and not extra, (missing, extra)
    for name, (params, ret) in hdr.items():
        rparams, rret = rs[name]
        assert len(rparams) == len(params), name
        assert rparams == params, (name, rparams, params)
        assert rret == ret, (name, rret, ret)


def test_repr_c_structs_match_header():
    hdr = header_structs()
    rs = rust_structs()
    # rblas's own Complex<T> (complex.rs:17-24, #[repr(C)] {re, im}) stands
    # for the header's complex structs
    assert hdr.pop("C32") == [("re", "f32"), ("im", "f32")]
    assert hdr.pop("C64") == [("re", "f64"), ("im", "f64")]
    for name, fields in hdr.items():
        assert name in rs, name
        assert rs[name] == fields, (name, rs[name], fields)
    assert "Grouped" in rs  # the opaque plan handle


# rblas's public signatures of the routed functions (parameter name, type),
# level3.rs:32-47, 68-83, 236-247, 301-314, 390-402; complex3.rs:27-41,
# 130-143, 151-164, 252-264; level1.rs:17, 45, 96; level2.rs:11-24, 65-76;
# complex2.rs:11-24, 80-91, 97-108. T is the crate's element type.
_C = "Complex<T>"
RBLAS_SIGNATURES = {
    ("level3", "gemm"): ([("layout", "Layout"), ("transa", "Transpose"), ("transb", "Transpose"),
                          ("m", "usize"), ("n", "usize"), ("k", "usize"), ("alpha", "T"),
                          ("a", "&[T]"), ("lda", "usize"), ("b", "&[T]"), ("ldb", "usize"),
                          ("beta", "T"), ("c", "&mut [T]"), ("ldc", "usize")], None),
    ("level3", "try_gemm"): ("gemm", "rblas::Result<()>"),
    ("level3", "syrk"): ([("uplo", "Uplo"), ("trans", "Transpose"), ("n", "usize"),
                          ("k", "usize"), ("alpha", "T"), ("a", "&[T]"), ("lda", "usize"),
                          ("beta", "T"), ("c", "&mut [T]"), ("ldc", "usize")], None),
    ("level3", "symm"): ([("side", "Side"), ("uplo", "Uplo"), ("m", "usize"), ("n", "usize"),
                          ("alpha", "T"), ("a", "&[T]"), ("lda", "usize"), ("b", "&[T]"),
                          ("ldb", "usize"), ("beta", "T"), ("c", "&mut [T]"),
                          ("ldc", "usize")], None),
    ("level3", "trsm"): ([("side", "Side"), ("uplo", "Uplo"), ("trans", "Transpose"),
                          ("diag", "Diag"), ("m", "usize"), ("n", "usize"), ("alpha", "T"),
                          ("a", "&[T]"), ("lda", "usize"), ("b", "&mut [T]"),
                          ("ldb", "usize")], None),
    ("complex3", "gemm"): ([("transa", "Transpose"), ("transb", "Transpose"), ("m", "usize"),
                            ("n", "usize"), ("k", "usize"), ("alpha", _C), ("a", f"&[{_C}]"),
                            ("lda", "usize"), ("b", f"&[{_C}]"), ("ldb", "usize"),
                            ("beta", _C), ("c", f"&mut [{_C}]"), ("ldc", "usize")], None),
    ("complex3", "symm"): ([("side", "Side"), ("uplo", "Uplo"), ("m", "usize"), ("n", "usize"),
                            ("alpha", _C), ("a", f"&[{_C}]"), ("lda", "usize"),
                            ("b", f"&[{_C}]"), ("ldb", "usize"), ("beta", _C),
                            ("c", f"&mut [{_C}]"), ("ldc", "usize")], None),
    ("complex3", "hemm"): ("complex3.symm", None),
    ("complex3", "trsm"): ([("side", "Side"), ("uplo", "Uplo"), ("trans", "Transpose"),
                            ("diag", "Diag"), ("m", "usize"), ("n", "usize"), ("alpha", _C),
                            ("a", f"&[{_C}]"), ("lda", "usize"), ("b", f"&mut [{_C}]"),
                            ("ldb", "usize")], None),
    ("level1", "axpy"): ([("n", "usize"), ("alpha", "T"), ("x", "&[T]"), ("incx", "usize"),
                          ("y", "&mut [T]"), ("incy", "usize")], None),
    ("level1", "dot"): ([("n", "usize"), ("x", "&[T]"), ("incx", "usize"), ("y", "&[T]"),
                         ("incy", "usize")], "T"),
    ("level1", "nrm2"): ([("n", "usize"), ("x", "&[T]"), ("incx", "usize")], "T"),
    ("level2", "gemv"): ([("layout", "Layout"), ("trans", "Transpose"), ("m", "usize"),
                          ("n", "usize"), ("alpha", "T"), ("a", "&[T]"), ("lda", "usize"),
                          ("x", "&[T]"), ("incx", "usize"), ("beta", "T"), ("y", "&mut [T]"),
                          ("incy", "usize")], None),
    ("level2", "ger"): ([("layout", "Layout"), ("m", "usize"), ("n", "usize"), ("alpha", "T"),
                         ("x", "&[T]"), ("incx", "usize"), ("y", "&[T]"), ("incy", "usize"),
                         ("a", "&mut [T]"), ("lda", "usize")], None),
    ("complex2", "gemv"): ([("layout", "Layout"), ("trans", "Transpose"), ("m", "usize"),
                            ("n", "usize"), ("alpha", _C), ("a", f"&[{_C}]"), ("lda", "usize"),
                            ("x", f"&[{_C}]"), ("incx", "usize"), ("beta", _C),
                            ("y", f"&mut [{_C}]"), ("incy", "usize")], None),
    ("complex2", "geru"): ([("layout", "Layout"), ("m", "usize"), ("n", "usize"),
                            ("alpha", _C), ("x", f"&[{_C}]"), ("incx", "usize"),
                            ("y", f"&[{_C}]"), ("incy", "usize"), ("a", f"&mut [{_C}]"),
                            ("lda", "usize")], None),
    ("complex2", "gerc"): ("complex2.geru", None),
}


def _expected(key):
    params, ret = RBLAS_SIGNATURES[key]
    if isinstance(params, str):  # same parameters as another entry
        other = tuple(params.split(".")) if "." in params else (key[0], params)
        params = RBLAS_SIGNATURES[other][0]
    return params, ret


def rust_module_fns(module):
    text = rust_text()
    m = re.search(r"pub mod " + module + r" \{(.*?)\n\}", text, flags=re.S)
    assert m, module
    fns = {}
    for name, args, ret in re.findall(r"pub fn (\w+)<T: HipFloat>\((.*?)\)\s*(?:->\s*([^{]+?))?\s*\{",
                                      m.group(1), flags=re.S):
        params = []
        for a in re.split(r",(?![^\[<]*[\]>])", " ".join(args.split())):
            a = a.strip()
            if a:
                pn, pt = a.split(":", 1)
                params.append((pn.strip(), pt.strip()))
        fns[name] = (params, ret.strip() if ret else None)
    return fns


@pytest.mark.parametrize("module,name", sorted(RBLAS_SIGNATURES))
def test_wrapper_keeps_rblas_signature(module, name):
    fns = rust_module_fns(module)
    assert name in fns, f"{module}::{name} missing"
    want_params, want_ret = _expected((module, name))
    got_params, got_ret = fns[name]
    assert got_params == want_params, (module, name, got_params)
    assert got_ret == want_ret, (module, name, got_ret)


def test_trait_routes_each_precision():
    text = rust_text()
    order = re.search(r"macro_rules! impl_hip_float \{\s*\((.*?)\) =>", text, flags=re.S).group(1)
    slots = re.findall(r"\$(\w+):ident", order)
    assert slots[0] == "gemm" and len(slots) == 17
    calls = re.findall(r"impl_hip_float!\((.*?)\);", text, flags=re.S)
    assert len(calls) == 2
    for call in calls:
        args = [a.strip() for a in call.split(",")]
        ty, names = args[0], args[1:]
        real, cplx = ("s", "c") if ty == "f32" else ("d", "z")
        assert ty in ("f32", "f64")
        want = []
        for s in slots:
            if s.startswith("c") and s[1:] in ("gemm", "symm", "hemm", "trsm", "gemv", "geru",
                                                "gerc"):
                want.append(f"rblas_{cplx}{s[1:]}")
            else:
                want.append(f"rblas_{real}{s}")
        assert names == want, (ty, names)
    hdr = header_prototypes()
    for call in calls:
        for nm in [a.strip() for a in call.split(",")][1:]:
            assert nm in hdr, nm
