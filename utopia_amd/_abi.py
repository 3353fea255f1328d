"""ctypes binding of the C ABI declared in include/utopia_hip.h.

The prototypes are read from the header itself (the one macro it uses is
expanded here), so the Python signatures cannot drift from the C ones.
Loading fails loudly when libutopia_hip.so has not been built: there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
HEADER = ROOT / "include" / "utopia_hip.h"
ISOLVER_HEADER = ROOT / "include" / "utopia_hip_isolver.h"
# (header, prefix of the functions it declares)
HEADERS = ((HEADER, "uhip_"), (ISOLVER_HEADER, "isolver_lsolve_"))
LIB_PATH = PKG / "lib" / "libutopia_hip.so"


class uhip_csr(ctypes.Structure):
    """Mirror of `struct uhip_csr` (include/utopia_hip.h)."""

    _fields_ = [
        ("nrows", ctypes.c_int32),
        ("ncols", ctypes.c_int32),
        ("nnz", ctypes.c_int64),
        ("rowptr", ctypes.c_void_p),
        ("col", ctypes.c_void_p),
        ("val", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("format", ctypes.c_int32),
        ("blocks", ctypes.c_void_p),
        ("nblocks", ctypes.c_int32),
        ("block_capk", ctypes.c_int32),
        ("nslices", ctypes.c_int32),
        ("slice_off", ctypes.c_void_p),
        ("sell_col", ctypes.c_void_p),
        ("sell_val", ctypes.c_void_p),
        ("nparts", ctypes.c_int32),
        ("ndiag", ctypes.c_int32),
        ("dia_off", ctypes.c_void_p),
        ("dia_val", ctypes.c_void_p),
        ("dia_mask", ctypes.c_void_p),
        ("dia_sym", ctypes.c_int32),
        ("spmm_kernel", ctypes.c_int32),
        ("vi_base", ctypes.c_void_p),
        ("vi_table", ctypes.c_void_p),
        ("vi_nvals", ctypes.c_int32),
        ("tr", ctypes.c_void_p),
        ("dia_nan", ctypes.c_int32),
        ("dia_s1", ctypes.c_int32),
        ("dia_s2", ctypes.c_int32),
        ("vi_nsig", ctypes.c_int32),
        ("dia_w27", ctypes.c_int32),
        ("dia_delta", ctypes.c_void_p),
    ]


class uhip_cheb_step(ctypes.Structure):
    """Mirror of `struct uhip_cheb_step` (include/utopia_hip.h)."""

    _fields_ = [
        ("pprev", ctypes.c_void_p),
        ("p_dtype", ctypes.c_int32),
        ("rprev", ctypes.c_void_p),
        ("xin", ctypes.c_void_p),
        ("xout", ctypes.c_void_p),
        ("x_dtype", ctypes.c_int32),
        ("rout", ctypes.c_void_p),
        ("pout", ctypes.c_void_p),
        ("a_prev", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("a_cur", ctypes.c_double),
        ("last", ctypes.c_int32),
        ("dot_with", ctypes.c_void_p),
        ("dot_d", ctypes.c_void_p),
    ]


_SCALARS = {
    "int": ctypes.c_int,
    "int32_t": ctypes.c_int32,
    "int64_t": ctypes.c_int64,
    "uint64_t": ctypes.c_uint64,
    "size_t": ctypes.c_size_t,
    "double": ctypes.c_double,
    "float": ctypes.c_float,
    "ptrdiff_t": ctypes.c_ssize_t,
    "isolver_idx_t": ctypes.c_int,
    "isolver_scalar_t": ctypes.c_double,
    "void": None,
}


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _expand_macros(text: str) -> str:
    """Expand the function-like declaration macros defined in the header."""
    macros = {}
    pattern = re.compile(r"#define\s+(\w+)\(([^)]*)\)\s*((?:[^\n]*\\\n)*[^\n]*)")
    for m in pattern.finditer(text):
        params = [p.strip() for p in m.group(2).split(",")]
        body = m.group(3).replace("\\\n", " ")
        macros[m.group(1)] = (params, body)
    text = pattern.sub(" ", text)
    for name, (params, body) in macros.items():
        def repl(m, params=params, body=body):
            args = [a.strip() for a in m.group(1).split(",")]
            out = body
            for p, a in zip(params, args):
                out = re.sub(r"##\s*%s\b|\b%s\s*##|\b%s\b" % (p, p, p), a, out)
            return out.replace("##", "")
        text = re.sub(r"\b%s\(([^)]*)\)" % name, repl, text)
    return text


def parse_header(path: Path = HEADER, prefix: str = "uhip_"):
    """Return {name: (restype_str, [argtype_str, ...])} for every prototype named prefix*."""
    text = _expand_macros(_strip_comments(path.read_text()))
    text = re.sub(r"#[^\n]*", " ", text)
    protos = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(%s\w+)\s*\(([^;{]*?)\)\s*;" % prefix, text):
        ret, name, args = m.group(1).strip(), m.group(2), m.group(3).strip()
        ret = ret.split()[-1] if "*" not in ret else ret
        argl = [] if args in ("", "void") else [a.strip() for a in args.split(",")]
        protos[name] = (" ".join(ret.split()), argl)
    return protos


def _ctype(decl: str):
    d = " ".join(decl.replace("const", " ").split())
    if "*" in d:
        base = d.split("*")[0].strip()
        if base == "char" and d.count("*") == 1:
            return ctypes.c_char_p
        return ctypes.c_void_p
    toks = d.split()
    if not toks:
        return None
    if toks[0] in ("uhip_ctx", "uhip_solver", "uhip_bsolver"):
        return ctypes.c_void_p
    if toks[0] in _SCALARS:
        return _SCALARS[toks[0]]
    raise TypeError(f"unmapped C type: {decl!r}")


_LIB = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Load libutopia_hip.so once and attach argtypes/restype from the header."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise LibraryMissing(
            f"{LIB_PATH} is not built; run `python -m utopia_amd.build` (or __graft_entry__.build()).")
    # torch ships its own HIP runtime; when it is loaded first the library binds
    # to that one, so torch tensors and this library share one device runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(str(LIB_PATH))
    protos = {}
    for path, prefix in HEADERS:
        protos.update(parse_header(path, prefix))
    for name, (ret, args) in protos.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = [_ctype(a) for a in args]
        fn.restype = _ctype(ret) if ret != "void" else None
    _LIB = L
    return L


def declared_symbols():
    """Every function declared by the ABI headers (include/utopia_hip.h, include/utopia_hip_isolver.h)."""
    names = set()
    for path, prefix in HEADERS:
        names.update(parse_header(path, prefix).keys())
    return sorted(names)
