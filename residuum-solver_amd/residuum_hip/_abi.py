"""Signatures of libresiduum_hip.so, read from include/residuum_hip.h.

Every exported function gets its ctypes restype and argtypes from the
header's own declaration, so a call with the wrong number of arguments or a
scalar of the wrong width raises instead of passing garbage to the library.
Pointers of every kind are declared as c_void_p (device pointers are plain
addresses; host arrays arrive as ctypes arrays, pointers or byref()), C
strings as c_char_p.

The library reports the SHA-256 of the header it was compiled against
(rsd_abi_hash, the first 16 hex digits); a library whose hash is not that of
the header next to it was built from other declarations and is refused.
"""
import ctypes as C
import hashlib
import os
import re

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))), "include",
                      "residuum_hip.h")

_SCALARS = {
    "int": C.c_int,
    "int32_t": C.c_int32,
    "int64_t": C.c_int64,
    "uint32_t": C.c_uint32,
    "uint64_t": C.c_uint64,
    "size_t": C.c_size_t,
    "double": C.c_double,
    "float": C.c_float,
}


def header_hash(path=HEADER):
    """First 16 hex digits of the header's SHA-256 (the library's build stamp)."""
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def _ctype(decl):
    """ctypes type of one C type (a parameter with its name removed)."""
    t = " ".join(decl.replace("*", " * ").split())
    stars = t.count("*")
    base = t.replace("*", " ").replace("const", " ").split()
    if stars == 1 and base == ["char"]:
        return C.c_char_p
    if stars:
        return C.c_void_p
    if base == ["void"]:
        return None
    if len(base) != 1 or base[0] not in _SCALARS:
        raise ValueError(f"residuum_hip.h: unsupported C type {decl!r}")
    return _SCALARS[base[0]]


def _param_type(p):
    p = p.strip()
    if p in ("", "void"):
        return None
    # drop the parameter name: the last identifier, unless the whole thing is a type
    m = re.match(r"^(.*?[\s*])([A-Za-z_]\w*)$", p)
    return _ctype(m.group(1) if m else p)


def declarations(path=HEADER):
    """{name: (restype, [argtypes])} for every rsd_* function in the header."""
    with open(path) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"^\s*#[^\n]*", " ", text, flags=re.M)
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(rsd_\w+)\s*\(([^()]*)\)\s*;", text):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        ret = " ".join(ret.split())
        if ret.startswith("typedef"):
            continue
        args = [t for t in (_param_type(p) for p in params.split(",")) if t is not None]
        out[name] = (_ctype(ret), args)
    return out


def bind(lib, path=HEADER):
    """Check the library's build stamp against the header and declare every
    function of the header on it; returns the declarations."""
    decls = declarations(path)
    missing = [n for n in decls if not hasattr(lib, n)]
    if missing:
        raise RuntimeError(f"libresiduum_hip.so lacks {len(missing)} declared symbols, e.g. {missing[:3]}: rebuild it")
    lib.rsd_abi_hash.restype = C.c_char_p
    lib.rsd_abi_hash.argtypes = []
    built = lib.rsd_abi_hash().decode()
    want = header_hash(path)
    if built != want:
        raise RuntimeError(f"libresiduum_hip.so was built against another residuum_hip.h (stamp {built}, "
                           f"header {want}): rebuild it")
    for name, (res, args) in decls.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return decls
