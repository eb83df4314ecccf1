"""The C ABI library loads without a GPU and exports every entry point include/*.h declares."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("ust.h", "ust_kernels.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(ust_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from earthbyte_underworld_amd import _lib
    _lib.load()
    raw = C.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert len(declared) > 60
    missing = [n for n in sorted(declared) if not hasattr(raw, n)]
    assert not missing, missing
    # the Python binding covers only declared symbols
    assert set(_lib.exported_symbols()) <= declared


def test_rheology_xml_names_and_defaults():
    import earthbyte_underworld_amd as ust
    r = ust.Arrhenius(eta0=1e-6, activationEnergy=27.63102112)
    assert r.type == 2 and list(r.p)[:4] == [1e-6, 27.63102112, 0.0, 1.0]
    lim = ust.ViscosityLimiter()
    assert list(lim.p)[:2] == [1e4, 1e-2]  # ViscosityLimiter.c:102-103 defaults
    dp = ust.DruckerPrager(cohesion=1e3, frictionCoefficient=0.6)
    assert dp.type == 5 and dp.p[0] == 1e3 and dp.p[1] == 0.6
    import pytest
    with pytest.raises(ust.UstError):
        ust.VonMises(frictionCoefficient=0.3)  # not a VonMises key
    with pytest.raises(ust.UstError):
        ust.stokes._rh("FrankKamenetskii", eta0=1.0)  # not on this path


def test_no_torch_types_in_abi():
    for h in ("ust.h", "ust_kernels.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        assert "torch" not in src.lower() and "at::" not in src


def _prototypes():
    """name -> (return type, [parameter types]) of every ust_* declaration in include/*.h."""
    protos = {}
    for h in ("ust.h", "ust_kernels.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
        src = re.sub(r"//[^\n]*", " ", src)
        src = re.sub(r"^\s*#[^\n]*", " ", src, flags=re.M)
        for m in re.finditer(r"([\w\s\*]+?)\b(ust_\w+)\s*\(([^()]*)\)\s*;", src):
            ret, name, params = m.group(1).strip(), m.group(2), m.group(3).strip()
            ret = re.sub(r"\s+", " ", ret.split(";")[-1].split("}")[-1]).strip()
            plist = [] if params in ("", "void") else [re.sub(r"\s+", " ", p.strip()) for p in params.split(",")]
            protos[name] = (ret, plist)
    return protos


def _is_pointer_ctype(t):
    return t in (C.c_void_p, C.c_char_p) or (isinstance(t, type) and issubclass(t, C._Pointer))


_SCALARS = {"int": C.c_int, "double": C.c_double, "float": C.c_float, "long": C.c_long, "size_t": C.c_size_t,
            "unsigned long long": C.c_ulonglong}


def _check_type(cdecl, ctype, is_return=False):
    """The ctypes type agrees with the C type: pointers (and arrays) as pointers, scalars of the same width."""
    pointer = "*" in cdecl or "[" in cdecl
    if pointer:
        return _is_pointer_ctype(ctype)
    base = re.sub(r"\b(const|struct)\b", "", cdecl).strip()
    if is_return:
        return (ctype is None) if base == "void" else _SCALARS.get(base) is ctype
    words = base.split()
    scalar = " ".join(words[:-1])  # drop the parameter name
    return _SCALARS.get(scalar) is ctype


def test_binding_signatures_match_headers():
    """Every signature in the ctypes table agrees with its C prototype: the same arity, every pointer (and
    array) parameter and pointer return declared as a pointer type, every scalar at its C width.  An
    undeclared or int-typed pointer is what truncates 64-bit device and stream pointers to 32 bits."""
    from earthbyte_underworld_amd import _lib
    protos = _prototypes()
    bad = []
    for name, (res, args) in _lib._SIGS.items():
        assert name in protos, f"{name} is bound but not declared in include/*.h"
        ret, params = protos[name]
        if len(params) != len(args):
            bad.append(f"{name}: {len(args)} argtypes for {len(params)} parameters")
            continue
        if not _check_type(ret, res, is_return=True):
            bad.append(f"{name}: return {ret!r} bound as {res}")
        for p, a in zip(params, args):
            if not _check_type(p, a):
                bad.append(f"{name}: parameter {p!r} bound as {a}")
    assert not bad, bad


def test_undeclared_entry_points_are_refused():
    """Names the table does not declare are not reachable through the binding (ctypes would default their
    arguments and result to 32-bit int)."""
    import pytest
    from earthbyte_underworld_amd import _lib
    lib = _lib.load()
    raw = C.CDLL(_lib.LIB_PATH)
    undeclared = sorted(n for n in _prototypes() if n not in _lib._SIGS and hasattr(raw, n))
    assert undeclared, "every entry point is bound: pick another refusal probe"
    for n in undeclared[:5]:
        with pytest.raises(AttributeError, match="no signature declared"):
            getattr(lib, n)
