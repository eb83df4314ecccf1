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
