"""The C-ABI library loads and exports every symbol declared in include/fermihip.h (no device compute)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fermihip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    import fermihip
    assert names == sorted(fermihip.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    import fermihip
    path = fermihip.lib_path()
    if not os.path.exists(path):
        pytest.fail("libfermihip.so is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_default_options_mirror_fermi():
    import fermihip
    L = fermihip.load_library()
    o = fermihip._Opts()
    L.fh_default_opts(ctypes.byref(o))
    assert (o.cc_max_iter, o.cc_e_conv, o.cc_max_rms, o.cc_diis, o.cc_ndiis, o.diis_start, o.cc_diis_relax,
            o.diis_prec, o.cc_damp_ratio, o.precision_override) == (50, 1e-10, 1e-10, 1, 3, 3, 3, 1, 0.0, 0)


def test_missing_library_fails_loudly(monkeypatch):
    import fermihip
    monkeypatch.setattr(fermihip, "_LIB", None)
    monkeypatch.setenv("FERMIHIP_LIB", "/nonexistent/libfermihip.so")
    with pytest.raises(fermihip.FermiException):
        fermihip.load_library()


def test_host_option_validation():
    import fermihip
    with pytest.raises(fermihip.FermiException):
        fermihip.Options(no_such_key=1)
    with pytest.raises(fermihip.FermiException):
        fermihip.Options(diis_prec="half")
