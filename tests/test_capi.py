"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from wslda_amd import td

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|size_t|void|double|tdw_ctx)\s*\**\s*([a-zA-Z_]\w*)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(td.LIBPATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert declared_functions() == set(td.EXPORTS)


def test_params_layout_matches_oracle_cfg():
    from oracle import pyoracle as po
    # the library's tdw_params starts with the oracle's cfg layout field by field
    for (a, _), (b, _) in zip(po.Cfg._fields_, td.Params._fields_):
        assert a == b
        assert getattr(po.Cfg, a).offset == getattr(td.Params, b).offset
