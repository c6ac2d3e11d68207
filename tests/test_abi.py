"""The C-ABI library loads and exports every entry point include/evora_hip.h declares."""
import ctypes
import os
import re

from evora_amd import hydro

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "evora_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|long long)\s+(evh_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_python_exports():
    assert declared_symbols() == sorted(hydro.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(hydro.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_config_struct_matches_header():
    text = open(os.path.join(ROOT, "include", "evora_hip.h")).read()
    body = text[text.index("typedef struct evh_config {"):text.index("} evh_config;")]
    names = re.findall(r"\b(\w+)\s*[,;]", re.sub(r"/\*.*?\*/", "", body, flags=re.S))
    assert names == [f[0] for f in hydro.EvhConfig._fields_]


def test_library_is_built_for_gfx950():
    data = open(hydro.LIB_PATH, "rb").read()
    assert b"gfx950" in data
