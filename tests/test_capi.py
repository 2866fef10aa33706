"""The C-ABI library loads and exports every entry point include/*.h declares (no GPU calls)."""
import ctypes
import glob
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(tdag_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import tdag_amd

    lib = ctypes.CDLL(tdag_amd.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 7
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


def test_version_and_error_without_gpu_calls():
    import tdag_amd

    assert "gfx950" in tdag_amd.version()
    assert isinstance(tdag_amd.last_error(), str)


def test_config_layout_matches_reference_struct():
    import tdag_amd

    # TransformDAGConfig: 3 bools + 1 padding byte + uint32 (transform_dag.h:157-160)
    assert ctypes.sizeof(tdag_amd.TdagConfig) == 8
    assert tdag_amd.TdagConfig.max_translation_level.offset == 4


def test_tp_helpers_closed_forms():
    import tdag_amd

    assert tdag_amd.TP_SENTINEL == 0x82082400FFFFFFFF
    assert tdag_amd.tp_pack(5) == 0x8208240000000005
    u = tdag_amd.tp_unpack(tdag_amd.tp_pack(9, 5, (2, 0, 1), (-3, 0, 31)))
    assert u == {"ptr": 9, "sym": 5, "perm": (2, 0, 1), "t": (-3, 0, 31)}


def test_cpp_dropins_compile_and_link(tmp_path):
    """include/tdag.hpp's typed drop-ins, called as the reference calls its own entry points on
    stand-in types with the reference's layouts (tests/cpp/dropin_check.cpp), compile as C++20
    and link against libtdag.so. The program is run by tests/test_gpu_dropin.py."""
    import dropin_build

    assert os.path.exists(dropin_build.build(str(tmp_path)))
