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


def test_cpp_dropin_header_compiles(tmp_path):
    """include/tdag.hpp (the C++ drop-ins over the C-ABI) compiles as C++20."""
    import subprocess

    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "tdag.hpp"\n'
        "int main() {\n"
        "  tdag_config cfg = tdag_cpp::toC(true, true, true, 4);\n"
        "  std::vector<uint32_t> items(2); std::vector<uint64_t> grids(2), mapping;\n"
        "  auto u = tdag_cpp::findDuplicatesAmongFlattenedItems_gpu<2, uint32_t>(items, grids, mapping, cfg);\n"
        "  return (int)u.size();\n"
        "}\n")
    proc = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-I", os.path.join(REPO, "include"), str(src)], capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
