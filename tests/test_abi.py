"""The C-ABI library loads and exports every entry point declared in
include/eprem_hip.h, and the ctypes structures match the C layouts.
No device calls (runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from eprem_amd import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "eprem_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(ep_[a-z_0-9]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(abi.EXPORTED_SYMBOLS)


def test_library_exports_every_symbol():
    lib = runtime.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_layouts(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "eprem_hip.h"\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu\\n\","
                   "sizeof(ep_dims),sizeof(ep_params),sizeof(ep_node_soa),sizeof(ep_diag),"
                   "sizeof(ep_mas_params),sizeof(ep_mas_prev),sizeof(ep_mas_out),"
                   "offsetof(ep_diag,stream_sweeps));return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    want = [C.sizeof(abi.EpDims), C.sizeof(abi.EpParams), C.sizeof(abi.EpNodeSoa),
            C.sizeof(abi.EpDiag), C.sizeof(abi.EpMasParams), C.sizeof(abi.EpMasPrev),
            C.sizeof(abi.EpMasOut), abi.EpDiag.stream_sweeps.offset]
    assert got == want


def test_missing_device_is_reported_not_silent():
    """Without a GPU, ep_init fails loudly (no CPU fallback)."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU device node is present")
    with pytest.raises(runtime.EpError) as ei:
        runtime.EpSolver((6, 8, 4, 4), abi.EpParams())
    assert ei.value.code in (abi.EP_ERR_NO_DEVICE, abi.EP_ERR_INVALID_ARG)


def test_missing_library_is_reported_not_silent(monkeypatch, tmp_path):
    """Without the built HIP library the binding raises instead of falling
    back to any CPU path."""
    monkeypatch.setattr(runtime, "_lib", None)
    monkeypatch.setattr(runtime, "LIB_PATH", str(tmp_path / "libeprem_hip.so"))
    with pytest.raises(RuntimeError, match="not built"):
        runtime.lib()
    with pytest.raises(RuntimeError, match="not built"):
        runtime.EpSolver((6, 8, 4, 4), abi.EpParams())
