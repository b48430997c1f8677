"""CPU checks of the C ABI boundary: the library loads, exports every entry
point include/uhann.h declares, and reports "no device" as the reference's
runtime_error when no GPU is visible (its CUDA-disabled stubs throw
std::runtime_error, e.g. src/core/cuda_dense_l2_scan_index_stub.cpp)."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(uhann_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from ultrahigh_ann_amd import _native

    lib = _native.lib()
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "uhann.h"\nint main(void){return UHANN_OK;}\n')
    rc = os.system(f"gcc -std=c99 -Wall -Werror -I{ROOT}/include -c {src} -o {tmp_path}/t.o")
    assert rc == 0


def test_status_codes_match_python_exceptions():
    from ultrahigh_ann_amd import _native

    text = open(os.path.join(ROOT, "include", "uhann.h")).read()
    codes = dict((k, int(v)) for k, v in re.findall(r"#define (UHANN_ERR_[A-Z_]+) (\d+)", text))
    assert _native._STATUS[codes["UHANN_ERR_INVALID_ARGUMENT"]] is _native.UhannInvalidArgument
    assert _native._STATUS[codes["UHANN_ERR_RUNTIME"]] is _native.UhannError
    assert _native._STATUS[codes["UHANN_ERR_OVERFLOW"]] is _native.UhannOverflowError
    assert _native._STATUS[codes["UHANN_ERR_LENGTH"]] is _native.UhannLengthError


def test_no_device_raises_runtime_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import ultrahigh_ann_amd as uh

    assert not uh.device_available()
    with pytest.raises(uh.UhannError):
        uh.HipExactL2Index(np.zeros((2, 2), np.float32))


def test_null_handles_are_rejected():
    from ultrahigh_ann_amd import _native

    lib = _native.lib()
    out = ctypes.c_uint64()
    assert lib.uhann_scan_info(None, ctypes.byref(out), None, None, None) == 1
    assert lib.uhann_workspace_info(None, ctypes.byref(out), None) == 1
    assert b"null" in lib.uhann_last_error()


def test_struct_layouts_match_ctypes(tmp_path):
    """The ctypes mirrors of the ABI's structs have the C compiler's sizes and offsets."""
    from ultrahigh_ann_amd import _native

    structs = {
        "uhann_server_config": _native.ServerConfig,
        "uhann_index_spec": _native.IndexSpec,
        "uhann_serving_stats": _native.ServingStats,
        "uhann_trace_entry": _native.TraceEntry,
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "uhann.h"', "int main(void){"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    assert os.system(f"gcc -std=c99 -I{ROOT}/include {src} -o {exe}") == 0
    out = os.popen(str(exe)).read().split("\n")
    got = {}
    for line in out:
        if line:
            name, field, value = line.split()
            got[(name, field)] = int(value)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)
