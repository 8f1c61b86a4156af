"""The C-ABI library: it loads on a machine without a GPU, exports every entry point that
include/graphflow.h declares, and its enums and structs agree with the Python binding and the oracle."""
import ctypes as C
import os
import re

import pytest

import graphflow_amd as g
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "graphflow.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:gf_status|const char\*|uint32_t)\s+(gf_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = g.lib()
    names = _declared()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    # the binding covers exactly the declared surface
    assert sorted(g.SIGNATURES) == names


def test_header_enums_match_binding_and_oracle():
    text = open(HEADER).read()
    enum = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"GF_(\w+)\s*=\s*(\d+)", text))
    for name, value in g.OPS.items():
        assert enum[name] == value, name
        assert O.OPS[name] == value, name
    for name, value in g.UPDS.items():
        assert enum["UPD_" + name] == value
        assert O.UPDS[name] == value
    assert enum["PAYLOAD_PARTICLE"] == g.PAYLOAD_PARTICLE and enum["PAYLOAD_PHOTON"] == g.PAYLOAD_PHOTON
    assert enum["SCHEDULE_DEVICE"] == g.SCHEDULE_DEVICE


def test_struct_layouts():
    text = open(HEADER).read()
    body = re.search(r"typedef struct gf_tuning \{(.*?)\} gf_tuning;", text, re.S).group(1)
    fields = re.findall(r"\b(uint64_t|uint32_t|double)\s+([\w, ]+);", body)
    names = [n.strip() for _, group in fields for n in group.split(",")]
    assert names == [f for f, _ in g.TuningC._fields_]
    assert C.sizeof(g.NodeDesc) == 13 * 4 + 4 + 32  # 13 u32/f32 words, 4 B padding, ptr[2] and size[2]
    assert C.sizeof(g.ChunkDesc) == 16


def test_host_plan_helpers_match_oracle():
    for n, c, b in [(0, 128, 2), (4097, 128, 2), (17, 127, 2), (64, 128, 64), (10**9, 8192, 4096)]:
        assert g.output_chunk_bound(n, c, b) == O.output_chunk_bound(n, c, b)[1]
    with pytest.raises(ValueError):
        g.output_chunk_bound(4, 0, 2)
    with pytest.raises(g.BatchError):
        g.output_chunk_bound(2**64 - 1, 1, 2)


def test_status_names():
    lib = g.lib()
    assert lib.gf_status_name(0) == b"ok"
    assert lib.gf_status_name(g.STALL) == b"stall"
    assert lib.gf_abi_version() == 1
