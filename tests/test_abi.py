"""The C ABI library (CPU): it loads, exports every entry point include/tileink_hip.h
declares, and its record layouts match the reference's GPU ABI."""
import ctypes as C
import os
import re

import pytest

import tileink_amd as tk
from tileink_amd import _abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "tileink_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = C.CDLL(_abi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_functions()) == set(_abi.SIGNATURES)


def test_record_layouts_match_reference_abi():
    assert C.sizeof(_abi.Line) == 24
    assert C.sizeof(_abi.PathRecord) == 76
    assert _abi.PathRecord.segment_start.offset == 40
    assert _abi.PathRecord.transform.offset == 52
    assert C.sizeof(_abi.DrawRecord) == 124
    assert _abi.DrawRecord.tag.offset == 32 and _abi.DrawRecord.fill_rule.offset == 36
    assert _abi.DrawRecord.pixel_bounds.offset == 40 and _abi.DrawRecord.solid_rect.offset == 72
    assert _abi.DrawRecord.transform.offset == 76 and _abi.DrawRecord.inverse_transform.offset == 100
    assert _abi.SEGMENT_DTYPE.itemsize == 20 and _abi.PTCL_DTYPE.itemsize == 24
    assert C.sizeof(_abi.ScanConfig) == 44


def header_struct_fields():
    """Field names of every `typedef struct NAME {...} NAME;` in the header, in order."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, body in re.findall(r"typedef struct (tk_\w+) \{(.*?)\} \1;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = re.sub(r"\[[^\]]*\]", "", decl).strip()
            if not decl:
                continue
            head, *rest = decl.split(",")
            fields.append(head.split()[-1].lstrip("*"))
            fields += [r.strip().lstrip("*") for r in rest]
        out[name] = fields
    return out


MIRRORS = {"tk_scan_config": "ScanConfig", "tk_cumsum_config": "CumsumConfig",
           "tk_scene": "Scene", "tk_limits": "Limits", "tk_frame_stats": "FrameStats", "tk_stage_times": "StageTimes",
           "tk_coarse_config": "CoarseConfig", "tk_coarse_inputs": "CoarseInputs",
           "tk_coarse_outputs": "CoarseOutputs", "tk_fine_config": "FineConfig", "tk_fine_inputs": "FineInputs",
           "tk_brush": "Brush", "tk_gradient_stop": "GradientStop", "tk_tuning": "Tuning",
           "tk_stroke_style": "StrokeStyle"}


@pytest.mark.parametrize("struct", sorted(k for k, v in MIRRORS.items() if v))
def test_python_mirrors_follow_header_field_order(struct):
    """A ctypes mirror with fields in another order than the header passes wrong
    pointers through the ABI; the names must match position by position."""
    fields = header_struct_fields()[struct]
    mirror = [n for n, *_ in getattr(_abi, MIRRORS[struct])._fields_]
    assert mirror == fields, (struct, mirror, fields)


def test_missing_device_is_an_error_not_a_fallback():
    with pytest.raises(tk.TileinkError) as err:
        tk.NativeContext(-1)
    assert err.value.status == _abi.TK_ERR_UNAVAILABLE


def test_renderer_rejects_other_backends():
    class Vulkan:
        pass
    with pytest.raises(tk.TileinkError):
        tk.NativeRenderer(Vulkan, 16, 16)


def test_version_and_error_string():
    assert _abi.lib().tk_version() >= 1
    with pytest.raises(tk.TileinkError):
        tk.Canvas(16, 16).pop_layer()
    assert b"layer" in _abi.lib().tk_last_error()


def test_no_alloc_limits_match_the_header():
    text = open(HEADER).read()
    scratch = re.search(r"#define TK_NO_ALLOC_SCRATCH_BYTES \(1u << (\d+)\)", text)
    table = re.search(r"#define TK_NO_ALLOC_MAX_TABLE (\d+)ull", text)
    ticket = re.search(r"#define TK_NO_ALLOC_TICKET_BYTES (\d+)u", text)
    assert scratch and table and ticket
    assert _abi.TK_NO_ALLOC_SCRATCH_BYTES == 1 << int(scratch.group(1))
    assert _abi.TK_NO_ALLOC_MAX_TABLE == int(table.group(1))
    # one u64 lookback status word per coarse scan workgroup of 4096 counts
    assert _abi.TK_NO_ALLOC_MAX_TABLE == (_abi.TK_NO_ALLOC_SCRATCH_BYTES - int(ticket.group(1))) // 8 * 4096
