// Internal helpers shared by the host-side translation units of libtileink_hip.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "tileink_hip.h"

namespace tk {

// Thread-local error text behind tk_last_error().
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

inline tk_status fail(tk_status status, const char *what) {
    set_error("%s", what);
    return status;
}

// Per draw of an encoded scene (4 words each: x0, y0, x1, y1 in its canvas's
// tiles): the rectangle early occlusion hides from the draw's tile bins, empty
// when nothing does (canvas.cpp).  Returns whether any draw is occluded.
bool draw_covers(const tk_scene &s, std::vector<uint32_t> &cover);

// Device address of `device`'s copy of the fixed no-alloc scratch array
// (TK_NO_ALLOC_SCRATCH_BYTES bytes, 256-byte aligned); `device` must be current.
// Taken once per device with hipGetSymbolAddress (scratch.hip).
tk_status scratch_address(int device, void **out);

// Record sizes are part of the ABI; a mismatch is a build error, not a runtime one.
static_assert(sizeof(tk_line) == 24, "Line is 24 bytes (src/shared/line.rs:4-9)");
static_assert(sizeof(tk_path_record) == 76, "PathRecord is 76 bytes (src/shared/path.rs:6-23)");
static_assert(sizeof(tk_draw_record) == 124, "DrawRecord is 124 bytes (draw_records.hlsli:5)");
static_assert(sizeof(tk_line_segment) == 20, "LineSegment is 20 bytes");
static_assert(sizeof(tk_ptcl_record) == 24, "PtclRecord is 24 bytes");
static_assert(sizeof(tk_layer_record) == 12, "LayerStackRecord is 12 bytes");
static_assert(sizeof(tk_scan_chunk) == 16, "GpuScanChunk is 16 bytes");
static_assert(sizeof(tk_scan_config) == 44, "ScanConfig is 11 words");

}  // namespace tk
