// line_scanned_tile_count: how many tiles of a path's tile box the scan DDA
// visits for one line, with Rust f32 semantics (src/shared/scan_line.rs:17-151):
// round() is half-away-from-zero and float-to-int casts saturate.  Shared by the
// host recorder (canvas.cpp) and the device front end (frontend.hip), which must
// give the same segment capacities for the same lines; both are built without
// floating-point contraction.
#pragma once

#include <cmath>
#include <cstdint>

#include "tileink_hip.h"

#if defined(__HIPCC__)
#define TK_HD __host__ __device__
#else
#define TK_HD
#endif

namespace tk {

TK_HD inline uint32_t rs_u32(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967295.0f) return 0xFFFFFFFFu;
    return uint32_t(v);
}
TK_HD inline int32_t rs_i32(float v) {
    if (v != v) return 0;
    if (v <= -2147483648.0f) return INT32_MIN;
    if (v >= 2147483647.0f) return INT32_MAX;
    return int32_t(v);
}
TK_HD inline uint32_t rs_span(float a, float b) {
    float hi = std::ceil(std::fmax(a, b));
    float lo = std::floor(std::fmin(a, b));
    return rs_u32(std::fmax(hi - lo, 1.0f));
}

TK_HD inline uint32_t scanned_tile_count(float p0x, float p0y, float p1x, float p1y, uint32_t bx0, uint32_t by0,
                                         uint32_t bx1, uint32_t by1) {
    const float scale = 1.0f / 16.0f;
    const float top_touch_epsilon = 1.0e-12f;
    bool is_down = p1y >= p0y;
    float xy0x = is_down ? p0x : p1x, xy0y = is_down ? p0y : p1y;
    float xy1x = is_down ? p1x : p0x, xy1y = is_down ? p1y : p0y;
    float s0x = xy0x * scale, s0y = xy0y * scale, s1x = xy1x * scale, s1y = xy1y * scale;
    uint32_t count_x = rs_span(s0x, s1x) - 1u;
    uint32_t count = count_x + rs_span(s0y, s1y);
    float dx = std::fabs(s1x - s0x);
    float dy = s1y - s0y;
    if (dx + dy == 0.0f) return 0;
    if (dy == 0.0f && std::floor(s0y) == s0y) return 0;
    float idxdy = 1.0f / (dx + dy);
    float a = dx * idxdy;
    bool is_positive_slope = s1x >= s0x;
    float sign = is_positive_slope ? 1.0f : -1.0f;
    float xt0 = std::floor(s0x * sign);
    float c = s0x * sign - xt0;
    float y0 = std::floor(s0y);
    float ytop = s0y == s1y ? std::ceil(s0y) : y0 + 1.0f;
    float b = std::fmin((dy * c + dx * (ytop - s0y)) * idxdy, 0.99999994f);
    float robust_err = std::floor(a * (float(count) - 1.0f) + b) - float(count_x);
    if (robust_err != 0.0f) a -= std::copysign(2e-7f, robust_err);
    float x0 = xt0 * sign + (is_positive_slope ? 0.0f : -1.0f);
    float xmin = std::fmin(s0x, s1x);
    if (s0y >= float(by1) || s1y <= float(by0) + top_touch_epsilon || xmin >= float(bx1)) return 0;
    uint32_t imin = 0;
    if (s0y < float(by0)) {
        float iminf = std::round((float(by0) - y0 + b - a) / (1.0f - a)) - 1.0f;
        if (y0 + iminf - std::floor(a * iminf + b) < float(by0)) iminf += 1.0f;
        imin = rs_u32(iminf);
    }
    uint32_t imax = count;
    if (s1y > float(by1)) {
        float imaxf = std::round((float(by1) - y0 + b - a) / (1.0f - a)) - 1.0f;
        if (y0 + imaxf - std::floor(a * imaxf + b) < float(by1)) imaxf += 1.0f;
        imax = rs_u32(imaxf);
    }
    if (std::fmax(s0x, s1x) < float(bx0)) {
        imax = imin;
    } else {
        float fudge = is_positive_slope ? 0.0f : 1.0f;
        if (xmin < float(bx0)) {
            float f = std::round((sign * (float(bx0) - x0) - b + fudge) / a);
            if ((x0 + sign * std::floor(a * f + b) < float(bx0)) == is_positive_slope) f += 1.0f;
            if (is_positive_slope) {
                if (rs_u32(f) > imin) imin = rs_u32(f);
            } else if (rs_u32(f) < imax) {
                imax = rs_u32(f);
            }
        }
        if (std::fmax(s0x, s1x) > float(bx1)) {
            float f = std::round((sign * (float(bx1) - x0) - b + fudge) / a);
            if ((x0 + sign * std::floor(a * f + b) < float(bx1)) == is_positive_slope) f += 1.0f;
            if (is_positive_slope)
                imax = imax < rs_u32(f) ? imax : rs_u32(f);
            else
                imin = imin > rs_u32(f) ? imin : rs_u32(f);
        }
    }
    imax = imax > imin ? imax : imin;
    uint32_t total = 0;
    for (uint32_t i = imin; i < imax; ++i) {
        float z = std::floor(a * float(i) + b);
        int32_t y = rs_i32(y0 + float(i) - z);
        int32_t x = rs_i32(x0 + sign * z);
        total += (y >= int32_t(by0) && y < int32_t(by1) && x >= int32_t(bx0) && x < int32_t(bx1)) ? 1u : 0u;
    }
    return total;
}

// PixelBounds::tile_bbox (src/shared/bounds.rs:113-128): the tiles a pixel box
// touches, clamped to the canvas's tile grid.
TK_HD inline void tile_box_of(int32_t px0, int32_t py0, int32_t px1, int32_t py1, uint32_t tw, uint32_t th,
                              uint32_t out[4]) {
    const uint32_t x0 = uint32_t(px0 > 0 ? px0 : 0), y0 = uint32_t(py0 > 0 ? py0 : 0);
    const uint32_t x1 = uint32_t(px1 > 0 ? px1 : 0), y1 = uint32_t(py1 > 0 ? py1 : 0);
    const uint32_t t = TK_TILE_SIZE;
    out[0] = x0 / t < tw ? x0 / t : tw;
    out[1] = y0 / t < th ? y0 / t : th;
    const uint64_t cx = (uint64_t(x1) + t - 1) / t, cy = (uint64_t(y1) + t - 1) / t;
    out[2] = cx < tw ? uint32_t(cx) : tw;
    out[3] = cy < th ? uint32_t(cy) : th;
}

}  // namespace tk
