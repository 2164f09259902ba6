// Device-side arithmetic of the compute-raster stages for gfx950.
//
// Every function restates one reference shader helper; the arithmetic order is
// kept exactly and HLSL `mad` becomes fmaf, everything else is compiled with
// -ffp-contract=off so no other product is fused.  HLSL round() lowers to
// round-half-to-even (rintf); float->int conversions saturate and map NaN to 0
// (v_cvt_u32_f32 / v_cvt_i32_f32 semantics).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tileink_hip.h"

#define TKD static __device__ __forceinline__

namespace tkd {

constexpr float kScanEpsilon = 1.0e-6f;      // scan/geometry.hlsli:8
constexpr float kTopTouchEpsilon = 1.0e-12f; // scan/geometry.hlsli:9
constexpr float kTileClipNudge = 1.0e-3f;    // scan/geometry.hlsli:10
constexpr float kTileSize = 16.0f;
constexpr float kTileScale = 1.0f / 16.0f;
constexpr float kChannelScale = 1.0f / 255.0f;       // shared/pixel.hlsli:5
constexpr float kGradientEpsilon = 0.00000011920929f; // shared/brush/ramp.hlsli:6
constexpr uint32_t kNone = TK_INVALID_INDEX;

TKD uint32_t f2u(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}
TKD int32_t f2i(float v) {
    if (!(v == v)) return 0;
    if (v <= -2147483648.0f) return (int32_t)0x80000000u;
    if (v >= 2147483648.0f) return 0x7FFFFFFF;
    return (int32_t)v;
}
TKD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
TKD int clampi(int x, int lo, int hi) { return min(max(x, lo), hi); }
// clampi(f2i(v), 0, 16) for every v (NaN included, as fmaxf(NaN, 0) = 0) without
// f2i's range branches: v is clamped in FP32 first, so the conversion is exact.
TKD int clamp_cell16(float v) { return (int)fminf(fmaxf(v, 0.0f), 16.0f); }
TKD float asfloat(uint32_t u) { return __uint_as_float(u); }
TKD uint32_t asuint(float f) { return __float_as_uint(f); }

// ---------------------------------------------------------------------------
// Scan geometry (scan/geometry.hlsli:12-114)
// ---------------------------------------------------------------------------

struct ScanTraversal {
    uint32_t bx, by, bz, bw;  // tile bbox x0, y0, x1, y1
    uint32_t data_offset;
    float px0, py0, px1, py1;  // oriented so that py1 >= py0
    bool down;
    uint32_t count;
    float a, b, sign, x0, y0;
    uint32_t imin, imax;
    int32_t ymin, ymax;
};

TKD bool is_tile_boundary_y(float v) { return fabsf(v - floorf(v)) <= kTopTouchEpsilon; }
TKD int32_t ceil_tile_boundary_y(float v) { return is_tile_boundary_y(v) ? f2i(floorf(v)) : f2i(ceilf(v)); }
TKD uint32_t scan_span(float a, float b) {
    float hi = ceilf(a);
    if (b > a) hi = ceilf(b);
    float lo = floorf(a);
    if (b < a) lo = floorf(b);
    float value = hi - lo;
    if (value < 1.0f) value = 1.0f;
    return f2u(value);
}

TKD bool scan_geometry(const tk_line *lines, const tk_path_record *paths, uint32_t path_count,
                       uint32_t line_index, ScanTraversal &s) {
    s = ScanTraversal{};
    const tk_line line = lines[line_index];
    uint32_t path_id = line.path_id;
    if (path_id >= path_count) return false;
    const tk_path_record &path = paths[path_id];
    s.bx = path.tile_x0;
    s.by = path.tile_y0;
    s.bz = path.tile_x1;
    s.bw = path.tile_y1;
    s.data_offset = path.data_offset;
    if (s.bz - s.bx == 0u || s.by >= s.bw) return false;
    const tk_affine t = path.transform;
    float p0x = t.a * line.p0[0] + t.c * line.p0[1] + t.e;
    float p0y = t.b * line.p0[0] + t.d * line.p0[1] + t.f;
    float p1x = t.a * line.p1[0] + t.c * line.p1[1] + t.e;
    float p1y = t.b * line.p1[0] + t.d * line.p1[1] + t.f;
    s.down = p1y >= p0y;
    if (s.down) {
        s.px0 = p0x; s.py0 = p0y; s.px1 = p1x; s.py1 = p1y;
    } else {
        s.px0 = p1x; s.py0 = p1y; s.px1 = p0x; s.py1 = p0y;
    }
    float sx = s.px0 * kTileScale, sy = s.py0 * kTileScale;
    float sz = s.px1 * kTileScale, sw = s.py1 * kTileScale;
    uint32_t count_x = scan_span(sx, sz) - 1u;
    s.count = count_x + scan_span(sy, sw);
    float dx = fabsf(sz - sx);
    float dy = sw - sy;
    if (dx + dy == 0.0f || (dy == 0.0f && floorf(sy) == sy)) return false;
    float reciprocal = 1.0f / (dx + dy);
    s.a = dx * reciprocal;
    bool positive = sz >= sx;
    s.sign = positive ? 1.0f : -1.0f;
    float xt0 = floorf(sx * s.sign);
    float c = sx * s.sign - xt0;
    s.y0 = floorf(sy);
    float ytop = s.y0 + 1.0f;
    if (sy == sw) ytop = ceilf(sy);
    s.b = fminf((dy * c + dx * (ytop - sy)) * reciprocal, 0.99999994f);
    float error = floorf(s.a * ((float)s.count - 1.0f) + s.b) - (float)count_x;
    if (error != 0.0f) {
        if (error > 0.0f) s.a -= 0.0000002f;
        else s.a += 0.0000002f;
    }
    s.x0 = positive ? xt0 * s.sign : xt0 * s.sign - 1.0f;
    float xmin = fminf(sx, sz);
    if (sy >= (float)s.bw || sw <= (float)s.by + kTopTouchEpsilon || xmin >= (float)s.bz) return false;
    s.imin = 0u;
    if (sy < (float)s.by) {
        float first = rintf(((float)s.by - s.y0 + s.b - s.a) / (1.0f - s.a)) - 1.0f;
        if (s.y0 + first - floorf(s.a * first + s.b) < (float)s.by) first += 1.0f;
        s.imin = f2u(first);
    }
    s.imax = s.count;
    if (sw > (float)s.bw) {
        float last = rintf(((float)s.bw - s.y0 + s.b - s.a) / (1.0f - s.a)) - 1.0f;
        if (s.y0 + last - floorf(s.a * last + s.b) < (float)s.bw) last += 1.0f;
        s.imax = f2u(last);
    }
    s.ymin = 0;
    s.ymax = 0;
    if (fmaxf(sx, sz) < (float)s.bx) {
        s.ymin = ceil_tile_boundary_y(sy);
        s.ymax = ceil_tile_boundary_y(sw);
        s.imax = s.imin;
    } else {
        float fudge = positive ? 0.0f : 1.0f;
        if (xmin < (float)s.bx) {
            float cut = rintf((s.sign * ((float)s.bx - s.x0) - s.b + fudge) / s.a);
            if ((s.x0 + s.sign * floorf(s.a * cut + s.b) < (float)s.bx) == positive) cut += 1.0f;
            int32_t ynext = f2i(s.y0 + cut - floorf(s.a * cut + s.b) + 1.0f);
            if (positive) {
                if (f2u(cut) > s.imin) {
                    float ystart = s.y0 + 1.0f;
                    if (is_tile_boundary_y(sy)) ystart = s.y0;
                    s.ymin = f2i(ystart);
                    s.ymax = ynext;
                    s.imin = f2u(cut);
                }
            } else if (f2u(cut) < s.imax) {
                s.ymin = ynext;
                s.ymax = ceil_tile_boundary_y(sw);
                s.imax = f2u(cut);
            }
        }
        if (fmaxf(sx, sz) > (float)s.bz) {
            float cut = rintf((s.sign * ((float)s.bz - s.x0) - s.b + fudge) / s.a);
            if ((s.x0 + s.sign * floorf(s.a * cut + s.b) < (float)s.bz) == positive) cut += 1.0f;
            if (positive) s.imax = min(s.imax, f2u(cut));
            else s.imin = max(s.imin, f2u(cut));
        }
    }
    s.imax = max(s.imin, s.imax);
    s.ymin = max(s.ymin, (int32_t)s.by);
    s.ymax = min(s.ymax, (int32_t)s.bw);
    return true;
}

TKD void scan_tile(const ScanTraversal &s, uint32_t index, float z, int32_t &tx, int32_t &ty) {
    tx = f2i(s.x0 + s.sign * z);
    ty = f2i(s.y0 + (float)index - z);
}
TKD bool scan_tile_inside(const ScanTraversal &s, int32_t tx, int32_t ty) {
    return ty >= (int32_t)s.by && ty < (int32_t)s.bw && tx >= (int32_t)s.bx && tx < (int32_t)s.bz;
}
TKD uint32_t scan_tile_offset(const ScanTraversal &s, int32_t tx, int32_t ty) {
    return s.data_offset + ((uint32_t)ty - s.by) * (s.bz - s.bx) + (uint32_t)tx - s.bx;
}

// ---------------------------------------------------------------------------
// Tile-clipped segment (scan/clip.hlsli:10-75)
// ---------------------------------------------------------------------------

TKD float x_at_y(float px0, float py0, float px1, float py1, float y) {
    return fmaf(px1 - px0, (y - py0) / (py1 - py0), px0);
}
TKD float y_at_x(float px0, float py0, float px1, float py1, float x) {
    return fmaf(py1 - py0, (x - px0) / (px1 - px0), py0);
}
TKD float clip_y_at_x(const ScanTraversal &s, float x, float tmin, float tmax) {
    float y = y_at_x(s.px0, s.py0, s.px1, s.py1, x);
    if (y <= tmin + kScanEpsilon) {
        float top_x = x_at_y(s.px0, s.py0, s.px1, s.py1, tmin);
        if (fabsf(top_x - x) > kTileClipNudge) return tmin;
        return tmin + kTileClipNudge;
    }
    return clampf(y, tmin + kTileClipNudge, tmax);
}

TKD tk_line_segment clipped_segment(const ScanTraversal &s, uint32_t sub_index, float z, int32_t tx, int32_t ty) {
    const float tile_size = kTileSize;
    float tminx = (float)tx * tile_size, tminy = (float)ty * tile_size;
    float tmaxx = tminx + tile_size, tmaxy = tminy + tile_size;
    float qx0 = s.px0, qy0 = s.py0, qx1 = s.px1, qy1 = s.py1;
    if (sub_index > 0u) {
        float previous = floorf(s.a * ((float)sub_index - 1.0f) + s.b);
        if (z == previous) {
            float x = x_at_y(s.px0, s.py0, s.px1, s.py1, tminy);
            x = clampf(x, tminx + kTileClipNudge, tmaxx);
            qx0 = x;
            qy0 = tminy;
        } else {
            float x_clip = s.sign > 0.0f ? tminx : tmaxx;
            qx0 = x_clip;
            qy0 = clip_y_at_x(s, x_clip, tminy, tmaxy);
        }
    }
    if (sub_index < s.count - 1u) {
        float next = floorf(s.a * ((float)sub_index + 1.0f) + s.b);
        if (z == next) {
            float x = x_at_y(s.px0, s.py0, s.px1, s.py1, tmaxy);
            x = clampf(x, tminx + kTileClipNudge, tmaxx);
            qx1 = x;
            qy1 = tmaxy;
        } else {
            float x_clip = s.sign > 0.0f ? tmaxx : tminx;
            qx1 = x_clip;
            qy1 = clip_y_at_x(s, x_clip, tminy, tmaxy);
        }
    }
    float edge = 1000000000.0f;
    float p0x = clampf(qx0 - tminx, 0.0f, tile_size);
    float p0y = clampf(qy0 - tminy, 0.0f, tile_size);
    float p1x = clampf(qx1 - tminx, 0.0f, tile_size);
    float p1y = clampf(qy1 - tminy, 0.0f, tile_size);
    if (p0x <= kScanEpsilon) p0x = 0.0f; else if (tile_size - p0x <= kScanEpsilon) p0x = tile_size;
    if (p0y <= kScanEpsilon) p0y = 0.0f; else if (tile_size - p0y <= kScanEpsilon) p0y = tile_size;
    if (p1x <= kScanEpsilon) p1x = 0.0f; else if (tile_size - p1x <= kScanEpsilon) p1x = tile_size;
    if (p1y <= kScanEpsilon) p1y = 0.0f; else if (tile_size - p1y <= kScanEpsilon) p1y = tile_size;
    if (p0x == 0.0f) {
        if (p1x == 0.0f) {
            p0x = kScanEpsilon;
            if (p0y == 0.0f) {
                p1x = kScanEpsilon;
                p1y = tile_size;
            } else {
                p1x = 2.0f * kScanEpsilon;
                p1y = p0y;
            }
        } else if (p0y == 0.0f) {
            p0x = kScanEpsilon;
        } else {
            edge = p0y;
        }
    } else if (p1x == 0.0f) {
        if (p1y == 0.0f) p1x = kScanEpsilon;
        else edge = p1y;
    }
    if (floorf(p0x) == p0x && p0x != 0.0f) p0x -= kScanEpsilon;
    if (floorf(p1x) == p1x && p1x != 0.0f) p1x -= kScanEpsilon;
    tk_line_segment out;
    if (!s.down) {
        out.p0x = p1x; out.p0y = p1y; out.p1x = p0x; out.p1y = p0y;
    } else {
        out.p0x = p0x; out.p0y = p0y; out.p1x = p1x; out.p1y = p1y;
    }
    out.y_edge = edge;
    return out;
}

// ---------------------------------------------------------------------------
// Pixel math (shared/pixel.hlsli:7-66)
// ---------------------------------------------------------------------------

struct Px {
    float r, g, b, a;
};

// The clamps keep the value in [0.5, 255.5] (a NaN clamps to 0), where a plain
// truncating conversion equals f2u's, without f2u's range branches.
TKD uint32_t coverage_to_u8(float c) { return (uint32_t)(clampf(c, 0.0f, 1.0f) * 255.0f + 0.5f); }
TKD uint32_t coverage_to_alpha(float v, uint32_t rule) {
    float alpha = fminf(fabsf(v), 1.0f);
    if (rule == 1u) alpha = fabsf(v - 2.0f * rintf(0.5f * v));
    return coverage_to_u8(alpha);
}
TKD float signum_f32(float v) { return v < 0.0f ? -1.0f : 1.0f; }
TKD uint32_t rgba8_pack(uint32_t r, uint32_t g, uint32_t b, uint32_t a) { return r | (g << 8) | (b << 16) | (a << 24); }
TKD uint32_t chan_u8(float v) { return (uint32_t)clampf(v * 255.0f + 0.5f, 0.0f, 255.0f); }  // in [0, 255], as above
TKD uint32_t unorm_to_rgba8(Px p) { return rgba8_pack(chan_u8(p.r), chan_u8(p.g), chan_u8(p.b), chan_u8(p.a)); }
TKD Px rgba8_to_unorm(uint32_t p) {
    return Px{(float)(p & 255u) * kChannelScale, (float)((p >> 8) & 255u) * kChannelScale,
              (float)((p >> 16) & 255u) * kChannelScale, (float)(p >> 24) * kChannelScale};
}
TKD uint32_t mul_div255(uint32_t a, uint32_t b) {
    uint32_t product = a * b + 128u;
    return (product + (product >> 8)) >> 8;
}
TKD uint32_t combine_alpha(uint32_t a, uint32_t b) { return mul_div255(a, b); }
TKD float straight_channel(uint32_t premul, uint32_t alpha) { return alpha == 0u ? 0.0f : (float)premul / (float)alpha; }
TKD uint32_t scale_premul_u8(uint32_t source, uint32_t factor) {
    if (factor == 0u) return 0u;
    if (factor == 255u) return source;
    return rgba8_pack(mul_div255(source & 255u, factor), mul_div255((source >> 8) & 255u, factor),
                      mul_div255((source >> 16) & 255u, factor), mul_div255(source >> 24, factor));
}
TKD Px scale_premul_u8_to_unorm(uint32_t source, uint32_t factor) {
    if (factor == 0u || (source >> 24) == 0u) return Px{0.0f, 0.0f, 0.0f, 0.0f};
    float scale = (float)factor * (1.0f / 65025.0f);
    return Px{(float)(source & 255u) * scale, (float)((source >> 8) & 255u) * scale,
              (float)((source >> 16) & 255u) * scale, (float)(source >> 24) * scale};
}
TKD uint32_t src_over_premul_u8(uint32_t dst, uint32_t src) {
    uint32_t alpha = src >> 24;
    if (alpha == 0u) return dst;
    if (alpha == 255u) return src;
    uint32_t inv = 255u - alpha;
    return rgba8_pack((src & 255u) + mul_div255(dst & 255u, inv), ((src >> 8) & 255u) + mul_div255((dst >> 8) & 255u, inv),
                      ((src >> 16) & 255u) + mul_div255((dst >> 16) & 255u, inv), alpha + mul_div255(dst >> 24, inv));
}
TKD Px src_over_premul_unorm(Px dst, Px src) {
    if (src.a <= 0.0f) return dst;
    if (src.a >= 1.0f) return src;
    float inv = 1.0f - src.a;
    return Px{src.r + dst.r * inv, src.g + dst.g * inv, src.b + dst.b * inv, src.a + dst.a * inv};
}
TKD uint32_t pack_premul_rgba8(float r, float g, float b, float a) {
    return rgba8_pack(f2u(clampf(r, 0.0f, 1.0f) * 255.0f + 0.5f), f2u(clampf(g, 0.0f, 1.0f) * 255.0f + 0.5f),
                      f2u(clampf(b, 0.0f, 1.0f) * 255.0f + 0.5f), f2u(clampf(a, 0.0f, 1.0f) * 255.0f + 0.5f));
}
TKD uint32_t lerp_premul_u8(uint32_t a, uint32_t b, float t) {
    uint32_t out = 0;
    for (int k = 0; k < 4; ++k) {
        float left = (float)((a >> (8 * k)) & 255u);
        float right = (float)((b >> (8 * k)) & 255u);
        uint32_t v = f2u(clampf(fmaf(right - left, t, left) + 0.5f, 0.0f, 255.0f));
        out |= v << (8 * k);
    }
    return out;
}
TKD float rem_euclid_f32(float value, float modulus) { return fmaf(-floorf(value / modulus), modulus, value); }

// ---------------------------------------------------------------------------
// Row-sweep area coverage (shared/coverage.hlsli:7-51)
// ---------------------------------------------------------------------------

struct RowParts {
    float edge, dy, xmin, xmax;
};

TKD RowParts segment_row_parts(float p0x, float p0y, float p1x, float p1y, float y_edge, uint32_t y) {
    float delta_x = p1x - p0x, delta_y = p1y - p0y;
    float row_y = (float)y, local_y = p0y - row_y;
    float y0 = clampf(local_y, 0.0f, 1.0f), y1 = clampf(p1y - row_y, 0.0f, 1.0f);
    float dy = y0 - y1;
    float row_edge = signum_f32(delta_x) * clampf(row_y - y_edge + 1.0f, 0.0f, 1.0f);
    if (dy == 0.0f) return RowParts{row_edge, dy, 0.0f, 0.0f};
    float slope = delta_x / delta_y;
    bool nearer_p0 = fabsf(local_y - 0.5f) <= fabsf(p1y - row_y - 0.5f);
    float anchor_x = nearer_p0 ? p0x : p1x;
    float anchor_y = nearer_p0 ? local_y : p1y - row_y;
    float sx0 = fmaf(y0 - anchor_y, slope, anchor_x);
    float sx1 = fmaf(y1 - anchor_y, slope, anchor_x);
    return RowParts{row_edge, dy, fminf(sx0, sx1), fmaxf(sx0, sx1)};
}

// segment_area_at with the pixel's x already in FP32 (an integer 0..15, exact).
TKD float segment_area_at_f(float xmin_abs, float xmax_abs, float xf) {
    float xmin = xmin_abs - xf, xmax = xmax_abs - xf;
    float width = xmax - xmin;
    if (width == 0.0f) return clampf(1.0f - xmin, 0.0f, 1.0f);
    if (xmin >= 0.0f && xmax <= 1.0f) return fmaf(-0.5f, xmin + xmax, 1.0f);
    float left = clampf(xmin, 0.0f, 1.0f), right = clampf(xmax, 0.0f, 1.0f);
    float full = clampf(-xmin, 0.0f, width);
    return fmaf(right - left, fmaf(-0.5f, left + right, 1.0f), full) / width;
}
TKD float segment_area_at(float xmin_abs, float xmax_abs, uint32_t x) {
    return segment_area_at_f(xmin_abs, xmax_abs, (float)x);
}

// fill_alpha_at keeps three sums per pixel (base, running, partial) and combines
// them as (base + running) + partial in FP32.  Here every FP32 term is truncated
// to a multiple of 2^-30 and the terms are added as 64-bit integers, the three
// sums included, so the value does not depend on the order in which the tile's
// segments were emitted.  Only the one conversion of the total to FP32 rounds.
//
// Every term lies in (-2, 2): the edge term is signum * a clamp to [0, 1], dy is
// the difference of two values clamped to [0, 1], and the area factor is at most
// 1 up to rounding.  So t * 2^30 (exact, a power-of-two scaling) fits an int32
// and one truncating v_cvt_i32_f32 gives trunc(t * 2^30); a NaN term counts as 0.
TKD long long coverage_fixed(float t) {
    const int32_t q = (int32_t)(t * 1073741824.0f);
    return t == t ? (long long)q : 0ll;
}
// Back to FP32: the sum floored to a multiple of 2^-24 (at most 6e-8 below it),
// then one rounding.  hi = floor(v / 2^30) and the top 24 of the 30 fraction bits
// convert exactly (|hi| < 2^24 for any coverage sum), the scaling by 2^-24 is
// exact, and only the add rounds.
TKD float fixed_to_float(long long v) {
    const float hi = (float)(int32_t)(v >> 30);
    const float lo = (float)(((uint32_t)v & 0x3FFFFFFFu) >> 6) * 5.9604644775390625e-8f;
    return hi + lo;
}
TKD long long backdrop_fixed(int32_t backdrop) { return (long long)backdrop * 1073741824LL; }

// One pixel's coverage from segments [start, end), terms summed exactly.
TKD float exact_fill_value(const tk_line_segment *segs, int32_t backdrop, uint32_t start, uint32_t end, uint32_t x,
                           uint32_t y) {
    long long base = backdrop_fixed(backdrop), running = 0, partial = 0;
    for (uint32_t i = start; i < end; ++i) {
        const tk_line_segment seg = segs[i];
        RowParts parts = segment_row_parts(seg.p0x, seg.p0y, seg.p1x, seg.p1y, seg.y_edge, y);
        base += coverage_fixed(parts.edge);
        if (parts.dy != 0.0f) {
            int full_start = clamp_cell16(ceilf(parts.xmax));
            if (full_start < 16 && (int)x >= full_start) running += coverage_fixed(parts.dy);
            int partial_start = clamp_cell16(floorf(parts.xmin));
            if ((int)x >= partial_start && (int)x < full_start)
                partial += coverage_fixed(segment_area_at(parts.xmin, parts.xmax, x) * parts.dy);
        }
    }
    return fixed_to_float(base + running + partial);
}

// ---------------------------------------------------------------------------
// Paint words, SDF rect coverage, brush sampling
// (shared/sdf/{base,basic,coverage,data}.hlsli, shared/brush/*.hlsli, shared/affine.hlsli)
// ---------------------------------------------------------------------------

// Bounds-checked word read with robust-buffer semantics: out of range reads 0.
TKD uint32_t paint_word(const uint32_t *paint, uint32_t count, uint32_t index) {
    return index < count ? paint[index] : 0u;
}
// sdf_word / brush_word (shared/sdf/data.hlsli:3-8, shared/brush/data.hlsli:3-14).
TKD uint32_t blob_word(const uint32_t *paint, uint32_t count, uint32_t base, uint32_t index) {
    if (base < count) {
        if (index < count - base) return paint[base + index];
    }
    return 0u;
}
TKD float blob_float(const uint32_t *paint, uint32_t count, uint32_t base, uint32_t index) {
    return asfloat(blob_word(paint, count, base, index));
}

struct Affine {
    float a, b, c, d, e, f;
};
// affine_record_point (shared/affine.hlsli:14-18).
TKD void affine_point(const Affine &t, float x, float y, float &ox, float &oy) {
    ox = fmaf(t.a, x, fmaf(t.c, y, t.e));
    oy = fmaf(t.b, x, fmaf(t.d, y, t.f));
}

struct SdfSample {
    float distance, nx, ny;
};

TKD float vec_length(float x, float y) { return sqrtf(x * x + y * y); }
// sdf_length (shared/sdf/base.hlsli:8-13).
TKD float sdf_length(float x, float y) {
    float magnitude = vec_length(x, y);
    if (magnitude > 0.0f && isfinite(magnitude)) return magnitude;
    float axis = fmaxf(fabsf(x), fabsf(y));
    return axis == 0.0f ? 0.0f : axis * vec_length(x / axis, y / axis);
}

// rounded_box_sdf_sample (shared/sdf/basic.hlsli:6-19).
TKD SdfSample rounded_box_sdf_sample(float px, float py, float hx, float hy, float radius) {
    float qx = fabsf(px) - hx + radius, qy = fabsf(py) - hy + radius;
    float ox = fmaxf(qx, 0.0f), oy = fmaxf(qy, 0.0f);
    float outside_length = sdf_length(ox, oy);
    float distance = fminf(fmaxf(qx, qy), 0.0f) + outside_length - radius;
    float sx = px >= 0.0f ? 1.0f : -1.0f, sy = py >= 0.0f ? 1.0f : -1.0f;
    if (outside_length > 0.0f) return SdfSample{distance, ox / outside_length * sx, oy / outside_length * sy};
    if (qx > qy) return SdfSample{distance, sx, 0.0f};
    return SdfSample{distance, 0.0f, sy};
}

// rect_sdf_sample (shared/sdf/basic.hlsli:34-57).
TKD SdfSample rect_sdf_sample(float x, float y, float x0r, float y0r, float x1r, float y1r, float tl, float tr, float bl,
                              float br) {
    float x0 = fminf(x0r, x1r), y0 = fminf(y0r, y1r), x1 = fmaxf(x0r, x1r), y1 = fmaxf(y0r, y1r);
    float cx = (x0 + x1) * 0.5f, cy = (y0 + y1) * 0.5f;
    float hx = (x1 - x0) * 0.5f, hy = (y1 - y0) * 0.5f;
    float px = x - cx, py = y - cy;
    float radius = tl;
    if (px >= 0.0f) {
        radius = py <= 0.0f ? tr : br;
    } else if (py > 0.0f) {
        radius = bl;
    }
    float r = fmaxf(fminf(fminf(radius, hx), hy), 0.0f);
    return rounded_box_sdf_sample(px, py, hx, hy, r);
}

// sdf_coverage_from_sample (shared/sdf/base.hlsli:44-52).
TKD float sdf_coverage_from_sample(const SdfSample &s, const Affine &inv) {
    float gx = fmaf(inv.a, s.nx, inv.b * s.ny);
    float gy = fmaf(inv.c, s.nx, inv.d * s.ny);
    float footprint = sdf_length(gx, gy);
    if (footprint == 0.0f) return s.distance < 0.0f ? 1.0f : (s.distance > 0.0f ? 0.0f : 0.5f);
    return clampf(0.5f - s.distance / footprint, 0.0f, 1.0f);
}

// sdf_coverage_from_blob (shared/sdf/coverage.hlsli:13-74) for the rectangle kinds
// on this path; other SDF kinds contribute no coverage here.
TKD float sdf_coverage_from_blob(const uint32_t *paint, uint32_t count, uint32_t base, float x, float y,
                                 const Affine &inv) {
    uint32_t kind = blob_word(paint, count, base, 0u);
    float x0 = blob_float(paint, count, base, 1u), y0 = blob_float(paint, count, base, 2u);
    float x1 = blob_float(paint, count, base, 3u), y1 = blob_float(paint, count, base, 4u);
    float r0 = blob_float(paint, count, base, 5u), r1 = blob_float(paint, count, base, 6u);
    float r2 = blob_float(paint, count, base, 7u), r3 = blob_float(paint, count, base, 8u);
    if (kind == TK_SDF_KIND_RECT) return sdf_coverage_from_sample(rect_sdf_sample(x, y, x0, y0, x1, y1, r0, r1, r2, r3), inv);
    if (kind == 3u) {  // SDF_KIND_RECT_STROKE
        float st = blob_float(paint, count, base, 9u), sr = blob_float(paint, count, base, 10u);
        float sb = blob_float(paint, count, base, 11u), sl = blob_float(paint, count, base, 12u);
        float ht = fmaxf(st, 0.0f), hr = fmaxf(sr, 0.0f), hb = fmaxf(sb, 0.0f), hl = fmaxf(sl, 0.0f);
        float rx0 = fminf(x0, x1), ry0 = fminf(y0, y1), rx1 = fmaxf(x0, x1), ry1 = fmaxf(y0, y1);
        float outer = sdf_coverage_from_sample(
            rect_sdf_sample(x, y, rx0 - hl, ry0 - ht, rx1 + hr, ry1 + hb, r0 + fmaxf(ht, hl), r1 + fmaxf(ht, hr),
                            r2 + fmaxf(hb, hl), r3 + fmaxf(hb, hr)),
            inv);
        float ix0 = rx0 + hl, iy0 = ry0 + ht, ix1 = rx1 - hr, iy1 = ry1 - hb;
        float inner = 0.0f;
        if (ix0 < ix1 && iy0 < iy1) {
            inner = sdf_coverage_from_sample(
                rect_sdf_sample(x, y, ix0, iy0, ix1, iy1, fmaxf(r0 - fmaxf(ht, hl), 0.0f), fmaxf(r1 - fmaxf(ht, hr), 0.0f),
                                fmaxf(r2 - fmaxf(hb, hl), 0.0f), fmaxf(r3 - fmaxf(hb, hr), 0.0f)),
                inv);
        }
        return clampf(outer - inner, 0.0f, 1.0f);
    }
    return 0.0f;
}

// apply_extend / sample_ramp / sample_linear (shared/brush/{extend,ramp,linear}.hlsli).
TKD float apply_extend(float t, uint32_t extend) {
    if (extend == 1u) return rem_euclid_f32(t, 1.0f);
    if (extend == 2u) {
        float v = rem_euclid_f32(t, 2.0f);
        return v <= 1.0f ? v : 2.0f - v;
    }
    return clampf(t, 0.0f, 1.0f);
}
TKD uint32_t sample_ramp(const uint32_t *paint, uint32_t count, uint32_t brush_base, uint32_t payload_offset,
                         uint32_t payload_len, float t, uint32_t extend) {
    if (payload_len == 0u) return 0u;
    uint32_t last = payload_len - 1u;
    float extended = apply_extend(t, extend);
    float position = extended * (float)last;
    uint32_t left_ix = f2u(floorf(position)), right_ix = min(left_ix + 1u, last);
    float fraction = fmaf(extended, (float)last, -(float)left_ix);
    uint32_t left = blob_word(paint, count, brush_base, payload_offset + left_ix);
    uint32_t right = blob_word(paint, count, brush_base, payload_offset + right_ix);
    if (fraction <= kGradientEpsilon || left_ix == right_ix) return left;
    return lerp_premul_u8(left, right, fraction);
}
// The ramp parameter of each gradient kind at (x, y); the ramp lookup is shared
// (sample_brush below), so fine carries one inlined copy of it.
//
// linear_t: sample_linear (shared/brush/linear.hlsli:5-14).
TKD float linear_t(const uint32_t *paint, uint32_t count, uint32_t brush_base, float x, float y, uint32_t base) {
    float p[10];
    for (int i = 0; i < 10; ++i) p[i] = blob_float(paint, count, brush_base, base + (uint32_t)i);
    float tx = fmaf(p[4], x, fmaf(p[6], y, p[8]));
    float ty = fmaf(p[5], x, fmaf(p[7], y, p[9]));
    float dx = p[2] - p[0], dy = p[3] - p[1];
    float denominator = dx * dx + dy * dy, t = 0.0f;
    if (denominator > kGradientEpsilon) t = fmaf(tx - p[0], dx, (ty - p[1]) * dy) / denominator;
    return t;
}

// radial_t: sample_radial (shared/brush/radial.hlsli:5-33), the two-point conical
// gradient: t from the quadratic |q - t dc| = r0 + t dr with r0 + t dr >= 0;
// has_t is false where no such t exists (the sample is transparent).
constexpr float kRadialQuadraticEpsilon = 0.000001f;  // shared/brush/radial.hlsli:4
TKD float radial_t(const uint32_t *paint, uint32_t count, uint32_t brush_base, float x, float y, uint32_t base,
                   bool &has_t) {
    float p[12];
    for (int i = 0; i < 12; ++i) p[i] = blob_float(paint, count, brush_base, base + (uint32_t)i);
    const float tx = fmaf(p[6], x, fmaf(p[8], y, p[10]));
    const float ty = fmaf(p[7], x, fmaf(p[9], y, p[11]));
    const float r0 = p[4], r1 = p[5];
    const float qx = tx - p[0], qy = ty - p[1], dcx = p[2] - p[0], dcy = p[3] - p[1], dr = r1 - r0;
    const float qa = fmaf(dcx, dcx, fmaf(dcy, dcy, -(dr * dr)));
    const float qb = -2.0f * fmaf(qx, dcx, fmaf(qy, dcy, r0 * dr));
    const float qc = fmaf(qx, qx, fmaf(qy, qy, -(r0 * r0)));
    has_t = false;
    float t = 0.0f;
    if (fabsf(qa) <= kRadialQuadraticEpsilon) {
        if (fabsf(qb) > kRadialQuadraticEpsilon) {
            const float candidate = -qc / qb;
            if (fmaf(candidate, dr, r0) >= 0.0f) {
                has_t = true;
                t = candidate;
            }
        }
    } else {
        const float discriminant = fmaf(qb, qb, (-4.0f * qa) * qc);
        if (discriminant >= 0.0f) {
            const float root = sqrtf(discriminant);
            const float t0 = (-qb - root) / (2.0f * qa), t1 = (-qb + root) / (2.0f * qa);
            const bool valid0 = fmaf(t0, dr, r0) >= 0.0f, valid1 = fmaf(t1, dr, r0) >= 0.0f;
            if (valid0) {
                has_t = true;
                t = valid1 ? fmaxf(t0, t1) : t0;
            } else if (valid1) {
                has_t = true;
                t = t1;
            }
        }
    }
    return t;
}

// atan2(y, x) for (x, y) != (0, 0): the octant reduction a = min / max of |x|, |y|
// and the odd polynomial of Abramowitz & Stegun 4.4.49 (Hastings) for atan(a) on
// [0, 1], absolute error 2e-8 before FP32 rounding (the ratio uses the hardware
// reciprocal, 1 ulp); then the quadrant.  It keeps fine's register budget that
// the device library's atan2f would raise; sweep samples stay within 1 LSB of an
// exact angle.
TKD float sweep_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float a = fminf(ax, ay) * __builtin_amdgcn_rcpf(fmaxf(ax, ay)), s = a * a;
    float p = 0.0028662257f;
    p = fmaf(p, s, -0.0161657367f);
    p = fmaf(p, s, 0.0429096138f);
    p = fmaf(p, s, -0.0752896400f);
    p = fmaf(p, s, 0.1065626393f);
    p = fmaf(p, s, -0.1420889944f);
    p = fmaf(p, s, 0.1999355085f);
    p = fmaf(p, s, -0.3333314528f);
    float r = fmaf(a * s, p, a);
    if (ay > ax) r = 1.57079637f - r;
    if (x < 0.0f) r = 3.14159274f - r;
    return y < 0.0f ? -r : r;
}

// sweep_t: sample_sweep (shared/brush/sweep.hlsli:5-20), the angle from the
// centre measured from the start orientation in the direction of the signed span.
constexpr float kGradientTau = 6.2831855f;  // shared/brush/sweep.hlsli:4
TKD float sweep_t(const uint32_t *paint, uint32_t count, uint32_t brush_base, float x, float y, uint32_t base) {
    const float cx = blob_float(paint, count, brush_base, base), cy = blob_float(paint, count, brush_base, base + 1u);
    const float start_angle = blob_float(paint, count, brush_base, base + 2u);
    const float span = blob_float(paint, count, brush_base, base + 3u);
    float t = 0.0f;
    if (fabsf(span) > kGradientEpsilon) {
        float angle = 0.0f;
        if (x != cx || y != cy) angle = sweep_atan2(y - cy, x - cx);
        float distance = angle - start_angle;
        if (span > 0.0f) {
            if (distance < 0.0f) distance += kGradientTau;
        } else if (distance > 0.0f) {
            distance -= kGradientTau;
        }
        t = distance / span;
    }
    return t;
}

// sample_four_corner (shared/brush/four_corner.hlsli:5-15): bilinear blend of the
// corner colours tl, tr, br, bl over the bounds.
TKD uint32_t sample_four_corner(const uint32_t *paint, uint32_t count, uint32_t brush_base, float x, float y,
                                uint32_t base, uint32_t payload) {
    const float x0 = blob_float(paint, count, brush_base, base), y0 = blob_float(paint, count, brush_base, base + 1u);
    const float x1 = blob_float(paint, count, brush_base, base + 2u), y1 = blob_float(paint, count, brush_base, base + 3u);
    const float width = x1 - x0, height = y1 - y0;
    float u = 0.0f, v = 0.0f;
    if (fabsf(width) > kGradientEpsilon) u = clampf((x - x0) / width, 0.0f, 1.0f);
    if (fabsf(height) > kGradientEpsilon) v = clampf((y - y0) / height, 0.0f, 1.0f);
    const uint32_t tl = blob_word(paint, count, brush_base, payload), tr = blob_word(paint, count, brush_base, payload + 1u);
    const uint32_t br = blob_word(paint, count, brush_base, payload + 2u), bl = blob_word(paint, count, brush_base, payload + 3u);
    return lerp_premul_u8(lerp_premul_u8(tl, tr, u), lerp_premul_u8(bl, br, u), v);
}

// ---------------------------------------------------------------------------
// Text compositing of path glyphs (PTCL_PATH_GLYPH, fine/interpreter.hlsli:205-206):
// fine/text/gamma.hlsli:5-21, fine/text/auto.hlsli:6-33, fine/text/basic.hlsli:101-115.
// HLSL lerp(a, b, t) is written fmaf(b - a, t, a), as lerp_premul_u8 does; every
// other product and sum rounds separately; dot() sums its products in order.
// ---------------------------------------------------------------------------

TKD float srgb_to_linear(float value) {
    const float v = clampf(value, 0.0f, 1.0f);
    float result = v / 12.92f;
    if (v > 0.04045f) result = powf((v + 0.055f) / 1.055f, 2.4f);
    return result;
}
TKD float linear_to_srgb(float value) {
    const float v = clampf(value, 0.0f, 1.0f);
    float result = v * 12.92f;
    if (v > 0.0031308f) result = 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
    return result;
}
// text_foreground: the straight colour, as a luminance weight (dot with Rec. 709).
TKD float text_foreground_luma(uint32_t source) {
    const float alpha = (float)(source >> 24);
    if (!(alpha > 0.0f)) return 0.0f;
    const float r = clampf((float)(source & 255u) / alpha, 0.0f, 1.0f);
    const float g = clampf((float)((source >> 8) & 255u) / alpha, 0.0f, 1.0f);
    const float b = clampf((float)((source >> 16) & 255u) / alpha, 0.0f, 1.0f);
    return r * 0.2126f + g * 0.7152f + b * 0.0722f;
}
TKD float text_correcting_coverage(uint32_t coverage, float foreground) {
    if (coverage == 0u || coverage == 255u) return (float)coverage * (1.0f / 255.0f);
    const float raw = (float)coverage * (1.0f / 255.0f);
    const float background = 1.0f - foreground;
    const float bg_linear = srgb_to_linear(background);
    const float enhanced = raw + raw * (1.0f - raw) * bg_linear;
    const float difference = foreground - background;
    if (fabsf(difference) < (1.0f / 256.0f)) return enhanced;
    const float fg_linear = srgb_to_linear(foreground);
    const float output = linear_to_srgb(fmaf(fg_linear - bg_linear, enhanced, bg_linear));
    return clampf((output - background) / difference, 0.0f, 1.0f);
}
// src_over_subpixel_mask_f32 with one coverage for the three channels (the
// path-glyph case; the grey reduction of a non-opaque destination is then the
// same value, (c + c + c) / 3 rounded).
TKD uint32_t src_over_mask_f32(uint32_t destination, uint32_t source, float coverage) {
    if ((source >> 24) == 0u || coverage == 0.0f) return destination;
    float cov = coverage;
    if ((destination >> 24) != 255u) cov = (coverage + coverage + coverage) / 3.0f;
    const float alpha = (float)(source >> 24) * (1.0f / 255.0f);
    const float amount = alpha * cov;
    uint32_t out = 0u;
    for (int k = 0; k < 3; ++k) {
        const float s = (float)((source >> (8 * k)) & 255u), d = (float)((destination >> (8 * k)) & 255u);
        out |= f2u(s * cov + d * (1.0f - amount) + 0.5f) << (8 * k);
    }
    const float out_alpha = amount * 255.0f + (float)(destination >> 24) * (1.0f - amount);
    return out | (f2u(out_alpha + 0.5f) << 24);
}
TKD uint32_t src_over_mask_linear_auto_u8(uint32_t destination, uint32_t source, uint32_t coverage) {
    if ((source >> 24) == 0u || coverage == 0u) return destination;
    const float corrected = text_correcting_coverage(coverage, text_foreground_luma(source));
    return src_over_mask_f32(destination, source, corrected);
}

// ---------------------------------------------------------------------------
// Image brushes (shared/brush/pattern.hlsli:9-55, shared/pattern_transform.hlsli,
// shared/integer.hlsli): the scene's atlas of premultiplied RGBA8 texels.
// ---------------------------------------------------------------------------

struct AtlasView {
    const uint32_t *texels;
    uint32_t width, height, pages;
};

// euclidean_remainder_i32 (shared/integer.hlsli:5-9).
TKD uint32_t euclidean_remainder_i32(int32_t value, uint32_t period) {
    const uint32_t magnitude = value < 0 ? 0u - (uint32_t)value : (uint32_t)value;
    const uint32_t remainder = magnitude % period;
    return value < 0 && remainder != 0u ? period - remainder : remainder;
}
// extend_coord_i32 (pattern.hlsli:9-17): pad clamps, repeat wraps, reflect mirrors.
TKD uint32_t extend_coord_i32(int32_t value, uint32_t size, uint32_t extend) {
    if (size <= 1u) return 0u;
    if (extend == 1u || extend == 2u) {
        const uint32_t period = size * (extend == 2u ? 2u : 1u);
        const uint32_t coord = euclidean_remainder_i32(value, period);
        return coord < size ? coord : period - coord - 1u;
    }
    return (uint32_t)clampi(value, 0, (int)size - 1);
}
// atlas.Load of one texel; outside the atlas reads 0, as a Load out of range does.
TKD uint32_t atlas_texel(const AtlasView &a, uint32_t page, uint32_t x, uint32_t y) {
    if (!a.texels || page >= a.pages || x >= a.width || y >= a.height) return 0u;
    return a.texels[((size_t)page * a.height + y) * a.width + x];
}
// atlas_pattern_pixel (pattern.hlsli:18-21).
TKD uint32_t atlas_pattern_pixel(const AtlasView &a, uint32_t page, uint32_t ox, uint32_t oy, uint32_t sx,
                                 uint32_t sy, uint32_t extend, int32_t px, int32_t py) {
    return atlas_texel(a, page, ox + extend_coord_i32(px, sx, extend), oy + extend_coord_i32(py, sy, extend));
}
// pattern_transform_component (shared/pattern_transform.hlsli:7-11): the second
// product's rounding error is recovered by a fused multiply-add.
TKD float pattern_transform_component(float a, float b, float offset, float x, float y) {
    const float by = b * y;
    const float dot = fmaf(a, x, by) + fmaf(b, y, -by);
    return dot + offset;
}
// The filtered read the reference makes through a bilinear sampler for a padded
// bilinear pattern (atlas.SampleLevel, pattern.hlsli:25-28): position clamped to
// the image, texel centres at half-integers, the four neighbours clamped to the
// image (an edge replicated, as an atlas gutter does).  A texture unit resolves
// the position to 8 fractional bits (the subtexel precision Direct3D requires),
// so each weight is k / 256 with k the position's fraction rounded to the
// nearest 1/256; the blend is then the integer sum of weighted texels, rounded
// once to 8 bits.
TKD void subtexel_position(float q, int32_t &index, uint32_t &weight) {
    const float base = floorf(q);
    weight = (uint32_t)rintf((q - base) * 256.0f);  // q - base is exact
    index = f2i(base);
    if (weight == 256u) {
        index += 1;
        weight = 0u;
    }
}
TKD uint32_t filtered_pad_sample(const AtlasView &a, uint32_t page, uint32_t ox, uint32_t oy, uint32_t sx, uint32_t sy,
                                 float px, float py) {
    int32_t ix, iy;
    uint32_t wx, wy;
    subtexel_position(clampf(px, 0.0f, (float)sx) - 0.5f, ix, wx);
    subtexel_position(clampf(py, 0.0f, (float)sy) - 0.5f, iy, wy);
    const uint32_t t[4] = {atlas_pattern_pixel(a, page, ox, oy, sx, sy, 0u, ix, iy),
                           atlas_pattern_pixel(a, page, ox, oy, sx, sy, 0u, ix + 1, iy),
                           atlas_pattern_pixel(a, page, ox, oy, sx, sy, 0u, ix, iy + 1),
                           atlas_pattern_pixel(a, page, ox, oy, sx, sy, 0u, ix + 1, iy + 1)};
    const uint32_t w[4] = {(256u - wx) * (256u - wy), wx * (256u - wy), (256u - wx) * wy, wx * wy};
    uint32_t out = 0u;
    for (int k = 0; k < 4; ++k) {
        uint32_t sum = 32768u;
        for (int n = 0; n < 4; ++n) sum += w[n] * ((t[n] >> (8 * k)) & 255u);
        out |= (sum >> 16) << (8 * k);
    }
    return out;
}
// sample_resource_pattern_atlas (pattern.hlsli:22-43).
TKD uint32_t sample_pattern_atlas(const AtlasView &a, float px, float py, uint32_t page, uint32_t ox, uint32_t oy,
                                  uint32_t sx, uint32_t sy, uint32_t opacity, uint32_t extend, uint32_t sampling) {
    if (sx == 0u || sy == 0u) return 0u;
    uint32_t color;
    if (sampling == 1u && extend == 0u) {
        color = filtered_pad_sample(a, page, ox, oy, sx, sy, px, py);
    } else if (sampling == 1u) {
        const float spx = px - 0.5f, spy = py - 0.5f;
        const float bx = floorf(spx), by = floorf(spy);
        const float fx = spx - bx, fy = spy - by;
        const int32_t ix = f2i(bx), iy = f2i(by);
        const uint32_t tl = atlas_pattern_pixel(a, page, ox, oy, sx, sy, extend, ix, iy);
        const uint32_t tr = atlas_pattern_pixel(a, page, ox, oy, sx, sy, extend, ix + 1, iy);
        const uint32_t bl = atlas_pattern_pixel(a, page, ox, oy, sx, sy, extend, ix, iy + 1);
        const uint32_t br = atlas_pattern_pixel(a, page, ox, oy, sx, sy, extend, ix + 1, iy + 1);
        color = lerp_premul_u8(lerp_premul_u8(tl, tr, fx), lerp_premul_u8(bl, br, fx), fy);
    } else {
        color = atlas_pattern_pixel(a, page, ox, oy, sx, sy, extend, f2i(floorf(px)), f2i(floorf(py)));
    }
    return scale_premul_u8(color, opacity);
}
// sample_atlas_brush (pattern.hlsli:44-54), for a header whose placement the
// encoding patched: words 2-6 = x, y, page, width, height.
TKD uint32_t sample_image_brush(const uint32_t *paint, uint32_t count, uint32_t brush_base, uint32_t data_base,
                                const AtlasView &atlas, float x, float y) {
    const uint32_t base = data_base + TK_BRUSH_HEADER_WORDS;
    const uint32_t sx = blob_word(paint, count, brush_base, data_base + 5u);
    const uint32_t sy = blob_word(paint, count, brush_base, data_base + 6u);
    if (sx == 0u || sy == 0u) return 0u;
    const uint32_t placement = blob_word(paint, count, brush_base, data_base + 4u);
    if (placement & 0x80000000u) return 0u;  // a texture-table image: this path holds only the atlas
    float p[6];
    for (int i = 0; i < 6; ++i) p[i] = blob_float(paint, count, brush_base, base + (uint32_t)i);
    const float tx = pattern_transform_component(p[0], p[2], p[4], x, y) * (float)sx;
    const float ty = pattern_transform_component(p[1], p[3], p[5], x, y) * (float)sy;
    return sample_pattern_atlas(atlas, tx, ty, placement, blob_word(paint, count, brush_base, data_base + 2u),
                                blob_word(paint, count, brush_base, data_base + 3u), sx, sy,
                                blob_word(paint, count, brush_base, data_base + 7u),
                                blob_word(paint, count, brush_base, data_base + 1u),
                                blob_word(paint, count, brush_base, data_base + 8u));
}

// sample_brush (shared/brush/sample.hlsli:15-42) for the solid, linear, radial,
// sweep, four-corner and image brushes.
TKD uint32_t sample_brush(const uint32_t *paint, uint32_t count, uint32_t brush_base, uint32_t data_base, float x,
                          float y, const AtlasView &atlas = AtlasView{nullptr, 0u, 0u, 0u}) {
    uint32_t kind = blob_word(paint, count, brush_base, data_base);
    uint32_t extend = blob_word(paint, count, brush_base, data_base + 1u);
    uint32_t payload = data_base + blob_word(paint, count, brush_base, data_base + 2u);
    uint32_t len = blob_word(paint, count, brush_base, data_base + 3u);
    uint32_t base = data_base + TK_BRUSH_HEADER_WORDS;
    uint32_t color = blob_word(paint, count, brush_base, data_base + 4u);
    if (kind == TK_BRUSH_LINEAR || kind == TK_BRUSH_RADIAL || kind == TK_BRUSH_SWEEP) {
        bool has_t = true;
        float t;
        if (kind == TK_BRUSH_LINEAR) t = linear_t(paint, count, brush_base, x, y, base);
        else if (kind == TK_BRUSH_RADIAL) t = radial_t(paint, count, brush_base, x, y, base, has_t);
        else t = sweep_t(paint, count, brush_base, x, y, base);
        return has_t ? sample_ramp(paint, count, brush_base, payload, len, t, extend) : 0u;
    }
    if (kind == TK_BRUSH_FOUR_CORNER) return sample_four_corner(paint, count, brush_base, x, y, base, payload);
    if (kind == TK_BRUSH_PATTERN) return 0u;
    if (kind == TK_BRUSH_IMAGE) return sample_image_brush(paint, count, brush_base, data_base, atlas, x, y);
    // Mesh and program brushes are not produced on this path.
    return color;
}

}  // namespace tkd
