// Host-side scene recording and planning for the compute-raster path.
//
// This is the recording layer the GPU stages consume: Canvas::push_* turn
// shapes into flattened Lines plus PathRecord / DrawRecord / paint words
// (src/canvas/paths.rs:65-156, src/canvas/primitives/rectangles.rs:33-43,
// src/canvas/layers.rs:7-78), compile groups draws into DrawBatches with fused
// clip stacks (src/canvas/compile.rs:134-217, src/shared/execution/batches.rs:
// 202-273), and the planner derives tile bins and scan/cumsum chunk metadata
// (src/shared/gpu_plan/tile_bins/build.rs:4-83, gpu_plan/path_plans.rs:463-504).
//
// Curve flattening and stroke outlining are this project's own geometry
// front end: the reference delegates both to kurbo 0.13.1, which is not part of
// the reference tree.  Flattening places chords by the parabola-integral method
// so the chord error stays within the tolerance with few chords; strokes are the
// exact union of per-segment rectangles, round joins and round caps, traced as
// one outline per side (inner joints pass through the joint vertex so that the
// nonzero winding number counts covering pieces).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <vector>

#include "flatten.h"
#include "scan_line.h"
#include "tk_internal.h"

namespace tk {
namespace {

using Pt = FPt;

struct BezPath {
    std::vector<uint8_t> verbs;
    std::vector<Pt> pts;
};

constexpr uint32_t kNone = TK_INVALID_INDEX;

inline Pt apply(const double *m, Pt p) {
    return Pt{m[0] * p.x + m[2] * p.y + m[4], m[1] * p.x + m[3] * p.y + m[5]};
}

int points_for_verb(uint8_t verb) {
    switch (verb) {
    case TK_PATH_MOVE:
    case TK_PATH_LINE: return 1;
    case TK_PATH_QUAD: return 2;
    case TK_PATH_CUBIC: return 3;
    case TK_PATH_CLOSE: return 0;
    default: return -1;
    }
}

bool parse_path(const uint8_t *verbs, size_t verb_count, const double *points, size_t point_count,
                BezPath *out) {
    size_t needed = 0;
    for (size_t i = 0; i < verb_count; ++i) {
        int n = points_for_verb(verbs[i]);
        if (n < 0) return false;
        if (i == 0 && verbs[i] != TK_PATH_MOVE) return false;
        needed += size_t(n);
    }
    if (needed != point_count) return false;
    out->verbs.assign(verbs, verbs + verb_count);
    out->pts.resize(point_count);
    for (size_t i = 0; i < point_count; ++i) {
        out->pts[i] = Pt{points[2 * i], points[2 * i + 1]};
        if (!std::isfinite(out->pts[i].x) || !std::isfinite(out->pts[i].y)) return false;
    }
    return true;
}

BezPath transform_path(const BezPath &path, const double *m) {
    BezPath out = path;
    for (Pt &p : out.pts) p = apply(m, p);
    return out;
}

// ---------------------------------------------------------------------------
// Exact curve bounds (kurbo's BezPath::bounding_box semantics: segment extrema).
// ---------------------------------------------------------------------------

struct Box {
    double x0 = std::numeric_limits<double>::infinity(), y0 = x0;
    double x1 = -std::numeric_limits<double>::infinity(), y1 = x1;
    bool empty() const { return !(x0 <= x1); }
    void add(Pt p) {
        x0 = std::min(x0, p.x);
        y0 = std::min(y0, p.y);
        x1 = std::max(x1, p.x);
        y1 = std::max(y1, p.y);
    }
};

void add_cubic_bounds(Box &box, Pt p0, Pt p1, Pt p2, Pt p3) {
    box.add(p0);
    box.add(p3);
    double ts[2];
    int n;
    cubic_axis_extrema(p0.x, p1.x, p2.x, p3.x, ts, &n);
    for (int i = 0; i < n; ++i) {
        double t = ts[i];
        box.add(Pt{cubic_coord(p0.x, p1.x, p2.x, p3.x, t), cubic_coord(p0.y, p1.y, p2.y, p3.y, t)});
    }
    cubic_axis_extrema(p0.y, p1.y, p2.y, p3.y, ts, &n);
    for (int i = 0; i < n; ++i) {
        double t = ts[i];
        box.add(Pt{cubic_coord(p0.x, p1.x, p2.x, p3.x, t), cubic_coord(p0.y, p1.y, p2.y, p3.y, t)});
    }
}

Box path_bounds(const BezPath &path) {
    Box box;
    size_t pi = 0;
    Pt last{0, 0};
    for (uint8_t verb : path.verbs) {
        switch (verb) {
        case TK_PATH_MOVE: last = path.pts[pi++]; break;
        case TK_PATH_LINE:
            box.add(last);
            box.add(path.pts[pi]);
            last = path.pts[pi++];
            break;
        case TK_PATH_QUAD: {
            Pt q0 = last, q1 = path.pts[pi], q2 = path.pts[pi + 1];
            Pt c1{q0.x + 2.0 / 3.0 * (q1.x - q0.x), q0.y + 2.0 / 3.0 * (q1.y - q0.y)};
            Pt c2{q2.x + 2.0 / 3.0 * (q1.x - q2.x), q2.y + 2.0 / 3.0 * (q1.y - q2.y)};
            add_cubic_bounds(box, q0, c1, c2, q2);
            last = q2;
            pi += 2;
            break;
        }
        case TK_PATH_CUBIC:
            add_cubic_bounds(box, last, path.pts[pi], path.pts[pi + 1], path.pts[pi + 2]);
            last = path.pts[pi + 2];
            pi += 3;
            break;
        default: break;
        }
    }
    return box;
}

// ---------------------------------------------------------------------------
// Flattening with the reference's contour rules
// (src/shared/path_flatten.rs:20-82): implicit close when the gap exceeds 1e-6,
// degenerate edges with |d|^2 <= 1e-24 dropped.
// ---------------------------------------------------------------------------

struct Polyline {
    std::vector<Pt> pts;
    bool closed = false;
};

// Curve flattening (flatten.h, shared with the device front end): quadratics by
// the parabola-integral method, cubics through quadratics at a tenth of the
// tolerance.
void flatten_quad_to(Pt p0, Pt p1, Pt p2, double tol, std::vector<Pt> &out) {
    flatten_quad(p0, p1, p2, tol, [&](Pt p) { out.push_back(p); });
}

void flatten_cubic_to(Pt p0, Pt p1, Pt p2, Pt p3, double tol, std::vector<Pt> &out) {
    flatten_cubic(p0, p1, p2, p3, tol, [&](Pt p) { out.push_back(p); });
}

// Flattens every contour into point lists; closed marks an explicit ClosePath.
std::vector<Polyline> flatten_contours(const BezPath &path, double tol) {
    std::vector<Polyline> out;
    size_t pi = 0;
    Pt last{0, 0};
    for (uint8_t verb : path.verbs) {
        switch (verb) {
        case TK_PATH_MOVE:
            out.emplace_back();
            last = path.pts[pi++];
            out.back().pts.push_back(last);
            break;
        case TK_PATH_LINE:
            last = path.pts[pi++];
            out.back().pts.push_back(last);
            break;
        case TK_PATH_QUAD:
            flatten_quad_to(last, path.pts[pi], path.pts[pi + 1], tol, out.back().pts);
            last = path.pts[pi + 1];
            pi += 2;
            break;
        case TK_PATH_CUBIC:
            flatten_cubic_to(last, path.pts[pi], path.pts[pi + 1], path.pts[pi + 2], tol, out.back().pts);
            last = path.pts[pi + 2];
            pi += 3;
            break;
        case TK_PATH_CLOSE:
            if (!out.empty()) {
                out.back().closed = true;
                last = out.back().pts.front();
            }
            break;
        default: break;
        }
    }
    return out;
}

bool is_non_degenerate(Pt a, Pt b) {
    double dx = a.x - b.x, dy = a.y - b.y;
    return dx * dx + dy * dy > 1.0e-12 * 1.0e-12;
}

bool is_open_contour_close(Pt start, Pt end) {
    double dx = start.x - end.x, dy = start.y - end.y;
    return dx * dx + dy * dy > 1.0e-6 * 1.0e-6;
}

void push_line(std::vector<tk_line> &out, uint32_t path_id, Pt a, Pt b) {
    if (!is_non_degenerate(a, b)) return;
    tk_line line;
    line.path_id = path_id;
    line.pad = 0.0f;
    line.p0[0] = float(a.x);
    line.p0[1] = float(a.y);
    line.p1[0] = float(b.x);
    line.p1[1] = float(b.y);
    out.push_back(line);
}

void flatten_to_lines(const BezPath &path, double tol, uint32_t path_id, std::vector<tk_line> &out) {
    for (const Polyline &contour : flatten_contours(path, tol)) {
        const auto &p = contour.pts;
        for (size_t i = 1; i < p.size(); ++i) push_line(out, path_id, p[i - 1], p[i]);
        if (p.empty()) continue;
        // Explicit ClosePath always emits the closing edge; an open contour is
        // closed only when its end is not already at its start.
        if (contour.closed || is_open_contour_close(p.front(), p.back()))
            push_line(out, path_id, p.back(), p.front());
    }
}

// ---------------------------------------------------------------------------
// Stroke outline (round joins, round caps).
// ---------------------------------------------------------------------------

struct OutlineWriter {
    BezPath path;
    void move(Pt p) {
        path.verbs.push_back(TK_PATH_MOVE);
        path.pts.push_back(p);
    }
    void line(Pt p) {
        path.verbs.push_back(TK_PATH_LINE);
        path.pts.push_back(p);
    }
    void close() { path.verbs.push_back(TK_PATH_CLOSE); }
};

// Appends the interior points of an arc around c of radius h from angle a0
// turning by delta (signed), with chord error at most tol.
void arc_points(std::vector<Pt> &out, Pt c, double h, double a0, double delta, double tol) {
    double step = M_PI;
    if (h > tol) step = 2.0 * std::acos(1.0 - tol / h);
    int n = std::max(1, int(std::ceil(std::fabs(delta) / step)));
    for (int i = 1; i < n; ++i) {
        double a = a0 + delta * double(i) / n;
        out.push_back(Pt{c.x + h * std::cos(a), c.y + h * std::sin(a)});
    }
}

// Intersection parameter of segments a0->a1 and b0->b1; false when parallel
// or when the crossing lies outside either segment.
bool segment_cross(Pt a0, Pt a1, Pt b0, Pt b1, Pt *x) {
    double rx = a1.x - a0.x, ry = a1.y - a0.y, sx = b1.x - b0.x, sy = b1.y - b0.y;
    double den = rx * sy - ry * sx;
    if (den == 0.0) return false;
    double qx = b0.x - a0.x, qy = b0.y - a0.y;
    double t = (qx * sy - qy * sx) / den, u = (qx * ry - qy * rx) / den;
    if (!(t >= 0.0 && t <= 1.0 && u >= 0.0 && u <= 1.0)) return false;
    *x = Pt{a0.x + t * rx, a0.y + t * ry};
    return true;
}

// Join and cap styles of a stroke (kurbo::Stroke: Join {Bevel, Miter, Round},
// Cap {Butt, Square, Round}, miter_limit).
struct StrokeStyle {
    uint32_t join = TK_JOIN_ROUND, start_cap = TK_CAP_ROUND, end_cap = TK_CAP_ROUND;
    double miter_limit = 4.0;
};

// Stroke outline of a polyline.
//
// The stroked region is the union of one rectangle per segment, a disc sector
// at every outer joint and a half disc at each open end.  Tracing the sum of
// those pieces' boundaries (all with one orientation) and cancelling shared
// half-edges leaves one chain per side: on the outer side of a joint the arc,
// on the inner side a detour through the joint vertex.  The nonzero winding of
// that outline counts covering pieces, so nonzero filling gives the union.
// Two reductions keep the line count down: an inner detour is replaced by the
// crossing of the two inner edges when that crossing lies on both (this only
// removes a loop lying inside both rectangles), and an outer arc whose miter
// is within the join tolerance of it is replaced by the miter point.  Miter and
// bevel joins replace the outer disc sector by the miter triangle, while the
// miter length stays within miter_limit times the stroke width (1 / cos of half
// the turn, as SVG and kurbo measure it), or else by nothing (the bevel).
void stroke_side(std::vector<Pt> &out, const std::vector<Pt> &p, const std::vector<Pt> &n, const std::vector<Pt> &d,
                 bool closed, double side, double h, double join_tol, const StrokeStyle &style) {
    size_t m = n.size();
    auto vert = [&](size_t i) { return p[i % p.size()]; };
    auto off = [&](size_t i, size_t seg) {
        Pt q = vert(i);
        return Pt{q.x + side * h * n[seg].x, q.y + side * h * n[seg].y};
    };
    size_t joints = closed ? m : m - 1;
    std::vector<Pt> pts;
    Pt start = off(0, 0);
    Pt cur = start;  // where the chain entered the current edge
    for (size_t i = 0; i < m; ++i) {
        Pt e = off(i + 1, i);
        if (i >= joints) {
            pts.push_back(e);
            break;
        }
        size_t j = (i + 1) % m;
        Pt c = vert(i + 1);
        double cr = d[i].x * d[j].y - d[i].y * d[j].x, dt = d[i].x * d[j].x + d[i].y * d[j].y;
        double ang = std::atan2(cr, dt);
        if (cr == 0.0 && dt < 0.0) ang = -M_PI;  // U-turn: the +n side is outer
        Pt sj = off(i + 1, j);
        bool outer = side > 0.0 ? ang < 0.0 : ang > 0.0;
        bool closing = closed && i + 1 == m;  // the joint at vertex 0 of a loop keeps the edge start exact
        if (ang == 0.0) {
            pts.push_back(e);
            cur = e;
        } else if (outer && style.join != TK_JOIN_ROUND) {
            double sweep = std::fabs(ang);
            double mx = side * (n[i].x + n[j].x), my = side * (n[i].y + n[j].y);
            double den = 1.0 + (n[i].x * n[j].x + n[i].y * n[j].y);
            pts.push_back(e);
            if (style.join == TK_JOIN_MITER && den > 0.0 && 1.0 / std::cos(0.5 * sweep) <= style.miter_limit)
                pts.push_back(Pt{c.x + h * mx / den, c.y + h * my / den});
            pts.push_back(sj);
            cur = sj;
        } else if (outer) {
            double sweep = std::fabs(ang);
            double mx = side * (n[i].x + n[j].x), my = side * (n[i].y + n[j].y);
            double den = 1.0 + (n[i].x * n[j].x + n[i].y * n[j].y);
            if (!closing && sweep < 0.5 * M_PI && h * (1.0 / std::cos(0.5 * sweep) - 1.0) <= join_tol && den > 0.0) {
                Pt mp{c.x + h * mx / den, c.y + h * my / den};
                pts.push_back(mp);
                cur = mp;
            } else {
                pts.push_back(e);
                double a0 = std::atan2(side * n[i].y, side * n[i].x);
                arc_points(pts, c, h, a0, side > 0.0 ? -sweep : sweep, join_tol);
                pts.push_back(sj);
                cur = sj;
            }
        } else {
            Pt x;
            if (!closing && segment_cross(cur, e, sj, off(i + 2, j), &x)) {
                pts.push_back(x);
                cur = x;
            } else {
                pts.push_back(e);
                pts.push_back(c);
                pts.push_back(sj);
                cur = sj;
            }
        }
    }
    if (closed && !pts.empty()) pts.pop_back();  // the closing joint ends on the loop's first point
    out.push_back(start);
    out.insert(out.end(), pts.begin(), pts.end());
}

void stroke_polyline(OutlineWriter &w, std::vector<Pt> pts, bool closed, double h, double tol,
                     const StrokeStyle &style) {
    std::vector<Pt> p;
    for (const Pt &q : pts)
        if (p.empty() || std::hypot(q.x - p.back().x, q.y - p.back().y) > 1e-9) p.push_back(q);
    if (closed && p.size() > 1 && std::hypot(p.front().x - p.back().x, p.front().y - p.back().y) <= 1e-9)
        p.pop_back();
    if (p.size() < 2) {
        if (p.size() == 1 && !closed && style.start_cap == TK_CAP_ROUND) {  // a dot: round caps make a disc
            Pt c = p[0];
            std::vector<Pt> ring{Pt{c.x + h, c.y}};
            arc_points(ring, c, h, 0.0, -2.0 * M_PI, tol);
            w.move(ring[0]);
            for (size_t i = 1; i < ring.size(); ++i) w.line(ring[i]);
            w.close();
        } else if (p.size() == 1 && !closed && style.start_cap == TK_CAP_SQUARE) {  // square caps: a square
            Pt c = p[0];
            w.move(Pt{c.x - h, c.y - h});
            w.line(Pt{c.x + h, c.y - h});
            w.line(Pt{c.x + h, c.y + h});
            w.line(Pt{c.x - h, c.y + h});
            w.close();
        }
        return;
    }
    size_t m = closed ? p.size() : p.size() - 1;
    std::vector<Pt> d(m), n(m);
    for (size_t i = 0; i < m; ++i) {
        Pt a = p[i], b = p[(i + 1) % p.size()];
        double len = std::hypot(b.x - a.x, b.y - a.y);
        d[i] = Pt{(b.x - a.x) / len, (b.y - a.y) / len};
        n[i] = Pt{-d[i].y, d[i].x};
    }
    std::vector<Pt> left, right;
    stroke_side(left, p, n, d, closed, 1.0, h, tol, style);
    stroke_side(right, p, n, d, closed, -1.0, h, tol, style);
    w.move(left[0]);
    for (size_t i = 1; i < left.size(); ++i) w.line(left[i]);
    if (closed) {
        w.close();
        w.move(right.back());
        for (size_t i = right.size() - 1; i-- > 0;) w.line(right[i]);
        w.close();
        return;
    }
    // End cap: from +n to -n through the direction of travel (butt: straight across).
    Pt pe = p.back();
    std::vector<Pt> cap;
    if (style.end_cap == TK_CAP_ROUND) {
        arc_points(cap, pe, h, std::atan2(n[m - 1].y, n[m - 1].x), -M_PI, tol);
    } else if (style.end_cap == TK_CAP_SQUARE) {
        const Pt dd = d[m - 1], nn = n[m - 1];
        cap.push_back(Pt{pe.x + h * (dd.x + nn.x), pe.y + h * (dd.y + nn.y)});
        cap.push_back(Pt{pe.x + h * (dd.x - nn.x), pe.y + h * (dd.y - nn.y)});
    }
    for (const Pt &q : cap) w.line(q);
    for (size_t i = right.size(); i-- > 0;) w.line(right[i]);
    // Start cap: from -n back to +n through the reverse direction.
    Pt ps = p.front();
    cap.clear();
    if (style.start_cap == TK_CAP_ROUND) {
        arc_points(cap, ps, h, std::atan2(-n[0].y, -n[0].x), -M_PI, tol);
    } else if (style.start_cap == TK_CAP_SQUARE) {
        const Pt dd = d[0], nn = n[0];
        cap.push_back(Pt{ps.x - h * (dd.x + nn.x), ps.y - h * (dd.y + nn.y)});
        cap.push_back(Pt{ps.x - h * (dd.x - nn.x), ps.y - h * (dd.y - nn.y)});
    }
    for (const Pt &q : cap) w.line(q);
    w.close();
}

BezPath stroke_outline(const BezPath &path, double width, double tol, const StrokeStyle &style = StrokeStyle{}) {
    OutlineWriter w;
    double h = 0.5 * width;
    if (!(h > 0.0)) return w.path;
    // The centreline gets three quarters of the tolerance, joins and caps the rest.
    for (const Polyline &c : flatten_contours(path, 0.75 * tol))
        stroke_polyline(w, c.pts, c.closed, h, 0.25 * tol, style);
    return w.path;
}

}  // namespace

// line_scanned_tile_count (src/shared/scan_line.rs:17-151); the arithmetic lives
// in scan_line.h, shared with the device front end.
uint32_t line_scanned_tile_count(const tk_line &line, uint32_t bx0, uint32_t by0, uint32_t bx1,
                                 uint32_t by1) {
    return scanned_tile_count(line.p0[0], line.p0[1], line.p1[0], line.p1[1], bx0, by0, bx1, by1);
}

namespace {

// ---------------------------------------------------------------------------
// Brush and SDF encoding (src/shared/brush.rs:249-274,638-767, gpu_sdf.rs:48-63)
// ---------------------------------------------------------------------------

uint32_t f2bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

float bits2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

uint32_t premul_f32_to_u32(const float c[4]) {
    // premul_f32_to_u32 (src/shared/pixel.rs:14-21): (c * 255 + 0.5) as u8, saturating.
    uint32_t out = 0;
    for (int i = 0; i < 4; ++i) {
        float v = c[i] * 255.0f + 0.5f;
        uint32_t b = v <= 0.0f ? 0u : (v >= 255.0f ? 255u : uint32_t(v));
        out |= b << (8 * i);
    }
    return out;
}

uint32_t premultiply_pack(const float rgba[4]) {
    float p[4] = {rgba[0] * rgba[3], rgba[1] * rgba[3], rgba[2] * rgba[3], rgba[3]};
    return premul_f32_to_u32(p);
}

size_t quantize_ramp_size(float span, size_t stop_count) {
    size_t span_samples = size_t(std::ceil(std::fmax(span, 1.0f) * 2.0f));
    size_t stop_samples = (std::max<size_t>(stop_count, 2) - 1) * 16;
    size_t target = std::max(span_samples, stop_samples);
    target = std::min<size_t>(std::max<size_t>(target, 64), 4096);
    size_t p = 1;
    while (p < target) p <<= 1;
    return std::min<size_t>(p, 4096);
}

// build_ramp (src/shared/brush.rs:638-687): sRGB interpolation between sorted
// stops, premultiplied (peniko's default) or in straight alpha and premultiplied
// after (InterpolationAlphaSpace::Unpremultiplied); peniko's exact colour
// arithmetic is not restated.
std::vector<uint32_t> build_ramp(std::vector<tk_gradient_stop> stops, size_t ramp_size, bool unpremultiplied) {
    ramp_size = std::max<size_t>(ramp_size, 2);
    std::vector<uint32_t> ramp(ramp_size, 0u);
    if (stops.empty()) return ramp;
    std::stable_sort(stops.begin(), stops.end(),
                     [](const tk_gradient_stop &a, const tk_gradient_stop &b) { return a.offset < b.offset; });
    for (size_t ix = 0; ix < ramp_size; ++ix) {
        float t = float(ix) / float(ramp_size - 1);
        size_t upper = 0;
        while (upper < stops.size() && stops[upper].offset < t) ++upper;
        tk_gradient_stop left, right;
        if (upper == 0) left = right = stops[0];
        else if (upper >= stops.size()) left = right = stops.back();
        else { left = stops[upper - 1]; right = stops[upper]; }
        float span = right.offset - left.offset;
        float lt = std::fabs(span) <= 1.1920929e-7f ? 0.0f : std::fmin(std::fmax((t - left.offset) / span, 0.0f), 1.0f);
        float c[4];
        if (unpremultiplied) {
            float u[4];
            for (int k = 0; k < 4; ++k) u[k] = left.rgba[k] + (right.rgba[k] - left.rgba[k]) * lt;
            c[0] = u[0] * u[3];
            c[1] = u[1] * u[3];
            c[2] = u[2] * u[3];
            c[3] = u[3];
        } else {
            float lp[4] = {left.rgba[0] * left.rgba[3], left.rgba[1] * left.rgba[3], left.rgba[2] * left.rgba[3],
                           left.rgba[3]};
            float rp[4] = {right.rgba[0] * right.rgba[3], right.rgba[1] * right.rgba[3],
                           right.rgba[2] * right.rgba[3], right.rgba[3]};
            for (int k = 0; k < 4; ++k) c[k] = lp[k] + (rp[k] - lp[k]) * lt;
        }
        ramp[ix] = premul_f32_to_u32(c);
    }
    return ramp;
}

// Ramp sizes (src/shared/brush.rs:503-535).
size_t radial_ramp_size(const tk_brush *b, size_t stop_count) {
    float center_span = std::hypot(b->end[0] - b->start[0], b->end[1] - b->start[1]);
    float radius_span = std::fabs(b->start_radius) + std::fabs(b->end_radius);
    float span = std::fmax(center_span + radius_span, std::fabs(b->end_radius - b->start_radius)) * 1.0f;
    return quantize_ramp_size(span, stop_count);
}
size_t sweep_ramp_size(float start_angle, float end_angle, size_t stop_count) {
    float turns = std::fmax(std::fabs(end_angle - start_angle) / 6.2831855f, 0.25f);
    return quantize_ramp_size(turns * 256.0f, stop_count);
}

// f64::rem_euclid.
double rem_euclid_f64(double a, double b) {
    double r = std::fmod(a, b);
    return r < 0.0 ? r + std::fabs(b) : r;
}

// The scope word of a canvas image's id (ImageResourceId::scene).
constexpr uint32_t kImageScopeScene = 1u;

// Encodes a brush into its 9 header + 12 parameter words and its payload
// (push_encoded_brush, src/shared/brush.rs:249-345, after Surface::physical_brush,
// src/canvas/surface.rs:395-420, has applied the canvas scale).
bool encode_brush(const tk_brush *brush, double scale, std::vector<uint32_t> &blob, uint32_t *offset,
                  uint32_t *len, uint32_t image_count) {
    uint32_t data[TK_BRUSH_HEADER_WORDS] = {TK_BRUSH_SOLID, 0u, 0u, 0u, 0u, 0u, 0u, 255u, 0u};
    uint32_t params[TK_BRUSH_PARAM_WORDS] = {};
    std::vector<uint32_t> payload;
    const uint32_t kind = brush ? brush->kind : TK_BRUSH_SOLID;
    const float scale_f = float(scale), inv = float(1.0 / scale);
    // scale_brush_transform(the brush's transform, 1 / scale) (src/canvas.rs:203-212).
    float xf[6] = {inv, 0.0f, 0.0f, inv, 0.0f, 0.0f};
    if (brush && (brush->kind == TK_BRUSH_LINEAR || brush->kind == TK_BRUSH_RADIAL) &&
        (brush->gradient_flags & TK_GRADIENT_TRANSFORM)) {
        for (int i = 0; i < 6; ++i)
            if (!std::isfinite(brush->gradient_transform[i])) return false;
        for (int i = 0; i < 4; ++i) xf[i] = brush->gradient_transform[i] * inv;
        xf[4] = brush->gradient_transform[4];
        xf[5] = brush->gradient_transform[5];
    }
    auto ramp_of = [&](size_t size) {
        std::vector<tk_gradient_stop> stops(brush->stops, brush->stops + brush->stop_count);
        return build_ramp(stops, size, (brush->gradient_flags & TK_GRADIENT_UNPREMULTIPLIED) != 0u);
    };
    if (kind == TK_BRUSH_SOLID) {
        float transparent[4] = {0, 0, 0, 0};
        data[4] = premultiply_pack(brush ? brush->rgba : transparent);
    } else if (kind == TK_BRUSH_LINEAR) {
        if (brush->extend > 2u) return false;
        float span = std::hypot(brush->end[0] - brush->start[0], brush->end[1] - brush->start[1]);
        payload = ramp_of(quantize_ramp_size(span, brush->stop_count));
        data[1] = brush->extend;
        params[0] = f2bits(brush->start[0]);
        params[1] = f2bits(brush->start[1]);
        params[2] = f2bits(brush->end[0]);
        params[3] = f2bits(brush->end[1]);
        for (int i = 0; i < 6; ++i) params[4 + i] = f2bits(xf[i]);
    } else if (kind == TK_BRUSH_RADIAL) {
        if (brush->extend > 2u) return false;
        payload = ramp_of(radial_ramp_size(brush, brush->stop_count));
        data[1] = brush->extend;
        params[0] = f2bits(brush->start[0]);
        params[1] = f2bits(brush->start[1]);
        params[2] = f2bits(brush->end[0]);
        params[3] = f2bits(brush->end[1]);
        params[4] = f2bits(brush->start_radius);
        params[5] = f2bits(brush->end_radius);
        for (int i = 0; i < 6; ++i) params[6 + i] = f2bits(xf[i]);
    } else if (kind == TK_BRUSH_SWEEP) {
        if (brush->extend > 2u) return false;
        const double start = double(brush->start_angle), end = double(brush->end_angle);
        if (!std::isfinite(start) || !std::isfinite(end)) return false;  // the reference panics
        // One reduction in f64 to [-pi, pi]; the signed span is kept separately
        // (Brush::from_gradient_with_ramp_size, src/shared/brush.rs:200-226).
        const double tau = 6.283185307179586, pi = 3.141592653589793;
        double orientation = rem_euclid_f64(start, tau);
        if (orientation > pi) orientation -= tau;
        payload = ramp_of(sweep_ramp_size(brush->start_angle, brush->end_angle, brush->stop_count));
        data[1] = brush->extend;
        params[0] = f2bits(brush->start[0] * scale_f);
        params[1] = f2bits(brush->start[1] * scale_f);
        params[2] = f2bits(float(orientation));
        params[3] = f2bits(float(end - start));
    } else if (kind == TK_BRUSH_FOUR_CORNER) {
        for (int i = 0; i < 4; ++i) params[i] = f2bits(brush->bounds[i] * scale_f);
        for (int i = 0; i < 4; ++i) payload.push_back(premultiply_pack(brush->corners[i]));
    } else if (kind == TK_BRUSH_IMAGE) {
        // Brush::Pattern (src/shared/brush.rs:325-340): extend, the transform scaled
        // by 1 / scale (Surface::physical_brush, src/canvas/surface.rs:428-431),
        // opacity, sampling, and the three-word image id the atlas placement
        // replaces at encoding (resource_brush_header, src/shared/gpu_brush.rs:92-127).
        if (brush->extend > 2u || brush->sampling > 1u || brush->image >= image_count) return false;
        for (float v : brush->pattern_transform)
            if (!std::isfinite(v)) return false;
        const float inv_scale = 1.0f / scale_f;
        data[1] = brush->extend;
        data[7] = std::min(brush->opacity, 255u);
        data[8] = brush->sampling;
        for (int i = 0; i < 4; ++i) params[i] = f2bits(brush->pattern_transform[i] * inv_scale);
        params[4] = f2bits(brush->pattern_transform[4]);
        params[5] = f2bits(brush->pattern_transform[5]);
        payload = {kImageScopeScene, brush->image, 0u};
    } else {
        return false;
    }
    if (kind != TK_BRUSH_SOLID) {
        data[0] = kind;
        data[2] = TK_BRUSH_HEADER_WORDS + TK_BRUSH_PARAM_WORDS;  // set_local_payload
        data[3] = uint32_t(payload.size());
    }
    *offset = uint32_t(blob.size());
    blob.insert(blob.end(), data, data + TK_BRUSH_HEADER_WORDS);
    blob.insert(blob.end(), params, params + TK_BRUSH_PARAM_WORDS);
    blob.insert(blob.end(), payload.begin(), payload.end());
    *len = uint32_t(blob.size()) - *offset;
    return true;
}

tk_affine identity_affine() { return tk_affine{1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f}; }

struct TileBox {
    uint32_t x0, y0, x1, y1;
    uint64_t area() const { return uint64_t(x1 > x0 ? x1 - x0 : 0) * uint64_t(y1 > y0 ? y1 - y0 : 0); }
};

// PixelBounds::tile_bbox (src/shared/bounds.rs:113-128), shared with the device front end.
TileBox tile_bbox(const tk_pixel_bounds &b, uint32_t tw, uint32_t th) {
    uint32_t t[4];
    tile_box_of(b.x0, b.y0, b.x1, b.y1, tw, th, t);
    return TileBox{t[0], t[1], t[2], t[3]};
}

int32_t sat_i32(double v) {
    if (!(v == v)) return 0;
    if (v <= -2147483648.0) return INT32_MIN;
    if (v >= 2147483647.0) return INT32_MAX;
    return int32_t(v);
}

}  // namespace
}  // namespace tk

// ---------------------------------------------------------------------------
// Canvas
// ---------------------------------------------------------------------------

struct tk_canvas {
    uint32_t width = 0, height = 0;  // physical
    double scale = 1.0;
    uint32_t tiles_w = 0, tiles_h = 0;

    std::vector<tk_line> lines;
    std::vector<tk_path_record> paths;
    std::vector<tk_draw_record> draws;
    std::vector<uint32_t> sdf_blob, shadow_blob, brush_blob;
    uint32_t backdrop_pool = 0;
    uint64_t segment_total = 0;

    struct Cmd {
        uint8_t kind;  // 0 draw, 1 clip layer
        uint32_t draw;
        uint32_t children;
    };
    std::vector<std::vector<Cmd>> lists{1};
    std::vector<uint32_t> open{0};

    // Encoded plan (valid until the next push).
    std::vector<uint32_t> paint;
    std::vector<uint32_t> draw_batch_ids;
    std::vector<tk_layer_record> layer_records;
    std::vector<tk_batch_record> batch_records;
    std::vector<tk_tile_draw_record> tile_draws;
    std::vector<uint32_t> tile_draw_indices;
    std::vector<tk_scan_chunk> scan_chunks;
    std::vector<tk_scan_chunk_range> scan_chunk_ranges;
    std::vector<uint32_t> cumsum_offsets, cumsum_lens, cumsum_row_starts, cumsum_row_ends;

    // Images the canvas registered (premultiplied RGBA8), and the atlas the
    // encoding packs them into.
    struct Image {
        uint32_t width, height;
        std::vector<uint32_t> pixels;
    };
    std::vector<Image> images;
    std::vector<uint32_t> atlas;
    uint32_t atlas_width = 0, atlas_height = 0, atlas_pages = 0;
};

namespace tk {
namespace {

void device_matrix(const tk_canvas *c, const double *transform, double out[6]) {
    double t[6] = {1, 0, 0, 1, 0, 0};
    if (transform) std::memcpy(t, transform, sizeof(t));
    for (int i = 0; i < 6; ++i) out[i] = t[i] * c->scale;
}

// push_path_inner_with_tag (src/canvas/paths.rs:65-156) for an already
// transformed (physical) path.
uint32_t push_physical_path(tk_canvas *c, const BezPath &path, double device_tol, uint32_t rule,
                            uint32_t tag, const tk_brush *brush, bool emit_command, tk_status *status) {
    uint32_t line_start = uint32_t(c->lines.size());
    uint32_t path_id = uint32_t(c->paths.size());
    flatten_to_lines(path, device_tol, path_id, c->lines);
    uint32_t line_count = uint32_t(c->lines.size()) - line_start;
    Box box = path_bounds(path);
    tk_pixel_bounds pb{0, 0, 0, 0};
    if (!box.empty()) {
        pb.x0 = sat_i32(std::floor(box.x0));
        pb.y0 = sat_i32(std::floor(box.y0));
        pb.x1 = sat_i32(std::ceil(box.x1));
        pb.y1 = sat_i32(std::ceil(box.y1));
    }
    TileBox tb = tile_bbox(pb, c->tiles_w, c->tiles_h);
    uint32_t stride = tb.x1 > tb.x0 ? tb.x1 - tb.x0 : 0;
    uint32_t height = tb.y1 > tb.y0 ? tb.y1 - tb.y0 : 0;
    uint32_t capacity = 0;
    for (uint32_t i = line_start; i < line_start + line_count; ++i) {
        uint64_t next = uint64_t(capacity) + line_scanned_tile_count(c->lines[i], tb.x0, tb.y0, tb.x1, tb.y1);
        capacity = next > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(next);
    }
    tk_path_record rec{};
    rec.path_id = path_id;
    rec.line_count = line_count;
    rec.line_start = line_start;
    rec.flags = 0;
    rec.data_offset = c->backdrop_pool;
    rec.data_len = stride * height;
    rec.tile_x0 = tb.x0;
    rec.tile_y0 = tb.y0;
    rec.tile_x1 = tb.x1;
    rec.tile_y1 = tb.y1;
    rec.segment_start = uint32_t(c->segment_total);
    rec.segment_capacity = capacity;
    rec.segment_count = 0;
    rec.transform = identity_affine();
    c->backdrop_pool += stride * height;
    c->segment_total += capacity;
    c->paths.push_back(rec);

    uint32_t brush_offset = 0, brush_len = 0;
    if (!encode_brush(brush, c->scale, c->brush_blob, &brush_offset, &brush_len, uint32_t(c->images.size()))) {
        *status = TK_ERR_INVALID;
        return kNone;
    }
    tk_draw_record d{};
    d.path_id = path_id;
    d.glyph_run_id = kNone;
    d.sdf_offset = kNone;
    d.sdf_len = 0;
    d.sdf_shadow_offset = kNone;
    d.sdf_shadow_len = 0;
    d.brush_offset = brush_offset;
    d.brush_len = brush_len;
    d.tag = tag;
    d.fill_rule = rule;
    d.pixel_bounds = pb;
    d.local_pixel_bounds = pb;
    d.solid_rect = 0;
    d.transform = identity_affine();
    d.inverse_transform = identity_affine();
    uint32_t draw_ix = uint32_t(c->draws.size());
    c->draws.push_back(d);
    if (emit_command) c->lists[c->open.back()].push_back(tk_canvas::Cmd{0, draw_ix, 0});
    *status = TK_OK;
    return draw_ix;
}

// push_physical_sdf_record (src/canvas/paths.rs:207-250) for an SdfRect.
uint32_t push_sdf_rect(tk_canvas *c, double x0, double y0, double x1, double y1, const float radii[4],
                       const tk_brush *brush, uint32_t tag, bool emit_command, tk_status *status) {
    double s = c->scale;
    double ax0 = std::min(x0, x1) * s, ay0 = std::min(y0, y1) * s;
    double ax1 = std::max(x0, x1) * s, ay1 = std::max(y0, y1) * s;
    uint32_t sdf_offset = uint32_t(c->sdf_blob.size());
    float r[4] = {0, 0, 0, 0};
    if (radii) for (int i = 0; i < 4; ++i) r[i] = float(radii[i] * s);
    c->sdf_blob.push_back(TK_SDF_KIND_RECT);
    float coords[4] = {float(ax0), float(ay0), float(ax1), float(ay1)};
    for (float v : coords) c->sdf_blob.push_back(f2bits(v));
    for (float v : r) c->sdf_blob.push_back(f2bits(v));
    for (int i = 0; i < 8; ++i) c->sdf_blob.push_back(0u);  // stroke[4], shadow[4]
    uint32_t brush_offset = 0, brush_len = 0;
    if (!encode_brush(brush, s, c->brush_blob, &brush_offset, &brush_len, uint32_t(c->images.size()))) {
        *status = TK_ERR_INVALID;
        return kNone;
    }
    tk_pixel_bounds pb{sat_i32(std::floor(ax0)), sat_i32(std::floor(ay0)), sat_i32(std::ceil(ax1)),
                       sat_i32(std::ceil(ay1))};
    tk_draw_record d{};
    d.path_id = kNone;
    d.glyph_run_id = kNone;
    d.sdf_offset = sdf_offset;
    d.sdf_len = TK_SDF_RECORD_WORDS;
    d.sdf_shadow_offset = kNone;
    d.sdf_shadow_len = 0;
    d.brush_offset = brush_offset;
    d.brush_len = brush_len;
    d.tag = tag;
    d.fill_rule = TK_FILL_NONZERO;
    d.pixel_bounds = pb;
    d.local_pixel_bounds = pb;
    d.solid_rect = 0;
    d.transform = identity_affine();
    d.inverse_transform = identity_affine();
    uint32_t draw_ix = uint32_t(c->draws.size());
    c->draws.push_back(d);
    if (emit_command) c->lists[c->open.back()].push_back(tk_canvas::Cmd{0, draw_ix, 0});
    *status = TK_OK;
    return draw_ix;
}

void open_clip_layer(tk_canvas *c, uint32_t draw) {
    uint32_t child = uint32_t(c->lists.size());
    c->lists.emplace_back();
    c->lists[c->open.back()].push_back(tk_canvas::Cmd{1, draw, child});
    c->open.push_back(child);
}

// axis_aligned_rect_path (src/canvas/layers.rs:286+): a closed 4/5-point line
// loop whose edges are axis-aligned.
bool axis_aligned_rect(const BezPath &path, double *r) {
    std::vector<Pt> p;
    bool closed = false;
    size_t pi = 0;
    for (uint8_t v : path.verbs) {
        if (v == TK_PATH_MOVE && p.empty()) p.push_back(path.pts[pi++]);
        else if (v == TK_PATH_LINE && !closed) p.push_back(path.pts[pi++]);
        else if (v == TK_PATH_CLOSE && !closed) closed = true;
        else return false;
    }
    if (!closed || p.size() < 4 || p.size() > 5) return false;
    if (p.size() == 5) {
        if (p[4].x != p[0].x || p[4].y != p[0].y) return false;
        p.pop_back();
    }
    for (size_t i = 0; i < 4; ++i) {
        Pt a = p[i], b = p[(i + 1) % 4];
        if (a.x != b.x && a.y != b.y) return false;
        if (a.x == b.x && a.y == b.y) return false;
    }
    double x0 = p[0].x, x1 = p[0].x, y0 = p[0].y, y1 = p[0].y;
    for (const Pt &q : p) {
        x0 = std::min(x0, q.x);
        x1 = std::max(x1, q.x);
        y0 = std::min(y0, q.y);
        y1 = std::max(y1, q.y);
    }
    for (const Pt &q : p)
        if ((q.x != x0 && q.x != x1) || (q.y != y0 && q.y != y1)) return false;
    r[0] = x0;
    r[1] = y0;
    r[2] = x1;
    r[3] = y1;
    return true;
}

// ---------------------------------------------------------------------------
// Compile + plan
// ---------------------------------------------------------------------------

struct PlannedBatch {
    std::vector<uint32_t> draws;
    std::vector<uint32_t> stack;  // clip draw ids, outermost first
};

// ---------------------------------------------------------------------------
// Early occlusion of tile bins (src/shared/gpu_plan/tile_bins/occlusion.rs,
// occlusion/geometry.rs, occlusion/transient.rs): before coarse, an immediate
// canvas drops the tile references of draws that a later opaque solid SDF
// rectangle of the same DrawBatch proves hidden.
// ---------------------------------------------------------------------------

TileBox box_intersection(const TileBox &a, const TileBox &b) {
    return TileBox{std::max(a.x0, b.x0), std::max(a.y0, b.y0), std::min(a.x1, b.x1), std::min(a.y1, b.y1)};
}

// rect_tiles (occlusion/geometry.rs:79-135): the whole tiles inside the central
// rectangle of an axis-aligned SDF rect (rounded corners and the AA edge
// excluded), with a 1/32 px guard when any input is fractional.  Empty for a
// path, glyph, shadowed or non-translating draw, or a non-rect SDF.  `s` is an
// encoded scene (its paint blob begins with the SDF records).
TileBox rect_tiles(const tk_scene &s, const tk_draw_record &d) {
    const TileBox none{0, 0, 0, 0};
    if (d.path_id != kNone || d.glyph_run_id != kNone || d.sdf_shadow_offset != kNone) return none;
    if (!(d.transform.a == 1.0f && d.transform.b == 0.0f && d.transform.c == 0.0f && d.transform.d == 1.0f))
        return none;
    // encoded_sdf (src/shared/gpu_sdf.rs:290-304): exactly one 17-word record.
    if (d.sdf_offset == kNone || d.sdf_len != TK_SDF_RECORD_WORDS ||
        uint64_t(d.sdf_offset) + d.sdf_len > s.paint_sdf_shadow_base)
        return none;
    const uint32_t *w = s.paint + d.sdf_offset;
    if (w[0] != TK_SDF_KIND_RECT) return none;
    float v[10];  // coords x0 y0 x1 y1, radii tl tr bl br, translation e f
    for (int i = 0; i < 8; ++i) v[i] = bits2f(w[1 + i]);
    v[8] = d.transform.e;
    v[9] = d.transform.f;
    bool integral = true;
    for (float x : v) {
        if (!std::isfinite(x) || std::fabs(x) > 65536.0f) return none;
        if (x - std::trunc(x) != 0.0f) integral = false;
    }
    const double x0 = std::min<double>(v[0], v[2]), x1 = std::max<double>(v[0], v[2]);
    const double y0 = std::min<double>(v[1], v[3]), y1 = std::max<double>(v[1], v[3]);
    const double max_radius = std::max(std::min(x1 - x0, y1 - y0) * 0.5, 0.0);
    auto radius = [&](float r) { return std::min(std::max(double(r), 0.0), max_radius); };
    const double tl = radius(v[4]), tr = radius(v[5]), bl = radius(v[6]), br = radius(v[7]);
    const double guard = integral ? 0.0 : 1.0 / 32.0;
    const double tx = v[8], ty = v[9], tile = double(TK_TILE_SIZE);
    const TileBox bb = tile_bbox(d.pixel_bounds, s.tiles_width, s.tiles_height);
    auto whole = [](double t) { return uint32_t(std::max(t, 0.0)); };
    TileBox r;
    r.x0 = std::max(whole(std::ceil((x0 + std::max(tl, bl) + tx + guard) / tile)), bb.x0);
    r.y0 = std::max(whole(std::ceil((y0 + std::max(tl, tr) + ty + guard) / tile)), bb.y0);
    r.x1 = std::min(whole(std::floor((x1 - std::max(tr, br) + tx - guard) / tile)), bb.x1);
    r.y1 = std::min(whole(std::floor((y1 - std::max(bl, br) + ty - guard) / tile)), bb.y1);
    return r;
}

// opaque_tiles (occlusion/geometry.rs:60-77): rect_tiles of a Brush draw whose
// brush is solid with packed alpha 255.
TileBox opaque_tiles(const tk_scene &s, const tk_draw_record &d) {
    const TileBox none{0, 0, 0, 0};
    if (d.tag != TK_DRAW_BRUSH || d.brush_len < 5u) return none;
    if (uint64_t(d.brush_offset) + 5u > uint64_t(s.paint_words) - s.paint_brush_base) return none;
    const uint32_t *b = s.paint + s.paint_brush_base + d.brush_offset;
    if (b[0] != TK_BRUSH_SOLID || (b[4] >> 24) != 255u) return none;
    return rect_tiles(s, d);
}

}  // namespace

// Early occlusion (Occlusion::update, tile_bins/occlusion.rs:48-160, and
// for_each_visible, occlusion/transient.rs:9-66) as one rectangle per draw:
// the whole tiles a later opaque solid SDF rect of the same DrawBatch hides.
// A batch takes part when every member is a Brush or PathGlyph draw; its first
// draw never occludes; a proof is cut to the tiles its batch's clip rectangles
// cover; walking the batch in reverse painter order, the largest proof seen so
// far (a strictly larger one replaces it, never a union) is the cover of the
// draws before it.  Painter order is ascending draw id over the batched draws
// (the recorder lays batches out in recording order).  Returns whether any
// draw has a nonempty proof.
bool draw_covers(const tk_scene &s, std::vector<uint32_t> &cover) {
    const uint32_t D = s.draw_count, Q = s.batch_count;
    cover.assign(4 * size_t(D), 0u);
    struct BatchVis {
        uint32_t first = kNone;
        bool enabled = true, has_occluder = false;
        TileBox clip{0, 0, 0, 0};
    };
    std::vector<BatchVis> vis(Q);
    auto batch_of = [&](uint32_t d) { return s.draw_batch_ids[d] < Q ? s.draw_batch_ids[d] : kNone; };
    for (uint32_t d = 0; d < D; ++d) {
        const uint32_t b = batch_of(d);
        if (b == kNone) continue;
        vis[b].enabled = vis[b].enabled && (s.draws[d].tag == TK_DRAW_BRUSH || s.draws[d].tag == TK_DRAW_PATH_GLYPH);
        if (vis[b].first == kNone) vis[b].first = d;
    }
    std::vector<TileBox> opaque(D, TileBox{0, 0, 0, 0});
    for (uint32_t d = 0; d < D; ++d) {
        const uint32_t b = batch_of(d);
        if (b == kNone || !vis[b].enabled || vis[b].first == d) continue;
        opaque[d] = opaque_tiles(s, s.draws[d]);
        if (opaque[d].area() != 0) vis[b].has_occluder = true;
    }
    for (uint32_t b = 0; b < Q; ++b) {  // batch_clips (occlusion/geometry.rs:17-56)
        if (!vis[b].has_occluder) continue;
        TileBox clip{0, 0, s.tiles_width, s.tiles_height};
        for (uint32_t l = s.batches[b].layer_start; l < s.batches[b].layer_end && l < s.layer_count; ++l)
            if (s.layers[l].draw < D) clip = box_intersection(clip, rect_tiles(s, s.draws[s.layers[l].draw]));
        vis[b].clip = clip;
    }
    bool occluded = false;
    for (uint32_t d = 0; d < D; ++d) {
        const uint32_t b = batch_of(d);
        if (b == kNone || !vis[b].enabled) continue;
        opaque[d] = box_intersection(opaque[d], vis[b].clip);
        if (opaque[d].area() != 0) occluded = true;
    }
    uint32_t batch = kNone;
    TileBox cur{0, 0, 0, 0};
    for (uint32_t d = D; d-- > 0;) {
        const uint32_t id = batch_of(d);
        if (id == kNone) continue;  // not drawn by itself (a clip's own draw)
        const uint32_t vb = vis[id].enabled ? id : kNone;
        if (batch != vb || batch == kNone) cur = TileBox{0, 0, 0, 0};
        batch = vb;
        cover[4 * size_t(d)] = cur.x0;
        cover[4 * size_t(d) + 1] = cur.y0;
        cover[4 * size_t(d) + 2] = cur.x1;
        cover[4 * size_t(d) + 3] = cur.y1;
        if (vb != kNone && opaque[d].area() > cur.area()) cur = opaque[d];
    }
    return occluded;
}

namespace {

void compile_into(const tk_canvas *c, uint32_t list, std::vector<uint32_t> &stack,
                  std::vector<PlannedBatch> &ops) {
    std::vector<uint32_t> pending;
    auto flush = [&]() {
        if (pending.empty()) return;
        ops.push_back(PlannedBatch{pending, stack});
        pending.clear();
    };
    for (const tk_canvas::Cmd &cmd : c->lists[list]) {
        if (cmd.kind == 0) {
            pending.push_back(cmd.draw);
        } else {
            flush();
            stack.push_back(cmd.draw);
            compile_into(c, cmd.children, stack, ops);
            stack.pop_back();
        }
    }
    flush();
}

// The atlas of the canvas's images and their placement in the encoded paint.
// Images stack down pages of atlas_width (the widest image) by atlas_height (at
// least the tallest image, else 4096 rows) texels, each at x = 0 and the next
// free row; every image brush header then gets its image's placement in words
// 2-6 (x, y, page, width, height), as resource_brush_header patches a header
// from the renderer's resource table (src/shared/gpu_brush.rs:92-127).  Only
// the encoded paint changes: the brush blob keeps the image ids.
void pack_images(tk_canvas *c) {
    c->atlas.clear();
    c->atlas_width = c->atlas_height = c->atlas_pages = 0;
    if (c->images.empty()) return;
    struct Place {
        uint32_t x, y, page;
    };
    std::vector<Place> place(c->images.size());
    uint32_t width = 0, tallest = 0;
    for (const auto &im : c->images) {
        width = std::max(width, im.width);
        tallest = std::max(tallest, im.height);
    }
    const uint32_t height = std::max<uint32_t>(tallest, 4096u);
    uint32_t page = 0, row = 0, used = 0;
    for (size_t i = 0; i < c->images.size(); ++i) {
        if (row + c->images[i].height > height) {
            ++page;
            row = 0;
        }
        place[i] = Place{0u, row, page};
        row += c->images[i].height;
        used = std::max(used, row);
    }
    const uint32_t pages = page + 1;
    // a single page is cut to the rows it uses
    const uint32_t page_height = pages == 1 ? used : height;
    c->atlas.assign(size_t(pages) * page_height * width, 0u);
    for (size_t i = 0; i < c->images.size(); ++i) {
        const auto &im = c->images[i];
        for (uint32_t y = 0; y < im.height; ++y)
            std::memcpy(&c->atlas[(size_t(place[i].page) * page_height + place[i].y + y) * width],
                        &im.pixels[size_t(y) * im.width], size_t(im.width) * 4);
    }
    c->atlas_width = width;
    c->atlas_height = page_height;
    c->atlas_pages = pages;
    const size_t brush_base = c->sdf_blob.size() + c->shadow_blob.size();
    for (const tk_draw_record &d : c->draws) {
        if (d.brush_offset == kNone) continue;
        uint32_t *h = &c->paint[brush_base + d.brush_offset];
        // a header two draws share is patched once (its x word is then 0, never
        // the payload offset)
        if (h[0] != TK_BRUSH_IMAGE || h[2] != TK_BRUSH_HEADER_WORDS + TK_BRUSH_PARAM_WORDS) continue;
        const uint32_t id = h[h[2] + 1u];
        if (id < c->images.size()) {
            const uint32_t v[5] = {place[id].x, place[id].y, place[id].page, c->images[id].width, c->images[id].height};
            std::memcpy(h + 2, v, sizeof(v));
        } else {  // absent: a transparent pattern
            h[0] = TK_BRUSH_PATTERN;
            const uint32_t v[5] = {TK_BRUSH_HEADER_WORDS, 0u, 0u, 0u, 0u};
            std::memcpy(h + 2, v, sizeof(v));
        }
    }
}

tk_status encode(tk_canvas *c, uint32_t clear_color, tk_scene *out) {
    const uint32_t tw = c->tiles_w, th = c->tiles_h;
    const uint64_t tiles = uint64_t(tw) * th;

    // ExecPlan: batches in painter order, adjacent equal stacks coalesced.
    std::vector<PlannedBatch> ops;
    std::vector<uint32_t> stack;
    compile_into(c, 0, stack, ops);
    std::vector<PlannedBatch> batches;
    for (PlannedBatch &op : ops) {
        if (!batches.empty() && batches.back().stack == op.stack) {
            batches.back().draws.insert(batches.back().draws.end(), op.draws.begin(), op.draws.end());
        } else {
            batches.push_back(std::move(op));
        }
    }
    c->draw_batch_ids.assign(c->draws.size(), kNone);
    c->layer_records.clear();
    c->batch_records.clear();
    std::vector<uint32_t> draw_order;
    uint32_t max_depth = 0;
    for (size_t b = 0; b < batches.size(); ++b) {
        tk_batch_record rec;
        rec.layer_start = uint32_t(c->layer_records.size());
        for (uint32_t clip : batches[b].stack) c->layer_records.push_back(tk_layer_record{TK_LAYER_CLIP, clip, 0u});
        rec.layer_end = uint32_t(c->layer_records.size());
        max_depth = std::max<uint32_t>(max_depth, uint32_t(batches[b].stack.size()));
        c->batch_records.push_back(rec);
        for (uint32_t d : batches[b].draws) {
            c->draw_batch_ids[d] = uint32_t(b);
            draw_order.push_back(d);
        }
    }

    // Paint blob = sdf || sdf_shadow || brush (src/render/upload/paint.rs:112-113).
    c->paint.clear();
    c->paint.insert(c->paint.end(), c->sdf_blob.begin(), c->sdf_blob.end());
    c->paint.insert(c->paint.end(), c->shadow_blob.begin(), c->shadow_blob.end());
    c->paint.insert(c->paint.end(), c->brush_blob.begin(), c->brush_blob.end());

    // Tile bins: painter-ordered candidate lists per tile (tile_bins/transient.rs:7-146),
    // minus the tiles early occlusion hides (draw_covers).
    std::vector<TileBox> boxes(c->draws.size());
    for (uint32_t d : draw_order) boxes[d] = tile_bbox(c->draws[d].pixel_bounds, tw, th);
    tk_scene view{};
    view.tiles_width = tw;
    view.tiles_height = th;
    view.draws = c->draws.data();
    view.draw_count = uint32_t(c->draws.size());
    view.paint = c->paint.data();
    view.paint_words = uint32_t(c->paint.size());
    view.paint_sdf_shadow_base = uint32_t(c->sdf_blob.size());
    view.paint_brush_base = uint32_t(c->sdf_blob.size() + c->shadow_blob.size());
    view.draw_batch_ids = c->draw_batch_ids.data();
    view.layers = c->layer_records.data();
    view.layer_count = uint32_t(c->layer_records.size());
    view.batches = c->batch_records.data();
    view.batch_count = uint32_t(c->batch_records.size());
    std::vector<uint32_t> cover;
    draw_covers(view, cover);
    // A draw's list entries: the tiles of its box outside its cover.
    auto visible = [&](uint32_t d, auto visit) {
        const TileBox &bb = boxes[d];
        const TileBox hidden =
            box_intersection(bb, TileBox{cover[4 * d], cover[4 * d + 1], cover[4 * d + 2], cover[4 * d + 3]});
        auto region = [&](const TileBox &r) {
            for (uint32_t y = r.y0; y < r.y1; ++y)
                for (uint32_t x = r.x0; x < r.x1; ++x) visit(size_t(y) * tw + x);
        };
        if (hidden.area() == 0) {
            region(bb);
        } else {
            region(TileBox{bb.x0, bb.y0, bb.x1, hidden.y0});
            region(TileBox{bb.x0, hidden.y1, bb.x1, bb.y1});
            region(TileBox{bb.x0, hidden.y0, hidden.x0, hidden.y1});
            region(TileBox{hidden.x1, hidden.y0, bb.x1, hidden.y1});
        }
    };
    std::vector<uint32_t> counts(tiles, 0u);
    for (uint32_t d : draw_order) visible(d, [&](size_t t) { counts[t]++; });
    c->tile_draws.assign(tiles, tk_tile_draw_record{kNone, 0u});
    uint64_t total = 0;
    for (size_t t = 0; t < tiles; ++t) {
        if (counts[t] != 0) c->tile_draws[t].start = uint32_t(total);
        c->tile_draws[t].len = counts[t];
        total += counts[t];
    }
    if (total > 0xFFFFFFFFull) return fail(TK_ERR_OVERFLOW, "tile draw references exceed u32");
    c->tile_draw_indices.assign(total, 0u);
    std::vector<uint32_t> cursor(tiles, 0u);
    for (uint32_t d : draw_order)
        visible(d, [&](size_t t) { c->tile_draw_indices[c->tile_draws[t].start + cursor[t]++] = d; });

    // PTCL bound over all batches: per batch, particles plus (2 depth + 1) per
    // tile that can hold a particle; tiles outside any clip's bbox hold none.
    uint64_t ptcl_capacity = 0;
    for (size_t b = 0; b < batches.size(); ++b) {
        TileBox clip{0, 0, tw, th};
        for (uint32_t cd : batches[b].stack) {
            TileBox cb = tile_bbox(c->draws[cd].pixel_bounds, tw, th);
            clip.x0 = std::max(clip.x0, cb.x0);
            clip.y0 = std::max(clip.y0, cb.y0);
            clip.x1 = std::min(clip.x1, cb.x1);
            clip.y1 = std::min(clip.y1, cb.y1);
        }
        uint64_t area = 0;
        for (uint32_t d : batches[b].draws) {
            TileBox bb = boxes[d];
            bb.x0 = std::max(bb.x0, clip.x0);
            bb.y0 = std::max(bb.y0, clip.y0);
            bb.x1 = std::min(bb.x1, clip.x1);
            bb.y1 = std::min(bb.y1, clip.y1);
            area += bb.area();
        }
        uint64_t depth = batches[b].stack.size();
        ptcl_capacity += area + (2 * depth + 1) * std::min<uint64_t>(clip.area(), area);
    }

    // Scan and cumsum chunk plans.
    c->scan_chunks.clear();
    c->scan_chunk_ranges.assign(c->paths.size(), tk_scan_chunk_range{0, 0});
    c->cumsum_offsets.clear();
    c->cumsum_lens.clear();
    c->cumsum_row_starts.clear();
    c->cumsum_row_ends.clear();
    for (size_t p = 0; p < c->paths.size(); ++p) {
        const tk_path_record &rec = c->paths[p];
        c->scan_chunk_ranges[p].start = uint32_t(c->scan_chunks.size());
        for (uint32_t local = 0; local < rec.data_len;) {
            uint32_t len = std::min(rec.data_len - local, TK_SCAN_CHUNK_SIZE);
            c->scan_chunks.push_back(tk_scan_chunk{rec.path_id, rec.data_offset + local, rec.segment_start, len});
            local += len;
        }
        c->scan_chunk_ranges[p].end = uint32_t(c->scan_chunks.size());
        uint32_t stride = rec.tile_x1 > rec.tile_x0 ? rec.tile_x1 - rec.tile_x0 : 0;
        uint32_t height = rec.tile_y1 > rec.tile_y0 ? rec.tile_y1 - rec.tile_y0 : 0;
        if (stride == 0 || height == 0) continue;
        for (uint32_t row = 0; row < height; ++row) {
            c->cumsum_row_starts.push_back(uint32_t(c->cumsum_offsets.size()));
            uint32_t row_offset = rec.data_offset + row * stride;
            for (uint32_t lx = 0; lx < stride;) {
                uint32_t len = std::min(stride - lx, TK_CUMSUM_CHUNK_SIZE);
                c->cumsum_offsets.push_back(row_offset + lx);
                c->cumsum_lens.push_back(len);
                lx += len;
            }
            c->cumsum_row_ends.push_back(uint32_t(c->cumsum_offsets.size()));
        }
    }
    if (c->segment_total > 0xFFFFFFFFull) return fail(TK_ERR_OVERFLOW, "segment capacity exceeds u32");

    std::memset(out, 0, sizeof(*out));
    out->width = c->width;
    out->height = c->height;
    out->tiles_width = tw;
    out->tiles_height = th;
    out->clear_color = clear_color;
    out->lines = c->lines.data();
    out->line_count = uint32_t(c->lines.size());
    out->paths = c->paths.data();
    out->path_count = uint32_t(c->paths.size());
    out->draws = c->draws.data();
    out->draw_count = uint32_t(c->draws.size());
    out->paint = c->paint.data();
    out->paint_words = uint32_t(c->paint.size());
    out->paint_sdf_shadow_base = uint32_t(c->sdf_blob.size());
    out->paint_brush_base = uint32_t(c->sdf_blob.size() + c->shadow_blob.size());
    out->draw_batch_ids = c->draw_batch_ids.data();
    out->layers = c->layer_records.data();
    out->layer_count = uint32_t(c->layer_records.size());
    out->batches = c->batch_records.data();
    out->batch_count = uint32_t(c->batch_records.size());
    out->tile_draws = c->tile_draws.data();
    out->tile_draw_indices = c->tile_draw_indices.data();
    out->tile_draw_index_count = uint32_t(c->tile_draw_indices.size());
    out->scan_chunks = c->scan_chunks.data();
    out->scan_chunk_count = uint32_t(c->scan_chunks.size());
    out->scan_chunk_ranges = c->scan_chunk_ranges.data();
    out->cumsum_chunk_offsets = c->cumsum_offsets.data();
    out->cumsum_chunk_lens = c->cumsum_lens.data();
    out->cumsum_chunk_count = uint32_t(c->cumsum_offsets.size());
    out->cumsum_row_starts = c->cumsum_row_starts.data();
    out->cumsum_row_ends = c->cumsum_row_ends.data();
    out->cumsum_row_count = uint32_t(c->cumsum_row_starts.size());
    out->backdrop_len = c->backdrop_pool;
    out->segment_capacity = uint32_t(c->segment_total);
    out->max_clip_depth = max_depth;
    out->ptcl_capacity = ptcl_capacity;
    pack_images(c);
    out->atlas = c->atlas_pages ? c->atlas.data() : nullptr;
    out->atlas_width = c->atlas_width;
    out->atlas_height = c->atlas_height;
    out->atlas_pages = c->atlas_pages;
    out->atlas_reserved = 0;
    return TK_OK;
}

}  // namespace
}  // namespace tk

using namespace tk;

extern "C" {

tk_status tk_canvas_create(uint32_t width, uint32_t height, double scale, tk_canvas **out) {
    if (!out) return fail(TK_ERR_INVALID, "null output");
    if (width == 0 || height == 0 || !(scale > 0.0) || !std::isfinite(scale))
        return fail(TK_ERR_INVALID, "canvas size and scale must be positive");
    double pw = std::ceil(double(width) * scale), ph = std::ceil(double(height) * scale);
    if (pw > 65536.0 || ph > 65536.0) return fail(TK_ERR_INVALID, "canvas exceeds 65536 physical pixels");
    tk_canvas *c = new (std::nothrow) tk_canvas();
    if (!c) return fail(TK_ERR_OOM, "canvas allocation failed");
    c->width = uint32_t(pw);
    c->height = uint32_t(ph);
    c->scale = scale;
    c->tiles_w = (c->width + TK_TILE_SIZE - 1) / TK_TILE_SIZE;
    c->tiles_h = (c->height + TK_TILE_SIZE - 1) / TK_TILE_SIZE;
    *out = c;
    return TK_OK;
}

void tk_canvas_destroy(tk_canvas *canvas) { delete canvas; }

uint32_t tk_canvas_line_count(const tk_canvas *canvas) { return canvas ? uint32_t(canvas->lines.size()) : 0u; }

static tk_status push_path_common(tk_canvas *c, const uint8_t *verbs, size_t verb_count, const double *points,
                                  size_t point_count, double stroke_width, bool stroke, const tk_brush *brush,
                                  const double *transform, uint32_t fill_rule, double tolerance, uint32_t *draw_id,
                                  const StrokeStyle &style = StrokeStyle{}, uint32_t tag = TK_DRAW_BRUSH) {
    if (!c) return fail(TK_ERR_INVALID, "null canvas");
    if (fill_rule > 1u) return fail(TK_ERR_INVALID, "fill rule must be 0 (nonzero) or 1 (even-odd)");
    if (!(tolerance > 0.0)) return fail(TK_ERR_INVALID, "tolerance must be positive");
    BezPath path;
    if (verb_count > 0 && !parse_path(verbs, verb_count, points, point_count, &path))
        return fail(TK_ERR_INVALID, "malformed path: verbs and points disagree or non-finite");
    if (stroke) {
        if (!(stroke_width >= 0.0) || !std::isfinite(stroke_width)) return fail(TK_ERR_INVALID, "invalid stroke width");
        path = stroke_outline(path, stroke_width, tolerance, style);
    }
    double m[6];
    device_matrix(c, transform, m);
    BezPath physical = transform_path(path, m);
    tk_status status;
    uint32_t d = push_physical_path(c, physical, tolerance * c->scale, fill_rule, tag, brush, true, &status);
    if (status != TK_OK) return fail(status, "unsupported brush");
    if (draw_id) *draw_id = d;
    return TK_OK;
}

tk_status tk_canvas_push_polygons(tk_canvas *canvas, const double *points, const uint32_t *counts, uint32_t polygon_count,
                                  const float *rgba, const uint32_t *fill_rules, uint32_t *first_draw_id) {
    if (!canvas) return fail(TK_ERR_INVALID, "null canvas");
    if (polygon_count && (!points || !counts || !rgba || !fill_rules)) return fail(TK_ERR_INVALID, "null argument");
    size_t base = 0;
    for (uint32_t i = 0; i < polygon_count; ++i) {
        if (fill_rules[i] > 1u) return fail(TK_ERR_INVALID, "fill rule must be 0 (nonzero) or 1 (even-odd)");
        base += counts[i];
    }
    std::vector<uint8_t> verbs;
    base = 0;
    for (uint32_t i = 0; i < polygon_count; ++i) {
        const uint32_t n = counts[i];
        verbs.assign(n + 1u, uint8_t(TK_PATH_LINE));
        if (n) verbs[0] = uint8_t(TK_PATH_MOVE);
        verbs[n] = uint8_t(TK_PATH_CLOSE);
        tk_brush brush{};
        brush.kind = TK_BRUSH_SOLID;
        for (int k = 0; k < 4; ++k) brush.rgba[k] = rgba[4 * size_t(i) + k];
        uint32_t id = 0;
        tk_status st = n ? push_path_common(canvas, verbs.data(), verbs.size(), points + 2 * base, n, 0.0, false, &brush,
                                            nullptr, fill_rules[i], 0.25, &id)
                         : fail(TK_ERR_INVALID, "empty polygon");
        if (st != TK_OK) return st;
        if (i == 0 && first_draw_id) *first_draw_id = id;
        base += n;
    }
    return TK_OK;
}

tk_status tk_canvas_push_path(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count, const double *points,
                              size_t point_count, const tk_brush *brush, const double *transform, uint32_t fill_rule,
                              double tolerance, uint32_t *draw_id) {
    return push_path_common(canvas, verbs, verb_count, points, point_count, 0.0, false, brush, transform, fill_rule,
                            tolerance, draw_id);
}

tk_status tk_canvas_push_text_outline(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count, const double *points,
                                      size_t point_count, const tk_brush *brush, const double *transform,
                                      double tolerance, uint32_t *draw_id) {
    if (draw_id) *draw_id = kNone;
    if (verb_count == 0) return TK_OK;  // an empty outline records nothing
    return push_path_common(canvas, verbs, verb_count, points, point_count, 0.0, false, brush, transform,
                            TK_FILL_NONZERO, tolerance, draw_id, StrokeStyle{}, TK_DRAW_PATH_GLYPH);
}

tk_status tk_canvas_push_stroke(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count, const double *points,
                                size_t point_count, double width, const tk_brush *brush, const double *transform,
                                uint32_t fill_rule, double tolerance, uint32_t *draw_id) {
    return push_path_common(canvas, verbs, verb_count, points, point_count, width, true, brush, transform, fill_rule,
                            tolerance, draw_id);
}

tk_status tk_canvas_push_stroke_styled(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count, const double *points,
                                       size_t point_count, const tk_stroke_style *style, const tk_brush *brush,
                                       const double *transform, uint32_t fill_rule, double tolerance,
                                       uint32_t *draw_id) {
    if (!style) return fail(TK_ERR_INVALID, "null stroke style");
    if (style->join > TK_JOIN_ROUND || style->start_cap > TK_CAP_ROUND || style->end_cap > TK_CAP_ROUND)
        return fail(TK_ERR_INVALID, "unknown join or cap");
    if (!(style->miter_limit >= 1.0) && style->join == TK_JOIN_MITER)
        return fail(TK_ERR_INVALID, "miter limit must be at least 1");
    StrokeStyle st;
    st.join = style->join;
    st.start_cap = style->start_cap;
    st.end_cap = style->end_cap;
    st.miter_limit = style->miter_limit;
    return push_path_common(canvas, verbs, verb_count, points, point_count, style->width, true, brush, transform,
                            fill_rule, tolerance, draw_id, st);
}

tk_status tk_canvas_push_rect(tk_canvas *canvas, double x0, double y0, double x1, double y1, const float radii[4],
                              const tk_brush *brush, uint32_t *draw_id) {
    if (!canvas) return fail(TK_ERR_INVALID, "null canvas");
    if (!std::isfinite(x0) || !std::isfinite(y0) || !std::isfinite(x1) || !std::isfinite(y1))
        return fail(TK_ERR_INVALID, "non-finite rectangle");
    tk_status status;
    uint32_t d = push_sdf_rect(canvas, x0, y0, x1, y1, radii, brush, TK_DRAW_BRUSH, true, &status);
    if (status != TK_OK) return fail(status, "unsupported brush");
    if (draw_id) *draw_id = d;
    return TK_OK;
}

tk_status tk_canvas_add_image(tk_canvas *canvas, uint32_t width, uint32_t height, const uint32_t *premul_rgba8,
                              uint32_t *image_id) {
    if (!canvas || !premul_rgba8 || !image_id) return fail(TK_ERR_INVALID, "null argument");
    if (width == 0 || height == 0) return fail(TK_ERR_INVALID, "an image has at least one pixel");
    if (uint64_t(width) * height > (1ull << 28)) return fail(TK_ERR_OVERFLOW, "image larger than 2^28 pixels");
    tk_canvas::Image im{width, height, std::vector<uint32_t>(premul_rgba8, premul_rgba8 + size_t(width) * height)};
    *image_id = uint32_t(canvas->images.size());
    canvas->images.push_back(std::move(im));
    return TK_OK;
}

tk_status tk_canvas_push_image(tk_canvas *canvas, double x0, double y0, double x1, double y1, uint32_t image,
                               uint32_t extend, uint32_t sampling, uint32_t *draw_id) {
    if (!canvas) return fail(TK_ERR_INVALID, "null canvas");
    if (draw_id) *draw_id = kNone;
    if (image >= canvas->images.size()) return fail(TK_ERR_INVALID, "unknown image");
    if (extend > 2u || sampling > 1u) return fail(TK_ERR_INVALID, "extend is 0-2 and sampling 0-1");
    // image_rect_is_valid (src/canvas.rs:221-228): nothing is recorded otherwise.
    if (!std::isfinite(x0) || !std::isfinite(y0) || !std::isfinite(x1) || !std::isfinite(y1) || !(x1 - x0 > 0.0) ||
        !(y1 - y0 > 0.0))
        return TK_OK;
    // PatternBrush::for_rect_resource (src/shared/brush.rs:557-583), in f32.
    tk_brush b{};
    b.kind = TK_BRUSH_IMAGE;
    b.image = image;
    b.extend = extend;
    b.sampling = sampling;
    b.opacity = 255u;
    const float sx = 1.0f / float(x1 - x0), sy = 1.0f / float(y1 - y0);
    const float t[6] = {sx, 0.0f, 0.0f, sy, -float(x0) * sx, -float(y0) * sy};
    std::memcpy(b.pattern_transform, t, sizeof(t));
    const float zero[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    tk_status status;
    uint32_t d = push_sdf_rect(canvas, x0, y0, x1, y1, zero, &b, TK_DRAW_BRUSH, true, &status);
    if (status != TK_OK) return fail(status, "image brush encoding failed");
    if (draw_id) *draw_id = d;
    return TK_OK;
}

tk_status tk_canvas_push_clip_sdf_rect_layer(tk_canvas *canvas, double x0, double y0, double x1, double y1,
                                             const float radii[4]) {
    if (!canvas) return fail(TK_ERR_INVALID, "null canvas");
    if (!std::isfinite(x0) || !std::isfinite(y0) || !std::isfinite(x1) || !std::isfinite(y1))
        return fail(TK_ERR_INVALID, "non-finite rectangle");
    tk_status status;
    uint32_t d = push_sdf_rect(canvas, x0, y0, x1, y1, radii, nullptr, TK_DRAW_CLIP, false, &status);
    if (status != TK_OK) return fail(status, "clip encoding failed");
    open_clip_layer(canvas, d);
    return TK_OK;
}

tk_status tk_canvas_push_clip_layer(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count, const double *points,
                                    size_t point_count, const double *transform, uint32_t fill_rule, double tolerance) {
    if (!canvas) return fail(TK_ERR_INVALID, "null canvas");
    if (fill_rule > 1u) return fail(TK_ERR_INVALID, "fill rule must be 0 or 1");
    if (!(tolerance > 0.0)) return fail(TK_ERR_INVALID, "tolerance must be positive");
    BezPath path;
    if (verb_count > 0 && !parse_path(verbs, verb_count, points, point_count, &path))
        return fail(TK_ERR_INVALID, "malformed path");
    double m[6];
    device_matrix(canvas, transform, m);
    BezPath physical = transform_path(path, m);
    double r[4];
    if (axis_aligned_rect(physical, r)) {
        bool aligned = true;
        for (double v : r) aligned = aligned && std::fabs(v - std::round(v)) <= 1.0e-9;
        if (aligned) {
            tk_status status;
            // physical coordinates: undo the scale push_sdf_rect applies
            uint32_t d = push_sdf_rect(canvas, r[0] / canvas->scale, r[1] / canvas->scale, r[2] / canvas->scale,
                                       r[3] / canvas->scale, nullptr, nullptr, TK_DRAW_CLIP, false, &status);
            if (status != TK_OK) return fail(status, "clip encoding failed");
            open_clip_layer(canvas, d);
            return TK_OK;
        }
    }
    tk_status status;
    tk_brush transparent{};
    transparent.kind = TK_BRUSH_SOLID;
    uint32_t d = push_physical_path(canvas, physical, tolerance * canvas->scale, fill_rule, TK_DRAW_CLIP, &transparent,
                                    false, &status);
    if (status != TK_OK) return fail(status, "clip encoding failed");
    open_clip_layer(canvas, d);
    return TK_OK;
}

tk_status tk_canvas_pop_layer(tk_canvas *canvas) {
    if (!canvas) return fail(TK_ERR_INVALID, "null canvas");
    if (canvas->open.size() <= 1) return fail(TK_ERR_INVALID, "no open layer");
    canvas->open.pop_back();
    return TK_OK;
}

tk_status tk_canvas_encode(tk_canvas *canvas, uint32_t clear_color, tk_scene *out) {
    if (!canvas || !out) return fail(TK_ERR_INVALID, "null argument");
    return encode(canvas, clear_color, out);
}

uint32_t tk_line_scanned_tile_count(const tk_line *line, uint32_t tile_x0, uint32_t tile_y0, uint32_t tile_x1,
                                    uint32_t tile_y1) {
    if (!line) return 0;
    return line_scanned_tile_count(*line, tile_x0, tile_y0, tile_x1, tile_y1);
}

}  // extern "C"
