// Curve flattening shared by the host recorder (canvas.cpp) and the device
// front end (frontend.hip), so that both produce the same vertices for the
// same curves: quadratics by the parabola-integral method (R. Levien,
// "Flattening quadratic Beziers", 2019), cubics split first into quadratics
// whose approximation error is a tenth of the tolerance.  The arithmetic is
// plain FP64 (+, -, *, /, sqrt, ceil) with no contraction, which both targets
// round identically; the square root of a sum of squares stands in for
// hypot(), and the count of quadratics per cubic is found in integers rather
// than through cbrt(), so no libm routine whose last bit may differ between the
// host and the device decides a vertex.
#pragma once

#include <cmath>
#include <cstdint>

#include "scan_line.h"

namespace tk {

struct FPt {
    double x, y;
};

TK_HD inline double parabola_integral(double x) {
    const double d = 0.67;
    return x / (1.0 - d + std::sqrt(std::sqrt(d * d * d * d + 0.25 * x * x)));
}
TK_HD inline double parabola_inv_integral(double x) {
    const double b = 0.39;
    return x * (1.0 - b + std::sqrt(b * b + 0.25 * x * x));
}
TK_HD inline double norm2(double x, double y) { return std::sqrt(x * x + y * y); }

struct QuadParams {
    FPt p0, p1, p2;
    double a0, a1, u0, uscale, val;
};

// The quadratic's chord budget `val` (negative: a straight chord) and the
// parabola mapping its chord points are placed with.
TK_HD inline QuadParams quad_params(FPt p0, FPt p1, FPt p2, double sqrt_tol) {
    QuadParams q{p0, p1, p2, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double ddx = 2.0 * p1.x - p0.x - p2.x, ddy = 2.0 * p1.y - p0.y - p2.y;
    const double cross = (p2.x - p0.x) * ddy - (p2.y - p0.y) * ddx;
    const double x0 = ((p1.x - p0.x) * ddx + (p1.y - p0.y) * ddy) / cross;
    const double x1 = ((p2.x - p1.x) * ddx + (p2.y - p1.y) * ddy) / cross;
    const double scale = std::fabs(cross / (norm2(ddx, ddy) * (x1 - x0)));
    q.a0 = parabola_integral(x0);
    q.a1 = parabola_integral(x1);
    if (std::isfinite(scale) && std::isfinite(q.a0) && std::isfinite(q.a1)) {
        const double da = std::fabs(q.a1 - q.a0);
        const double sqrt_scale = std::sqrt(scale);
        if ((x0 < 0.0) == (x1 < 0.0)) {
            q.val = da * sqrt_scale;
        } else {
            // The vertex of the parabola lies inside the segment: bound the
            // integrand near the cusp by the tolerance.
            const double xmin = sqrt_tol / sqrt_scale;
            q.val = sqrt_tol * da / parabola_integral(xmin);
        }
        const double u0 = parabola_inv_integral(q.a0), u2 = parabola_inv_integral(q.a1);
        q.u0 = u0;
        q.uscale = 1.0 / (u2 - u0);
        if (!std::isfinite(q.val) || !std::isfinite(q.uscale)) q.val = -1.0;
    } else {
        q.val = -1.0;  // (nearly) a straight line: one chord
    }
    return q;
}

TK_HD inline FPt quad_eval(const QuadParams &q, double t) {
    const double mt = 1.0 - t;
    return FPt{mt * mt * q.p0.x + 2.0 * mt * t * q.p1.x + t * t * q.p2.x,
               mt * mt * q.p0.y + 2.0 * mt * t * q.p1.y + t * t * q.p2.y};
}

// Parameter on the quadratic at a fraction u in [0, 1] of its integral.
TK_HD inline double quad_param_at(const QuadParams &q, double u) {
    const double a = q.a0 + (q.a1 - q.a0) * u;
    const double t = (parabola_inv_integral(a) - q.u0) * q.uscale;
    return std::fmin(std::fmax(t, 0.0), 1.0);
}

TK_HD inline double cubic_coord(double a, double b, double c, double d, double t) {
    const double mt = 1.0 - t;
    return mt * mt * mt * a + 3.0 * mt * mt * t * b + 3.0 * mt * t * t * c + t * t * t * d;
}

// Quadratics per cubic: the midpoint-quadratic error of a cubic on a span h is
// sqrt(3)/36 |p3 - 3p2 + 3p1 - p0| h^3, so the smallest q >= 1 with
// q^3 >= err / quad_tol keeps it within quad_tol (at most 4096).
TK_HD inline int cubic_quad_count(FPt p0, FPt p1, FPt p2, FPt p3, double quad_tol) {
    const double d3x = p3.x - 3.0 * p2.x + 3.0 * p1.x - p0.x, d3y = p3.y - 3.0 * p2.y + 3.0 * p1.y - p0.y;
    const double e = std::sqrt(3.0) / 36.0 * norm2(d3x, d3y) / quad_tol;
    if (!(e > 1.0)) return 1;
    if (!(e <= 4096.0 * 4096.0 * 4096.0)) return 4096;
    int q = 1;
    for (int bit = 2048; bit >= 1; bit >>= 1) {  // largest q with q^3 < e, then one more
        const double c = double(q + bit);
        if (c * c * c < e) q += bit;
    }
    return q + 1;
}

// Quadratic i of the nq that approximate the cubic (each with the cubic's end
// tangents of its span).
TK_HD inline QuadParams cubic_quad(FPt p0, FPt p1, FPt p2, FPt p3, int i, int nq, double sqrt_tol) {
    const double t0 = double(i) / nq, t1 = double(i + 1) / nq, h = t1 - t0;
    auto deriv = [](double t, double a, double b, double c, double d) {
        const double mt = 1.0 - t;
        return 3.0 * (mt * mt * (b - a) + 2.0 * mt * t * (c - b) + t * t * (d - c));
    };
    const FPt q0 = i == 0 ? p0 : FPt{cubic_coord(p0.x, p1.x, p2.x, p3.x, t0), cubic_coord(p0.y, p1.y, p2.y, p3.y, t0)};
    const FPt q2 = i + 1 == nq ? p3
                               : FPt{cubic_coord(p0.x, p1.x, p2.x, p3.x, t1), cubic_coord(p0.y, p1.y, p2.y, p3.y, t1)};
    const FPt c1{q0.x + deriv(t0, p0.x, p1.x, p2.x, p3.x) * h / 3.0, q0.y + deriv(t0, p0.y, p1.y, p2.y, p3.y) * h / 3.0};
    const FPt c2{q2.x - deriv(t1, p0.x, p1.x, p2.x, p3.x) * h / 3.0, q2.y - deriv(t1, p0.y, p1.y, p2.y, p3.y) * h / 3.0};
    const FPt qc{(3.0 * (c1.x + c2.x) - q0.x - q2.x) * 0.25, (3.0 * (c1.y + c2.y) - q0.y - q2.y) * 0.25};
    return quad_params(q0, qc, q2, sqrt_tol);
}

// Emits the chord end points (not the start) of a run of quadratics sharing one
// chord budget, then the run's end; quad(k) returns quadratic k of `count`.
template <class Quad, class Emit>
TK_HD inline void flatten_quads(int count, Quad quad, double sqrt_tol, FPt end, Emit emit) {
    double sum = 0.0;
    for (int k = 0; k < count; ++k) {
        const QuadParams q = quad(k);
        sum += std::fmax(q.val, 0.0);
    }
    int n = int(std::ceil(0.5 * sum / sqrt_tol));
    n = n > 1 ? n : 1;
    n = n < (1 << 16) ? n : (1 << 16);
    const double step = sum / n;
    double acc = 0.0;  // integral consumed before the current quadratic
    int emitted = 0;
    for (int k = 0; k < count; ++k) {
        const QuadParams q = quad(k);
        const double val = std::fmax(q.val, 0.0);
        if (q.val < 0.0) {
            // Degenerate parabola mapping: the quadratic is a straight chord.
            if (k + 1 < count) emit(q.p2);
            continue;
        }
        while (emitted + 1 < n) {
            const double target = step * (emitted + 1);
            if (target > acc + val) break;
            const double u = val > 0.0 ? (target - acc) / val : 1.0;
            emit(quad_eval(q, quad_param_at(q, u)));
            ++emitted;
        }
        acc += val;
        if (k + 1 < count) emit(q.p2);  // keep sub-curve joints exact
    }
    emit(end);
}

template <class Emit>
TK_HD inline void flatten_quad(FPt p0, FPt p1, FPt p2, double tol, Emit emit) {
    const double sqrt_tol = std::sqrt(tol);
    flatten_quads(1, [&](int) { return quad_params(p0, p1, p2, sqrt_tol); }, sqrt_tol, p2, emit);
}

template <class Emit>
TK_HD inline void flatten_cubic(FPt p0, FPt p1, FPt p2, FPt p3, double tol, Emit emit) {
    const double sqrt_tol = std::sqrt(0.9 * tol);
    const int nq = cubic_quad_count(p0, p1, p2, p3, 0.1 * tol);
    flatten_quads(nq, [&](int i) { return cubic_quad(p0, p1, p2, p3, i, nq, sqrt_tol); }, sqrt_tol, p3, emit);
}

// Exact bounds of a cubic (kurbo's segment extrema): its ends and the points
// where a coordinate's derivative vanishes inside (0, 1).
TK_HD inline void cubic_axis_extrema(double a, double b, double c, double d, double *t_out, int *n_out) {
    // derivative / 3 = qa t^2 + qb t + qc
    const double qa = d - 3.0 * c + 3.0 * b - a;
    const double qb = 2.0 * (c - 2.0 * b + a);
    const double qc = b - a;
    int n = 0;
    if (std::fabs(qa) < 1e-12) {
        if (std::fabs(qb) > 1e-12) {
            const double t = -qc / qb;
            if (t > 0.0 && t < 1.0) t_out[n++] = t;
        }
    } else {
        const double disc = qb * qb - 4.0 * qa * qc;
        if (disc >= 0.0) {
            const double sq = std::sqrt(disc);
            const double t1 = (-qb + sq) / (2.0 * qa), t2 = (-qb - sq) / (2.0 * qa);
            if (t1 > 0.0 && t1 < 1.0) t_out[n++] = t1;
            if (t2 > 0.0 && t2 < 1.0) t_out[n++] = t2;
        }
    }
    *n_out = n;
}

}  // namespace tk
