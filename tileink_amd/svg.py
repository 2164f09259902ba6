"""SVG import: a subset of SVG 1.1 recorded on a Canvas.

Mirrors what the reference's SVG front end does for plain vector content
(src/svg/geometry.rs, src/svg/paint.rs via usvg, loaded at a target width as
examples/common/svg.rs:9-52 does): shapes and paths become BezPaths in user
units, filled (fill, fill-opacity, fill-rule) and then stroked (stroke,
stroke-opacity, stroke-width, stroke-linejoin, stroke-linecap,
stroke-miterlimit), each under its accumulated transform times the viewBox
mapping.  A clip-path naming a clipPath of one shape becomes a clip layer of that
shape (src/svg/clip.rs:60-90, 172-186), in user space or the element's bounding
box.  Raster images (PNG, JPEG, GIF, WebP: embedded as base64 data URIs
or files beside the document) become image brushes on a rect path, as
src/svg/scene_image.rs:25-83 lowers usvg's image node, with the node's
placement computed as usvg computes it, in f32 (x, y, width, height, auto
sizes, preserveAspectRatio with a clip layer for slice, image-rendering).
Elements this subset does not take (text, SVG images, gradients, patterns,
clipPaths of several shapes or nested clips, masks, filters, markers, dashes,
group opacity, use) raise SvgUnsupported instead of being dropped silently.
"""
import base64
import io
import math
import os
import re
import xml.etree.ElementTree as ET

import numpy as np

import tileink_amd as tk

SVG_NS = "{http://www.w3.org/2000/svg}"
TOLERANCE = 0.1  # SvgOptions::default().tolerance (src/svg.rs:32-40)
FONT_SIZE = 12.0  # usvg's default font-size, the base of em lengths

COLORS = {
    "black": (0, 0, 0), "white": (255, 255, 255), "red": (255, 0, 0), "green": (0, 128, 0), "blue": (0, 0, 255),
    "yellow": (255, 255, 0), "cyan": (0, 255, 255), "aqua": (0, 255, 255), "magenta": (255, 0, 255),
    "fuchsia": (255, 0, 255), "gray": (128, 128, 128), "grey": (128, 128, 128), "silver": (192, 192, 192),
    "maroon": (128, 0, 0), "olive": (128, 128, 0), "lime": (0, 255, 0), "teal": (0, 128, 128), "navy": (0, 0, 128),
    "purple": (128, 0, 128), "orange": (255, 165, 0), "pink": (255, 192, 203), "brown": (165, 42, 42),
    "seagreen": (46, 139, 87), "lightblue": (173, 216, 230), "darkgreen": (0, 100, 0),
}
UNSUPPORTED = {"text", "pattern", "mask", "filter",
               "marker", "symbol", "foreignObject", "a", "tspan", "textPath"}
INHERITED = ("fill", "fill-opacity", "fill-rule", "stroke", "stroke-opacity", "stroke-width", "stroke-linejoin",
             "stroke-linecap", "stroke-miterlimit", "visibility", "font-size", "image-rendering", "clip-rule", "color",
             "paint-order")
SHAPES = ("path", "rect", "circle", "ellipse", "line", "polyline", "polygon")
XLINK_HREF = "{http://www.w3.org/1999/xlink}href"


class SvgUnsupported(ValueError):
    pass


class InvalidColor(SvgUnsupported):
    """A colour the reference's parser rejects (the property keeps its default)."""


# --- numbers, lengths, colours, transforms ------------------------------------------

_NUM = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?")


def _length(text, ref=None, font_size=FONT_SIZE):
    m = re.fullmatch(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)\s*([a-z%]*)\s*", text or "")
    if not m:
        raise SvgUnsupported(f"length {text!r}")
    v, unit = float(m.group(1)), m.group(2)
    scale = {"": 1.0, "px": 1.0, "in": 96.0, "cm": 96.0 / 2.54, "mm": 96.0 / 25.4, "pt": 4.0 / 3.0, "pc": 16.0,
             "em": font_size, "ex": font_size / 2.0}
    if unit == "%":
        if ref is None:
            raise SvgUnsupported("percentage without a reference length")
        return v / 100.0 * ref
    if unit not in scale:
        raise SvgUnsupported(f"length unit {unit!r}")
    return v * scale[unit]


def _number(text, default):
    """A plain number attribute; an absent or invalid value is the default."""
    try:
        return float(text) if text is not None else default
    except ValueError:
        return default


def _round_half_away(v):
    return int(math.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


def _css_number(c, text, percent_scale):
    c = c.strip()
    try:
        return float(c[:-1]) / 100.0 * percent_scale if c.endswith("%") else float(c)
    except ValueError:
        raise SvgUnsupported(f"colour {text!r}") from None


def _color(text, current=(0, 0, 0)):
    """(r, g, b) in 0-255 for an opaque colour, (r, g, b, alpha) with alpha in [0, 1)
    for a translucent one (rgba, hsla, transparent), or None for 'none'.
    currentColor is `current`."""
    t = text.strip().lower()
    if t == "none":
        return None
    if t == "currentcolor":
        return current
    if t == "transparent":
        return (0, 0, 0, 0.0)
    if t in COLORS:
        return COLORS[t]
    if t.startswith("#"):
        h = t[1:]
        if not re.fullmatch(r"[0-9a-f]{3}|[0-9a-f]{6}", h):
            raise SvgUnsupported(f"colour {text!r}")
        if len(h) == 3:
            return tuple(int(c * 2, 16) for c in h)
        return tuple(int(h[i:i + 2], 16) for i in (0, 2, 4))
    m = re.fullmatch(r"(rgba?|hsla?)\(([^)]*)\)", t)
    if not m:
        raise SvgUnsupported(f"colour {text!r}")
    parts = [x for x in m.group(2).split(",")]
    if len(parts) not in (3, 4):
        raise SvgUnsupported(f"colour {text!r}")
    alpha = 1.0
    if len(parts) == 4:
        if parts[3].strip().endswith("%"):
            raise InvalidColor(f"colour {text!r}")  # a percentage alpha is not taken
        alpha = min(max(_css_number(parts[3], text, 1.0), 0.0), 1.0)
    if m.group(1).startswith("rgb"):
        if len({c.strip().endswith("%") for c in parts[:3]}) > 1:
            raise InvalidColor(f"colour {text!r}")  # numbers and percentages mixed
        rgb = tuple(max(0, min(255, _round_half_away(_css_number(c, text, 255.0)))) for c in parts[:3])
    else:  # CSS hsl(): hue in degrees, saturation and lightness in percent
        hue = (_css_number(parts[0], text, 1.0) % 360.0) / 360.0
        sat = min(max(_css_number(parts[1], text, 1.0), 0.0), 1.0)
        light = min(max(_css_number(parts[2], text, 1.0), 0.0), 1.0)
        t2 = light * (sat + 1.0) if light <= 0.5 else light + sat - light * sat
        t1 = light * 2.0 - t2

        def channel(h):
            h = h % 1.0
            if h * 6.0 < 1.0:
                return t1 + (t2 - t1) * h * 6.0
            if h * 2.0 < 1.0:
                return t2
            if h * 3.0 < 2.0:
                return t1 + (t2 - t1) * (2.0 / 3.0 - h) * 6.0
            return t1

        rgb = tuple(max(0, min(255, _round_half_away(channel(hue + d) * 255.0))) for d in (1 / 3, 0.0, -1 / 3))
    return rgb if alpha >= 1.0 else rgb + (alpha,)


def _mul(a, b):
    """a after b, both (a, b, c, d, e, f) with x' = a x + c y + e, y' = b x + d y + f."""
    return (a[0] * b[0] + a[2] * b[1], a[1] * b[0] + a[3] * b[1], a[0] * b[2] + a[2] * b[3],
            a[1] * b[2] + a[3] * b[3], a[0] * b[4] + a[2] * b[5] + a[4], a[1] * b[4] + a[3] * b[5] + a[5])


IDENTITY = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)


def _transform(text):
    t = IDENTITY
    for name, args in re.findall(r"([a-zA-Z]+)\s*\(([^)]*)\)", text or ""):
        v = [float(x) for x in _NUM.findall(args)]
        if name == "matrix" and len(v) == 6:
            m = tuple(v)
        elif name == "translate":
            m = (1.0, 0.0, 0.0, 1.0, v[0], v[1] if len(v) > 1 else 0.0)
        elif name == "scale":
            m = (v[0], 0.0, 0.0, v[1] if len(v) > 1 else v[0], 0.0, 0.0)
        elif name == "rotate":
            a = math.radians(v[0])
            r = (math.cos(a), math.sin(a), -math.sin(a), math.cos(a), 0.0, 0.0)
            if len(v) == 3:
                r = _mul((1.0, 0.0, 0.0, 1.0, v[1], v[2]), _mul(r, (1.0, 0.0, 0.0, 1.0, -v[1], -v[2])))
            m = r
        elif name == "skewX":
            m = (1.0, 0.0, math.tan(math.radians(v[0])), 1.0, 0.0, 0.0)
        elif name == "skewY":
            m = (1.0, math.tan(math.radians(v[0])), 0.0, 1.0, 0.0, 0.0)
        else:
            raise SvgUnsupported(f"transform {name}")
        t = _mul(t, m)
    return t


def _f32(v):
    return float(np.float32(v))


def _to_f32(t):
    """An f64 transform as the f32 transform usvg keeps (each coefficient rounded)."""
    return tuple(_f32(v) for v in t)


def _concat_f32(a, b):
    """a after b in f32, as tiny-skia's Transform::pre_concat: identity shortcuts,
    separate f32 products for scale-and-translate pairs, else each pair of
    products summed in f64 and rounded once, then the translation added in f32."""
    if a == IDENTITY:
        return b
    if b == IDENTITY:
        return a
    f = np.float32
    A, B = [f(v) for v in a], [f(v) for v in b]
    if a[1] == 0 and a[2] == 0 and b[1] == 0 and b[2] == 0:
        return (float(A[0] * B[0]), 0.0, 0.0, float(A[3] * B[3]), float(A[0] * B[4] + A[4]),
                float(A[3] * B[5] + A[5]))

    def mam(p, q, r, t):
        return f(float(p) * float(q) + float(r) * float(t))

    return (float(mam(A[0], B[0], A[2], B[1])), float(mam(A[1], B[0], A[3], B[1])),
            float(mam(A[0], B[2], A[2], B[3])), float(mam(A[1], B[2], A[3], B[3])),
            float(mam(A[0], B[4], A[2], B[5]) + A[4]), float(mam(A[1], B[4], A[3], B[5]) + A[5]))


def _inverse(t):
    """The inverse of an f64 affine in kurbo's order of operations; None when it
    is not invertible (inverse_affine, src/svg/geometry.rs:143-153)."""
    det = t[0] * t[3] - t[1] * t[2]
    if not math.isfinite(det) or abs(det) <= 2.220446049250313e-16:
        return None
    inv = 1.0 / det
    out = (inv * t[3], -inv * t[1], -inv * t[2], inv * t[0], inv * (t[2] * t[5] - t[3] * t[4]),
           inv * (t[1] * t[4] - t[0] * t[5]))
    return out if all(math.isfinite(v) for v in out) else None


# --- raster images ----------------------------------------------------------------------


def _image_bytes(href, base_dir):
    """The encoded bytes an image href names, or None where usvg skips the image
    (a remote URL, a file that is not there)."""
    href = href.strip()
    if href.startswith("data:"):
        head, sep, data = href[5:].partition(",")
        if not sep or ";base64" not in head:
            raise SvgUnsupported("image data URI without base64")
        return base64.b64decode(re.sub(r"\s+", "", data))
    if re.match(r"^[A-Za-z][A-Za-z0-9+.-]*://", href) or base_dir is None:
        return None
    path = os.path.join(base_dir, href)
    if not os.path.isfile(path):
        return None
    with open(path, "rb") as f:
        return f.read()


def decode_image(data):
    """A tk.Image of encoded raster bytes: PNG with 16-bit channels cut to their
    high byte and every colour type expanded to RGBA8 (decode_png_image,
    src/svg/image.rs:11-59), JPEG, GIF (first frame) and WebP as RGBA8
    (decode_encoded_image, :61-80), each pixel premultiplied with mul_div255."""
    from PIL import Image as PILImage
    if data[:4] != b"\x89PNG" and data[:3] != b"\xff\xd8\xff" and data[:4] != b"GIF8" and not (
            data[:4] == b"RIFF" and data[8:12] == b"WEBP"):
        raise SvgUnsupported("image format (only PNG, JPEG, GIF and WebP rasters)")
    im = PILImage.open(io.BytesIO(data))
    im.seek(0)
    if im.mode in ("I;16", "I;16B", "I;16L", "I"):
        grey = (np.asarray(im, dtype=np.uint32) >> 8).astype(np.uint8)  # 16-bit grey: the high byte
        rgba = np.stack([grey, grey, grey, np.full_like(grey, 255)], -1)
    else:
        rgba = np.asarray(im.convert("RGBA"), dtype=np.uint8)
    return tk.Image.from_rgba8(rgba)


def _aspect(text):
    """(align, slice) of a preserveAspectRatio value; align 'none' or (x, y) in {0, 1, 2}."""
    words = (text or "").split()
    if words and words[0] == "defer":
        words = words[1:]
    if not words:
        return (1, 1), False
    if words[0] == "none":
        return "none", False
    m = re.fullmatch(r"x(Min|Mid|Max)Y(Min|Mid|Max)", words[0])
    if not m:
        return (1, 1), False
    pos = {"Min": 0, "Mid": 1, "Max": 2}
    return (pos[m.group(1)], pos[m.group(2)]), len(words) > 1 and words[1] == "slice"


def _image_placement(actual, x, y, width, height, aspect):
    """usvg's image node (converter/image.rs convert_inner): the image transform
    (x', y') = (sx x + tx, sy y + ty) taking image pixels to user space, in f32,
    and the viewport rect (x, y, w, h) a slice clips to."""
    f = np.float32
    iw, ih = f(actual[0]), f(actual[1])
    x, y, w, h = f(x), f(y), f(width), f(height)
    rw, rh = f(f(x + w) - x), f(f(y + h) - y)  # NonZeroRect::from_xywh keeps right and bottom
    align, slice_ = aspect
    if align == "none":
        aw, ah = rw, rh
    else:
        # Size::scale_to / expand_to
        scaled_w = f(f(rh * iw) / ih)
        keep_height = scaled_w <= rw if slice_ else scaled_w >= rw
        if not keep_height:
            aw, ah = scaled_w, rh
        else:
            aw, ah = rw, f(f(rw * ih) / iw)
    ax, ay = x, y
    if align != "none":
        dx, dy = f(rw - aw), f(rh - ah)
        ax = x if align[0] == 0 else (f(x + f(dx / f(2))) if align[0] == 1 else f(x + dx))
        ay = y if align[1] == 0 else (f(y + f(dy / f(2))) if align[1] == 1 else f(y + dy))
    vw, vh = f(f(ax + aw) - ax), f(f(ay + ah) - ay)
    ts = (float(f(vw / iw)), 0.0, 0.0, float(f(vh / ih)), float(ax), float(ay))
    return ts, (float(x), float(y), float(rw), float(rh))


# --- path data ------------------------------------------------------------------------


def _arc_to_cubics(p0, rx, ry, phi, large, sweep, p1):
    """Endpoint arc (SVG 1.1 F.6.5) as cubic Beziers, at most 90 degrees each."""
    if rx == 0 or ry == 0:
        return [("L", p1)]
    rx, ry = abs(rx), abs(ry)
    cp, sp = math.cos(phi), math.sin(phi)
    dx, dy = (p0[0] - p1[0]) / 2.0, (p0[1] - p1[1]) / 2.0
    x1, y1 = cp * dx + sp * dy, -sp * dx + cp * dy
    lam = x1 * x1 / (rx * rx) + y1 * y1 / (ry * ry)
    if lam > 1:
        rx, ry = rx * math.sqrt(lam), ry * math.sqrt(lam)
    num = rx * rx * ry * ry - rx * rx * y1 * y1 - ry * ry * x1 * x1
    den = rx * rx * y1 * y1 + ry * ry * x1 * x1
    co = math.sqrt(max(num, 0.0) / den) if den else 0.0
    if large == sweep:
        co = -co
    cx1, cy1 = co * rx * y1 / ry, -co * ry * x1 / rx
    cx = cp * cx1 - sp * cy1 + (p0[0] + p1[0]) / 2.0
    cy = sp * cx1 + cp * cy1 + (p0[1] + p1[1]) / 2.0

    def angle(ux, uy, vx, vy):
        a = math.atan2(ux * vy - uy * vx, ux * vx + uy * vy)
        return a

    t1 = angle(1.0, 0.0, (x1 - cx1) / rx, (y1 - cy1) / ry)
    dt = angle((x1 - cx1) / rx, (y1 - cy1) / ry, (-x1 - cx1) / rx, (-y1 - cy1) / ry)
    if not sweep and dt > 0:
        dt -= 2 * math.pi
    elif sweep and dt < 0:
        dt += 2 * math.pi
    n = max(1, int(math.ceil(abs(dt) / (math.pi / 2) - 1e-9)))
    step = dt / n
    k = 4.0 / 3.0 * math.tan(step / 4.0)
    out = []

    def point(t):
        x, y = rx * math.cos(t), ry * math.sin(t)
        return (cp * x - sp * y + cx, sp * x + cp * y + cy)

    def deriv(t):
        x, y = -rx * math.sin(t), ry * math.cos(t)
        return (cp * x - sp * y, sp * x + cp * y)

    t = t1
    for i in range(n):
        a, b = t, t + step
        pa, pb = point(a), point(b)
        da, db = deriv(a), deriv(b)
        c1 = (pa[0] + k * da[0], pa[1] + k * da[1])
        c2 = (pb[0] - k * db[0], pb[1] - k * db[1])
        out.append(("C", c1, c2, p1 if i == n - 1 else pb))
        t = b
    return out


_NUMBER_AT = re.compile(r"[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?")
_ARGS = {"M": 2, "L": 2, "H": 1, "V": 1, "C": 6, "S": 4, "Q": 4, "T": 2, "A": 7, "Z": 0}


class _Scanner:
    """SVG path data tokens: command letters, numbers, and the single-character
    flags of arcs (which need no separator); stops at the first invalid character."""

    def __init__(self, d):
        self.d, self.i = d or "", 0

    def skip(self):
        while self.i < len(self.d) and self.d[self.i] in " \t\r\n,":
            self.i += 1

    def command(self):
        self.skip()
        if self.i < len(self.d) and self.d[self.i] in "MmLlHhVvCcSsQqTtAaZz":
            self.i += 1
            return self.d[self.i - 1]
        return None

    def at_number(self):
        self.skip()
        return bool(_NUMBER_AT.match(self.d, self.i))

    def number(self):
        self.skip()
        m = _NUMBER_AT.match(self.d, self.i)
        if not m:
            raise ValueError("number expected")
        self.i = m.end()
        return float(m.group(0))

    def flag(self):
        self.skip()
        if self.i < len(self.d) and self.d[self.i] in "01":
            self.i += 1
            return self.d[self.i - 1] == "1"
        raise ValueError("flag expected")


def parse_path(d):
    """SVG path data as a BezPath (absolute coordinates).  Stops at the first error,
    keeping the segments read before it, as SVG's error handling does; a command
    after ClosePath starts its subpath at the closed subpath's start."""
    sc = _Scanner(d)
    path = tk.BezPath()
    cur = start = (0.0, 0.0)
    last_c2 = last_q = None
    cmd, have_move, need_move = None, False, False
    while True:
        c = sc.command()
        if c is None:
            if cmd is None or cmd.upper() == "Z" or not sc.at_number():
                break
            c = cmd  # implicit repetition
        elif not have_move and c not in "Mm":
            break  # a path must start with MoveTo
        cmd = c
        rel, C = c.islower(), c.upper()
        try:
            if C == "A":
                rx, ry, phi = sc.number(), sc.number(), sc.number()
                large, sweep = sc.flag(), sc.flag()
                args = [rx, ry, phi, large, sweep, sc.number(), sc.number()]
            else:
                args = [sc.number() for _ in range(_ARGS[C])]
        except ValueError:
            break
        ox, oy = cur if rel else (0.0, 0.0)
        if need_move and C != "M":
            path.move_to(*start)
            need_move = False
        if C == "M":
            cur = start = (ox + args[0], oy + args[1])
            path.move_to(*cur)
            have_move, need_move = True, False
            cmd = "l" if rel else "L"  # further pairs are implicit LineTo
            last_c2 = last_q = None
        elif C == "L":
            cur = (ox + args[0], oy + args[1])
            path.line_to(*cur)
            last_c2 = last_q = None
        elif C == "H":
            cur = ((cur[0] if rel else 0.0) + args[0], cur[1])
            path.line_to(*cur)
            last_c2 = last_q = None
        elif C == "V":
            cur = (cur[0], (cur[1] if rel else 0.0) + args[0])
            path.line_to(*cur)
            last_c2 = last_q = None
        elif C in "CS":
            if C == "C":
                c1 = (ox + args[0], oy + args[1])
                x2, y2, x, y = args[2:]
            else:
                c1 = (2 * cur[0] - last_c2[0], 2 * cur[1] - last_c2[1]) if last_c2 else cur
                x2, y2, x, y = args
            c2, end = (ox + x2, oy + y2), (ox + x, oy + y)
            path.curve_to(*c1, *c2, *end)
            cur, last_c2, last_q = end, c2, None
        elif C in "QT":
            if C == "Q":
                q = (ox + args[0], oy + args[1])
                x, y = args[2:]
            else:
                q = (2 * cur[0] - last_q[0], 2 * cur[1] - last_q[1]) if last_q else cur
                x, y = args
            end = (ox + x, oy + y)
            path.quad_to(*q, *end)
            cur, last_q, last_c2 = end, q, None
        elif C == "A":
            rx, ry, phi, large, sweep, x, y = args
            end = (ox + x, oy + y)
            if end != cur:
                for seg in _arc_to_cubics(cur, rx, ry, math.radians(phi), large, sweep, end):
                    if seg[0] == "L":
                        path.line_to(*seg[1])
                    else:
                        path.curve_to(*seg[1], *seg[2], *seg[3])
            cur, last_c2, last_q = end, None, None
        elif C == "Z":
            path.close_path()
            cur = start
            last_c2 = last_q = None
            need_move = True
    return path


def _ellipse_path(cx, cy, rx, ry):
    k = 4.0 / 3.0 * (math.sqrt(2.0) - 1.0)
    p = tk.BezPath().move_to(cx + rx, cy)
    p.curve_to(cx + rx, cy + k * ry, cx + k * rx, cy + ry, cx, cy + ry)
    p.curve_to(cx - k * rx, cy + ry, cx - rx, cy + k * ry, cx - rx, cy)
    p.curve_to(cx - rx, cy - k * ry, cx - k * rx, cy - ry, cx, cy - ry)
    p.curve_to(cx + k * rx, cy - ry, cx + rx, cy - k * ry, cx + rx, cy)
    return p.close_path()


def _polylines(path, pieces=64):
    """The subpaths of a BezPath as point lists, curves cut into `pieces` chords,
    each closed subpath ending back at its start (so its closing segment dashes too)."""
    verbs, pts = path.arrays()
    pts = np.asarray(pts, np.float64).reshape(-1, 2)
    out, cur, i = [], None, 0
    u = np.linspace(0.0, 1.0, pieces + 1)[1:, None]
    for v in np.asarray(verbs):
        if v == tk._abi.PATH_MOVE:
            if cur is not None and len(cur) > 1:
                out.append(cur)
            cur = [pts[i]]
            i += 1
        elif v == tk._abi.PATH_LINE:
            cur.append(pts[i])
            i += 1
        elif v == tk._abi.PATH_QUAD:
            p0, p1, p2 = cur[-1], pts[i], pts[i + 1]
            cur.extend((1 - u) ** 2 * p0 + 2 * (1 - u) * u * p1 + u * u * p2)
            i += 2
        elif v == tk._abi.PATH_CUBIC:
            p0, p1, p2, p3 = cur[-1], pts[i], pts[i + 1], pts[i + 2]
            cur.extend((1 - u) ** 3 * p0 + 3 * (1 - u) ** 2 * u * p1 + 3 * (1 - u) * u * u * p2 + u ** 3 * p3)
            i += 3
        elif v == tk._abi.PATH_CLOSE and cur is not None:
            cur.append(cur[0])
            out.append(cur)
            cur = [cur[0]]
    if cur is not None and len(cur) > 1:
        out.append(cur)
    return out


def _dash(path, dashes, offset):
    """The dashes of a path as open subpaths (kurbo's dash iterator on the
    flattened outline): the pattern restarts at every subpath, shifted by offset."""
    total = sum(dashes)
    out = tk.BezPath()
    for poly in _polylines(path):
        pos = offset % total
        k = 0
        while pos >= dashes[k]:
            pos -= dashes[k]
            k = (k + 1) % len(dashes)
        left, on, drawing = dashes[k] - pos, k % 2 == 0, False
        if on:
            out.move_to(*poly[0])
            drawing = True
        for a, b in zip(poly[:-1], poly[1:]):
            seg = float(np.hypot(*(b - a)))
            done = 0.0
            while seg - done > left:
                done += left
                q = a + (b - a) * (done / seg)
                if on:
                    out.line_to(*q)
                else:
                    out.move_to(*q)
                on = not on
                k = (k + 1) % len(dashes)
                left = dashes[k]
            left -= seg - done
            if on:
                out.line_to(*b)
    return out


def _rect_path(x, y, w, h, rx, ry):
    if rx <= 0 or ry <= 0:
        return tk.BezPath().move_to(x, y).line_to(x + w, y).line_to(x + w, y + h).line_to(x, y + h).close_path()
    k = 4.0 / 3.0 * (math.sqrt(2.0) - 1.0)
    p = tk.BezPath().move_to(x + rx, y)
    p.line_to(x + w - rx, y).curve_to(x + w - rx + k * rx, y, x + w, y + ry - k * ry, x + w, y + ry)
    p.line_to(x + w, y + h - ry).curve_to(x + w, y + h - ry + k * ry, x + w - rx + k * rx, y + h, x + w - rx, y + h)
    p.line_to(x + rx, y + h).curve_to(x + rx - k * rx, y + h, x, y + h - ry + k * ry, x, y + h - ry)
    p.line_to(x, y + ry).curve_to(x, y + ry - k * ry, x + rx - k * rx, y, x + rx, y)
    return p.close_path()


# --- the document ---------------------------------------------------------------------


# Properties a style attribute may set (presentation attributes, and transform):
# usvg ignores geometry such as width or height there.
PRESENTATION = set(INHERITED) | {"opacity", "stroke-dasharray", "stroke-dashoffset", "clip-path", "mask", "filter",
                                 "marker-start", "marker-mid", "marker-end", "display", "stop-color", "stop-opacity",
                                 "color", "mix-blend-mode", "isolation", "transform", "shape-rendering",
                                 "transform-origin", "marker"}


def _style(el):
    """The declarations of an element's style attribute, CSS comments removed."""
    text = re.sub(r"/\*.*?\*/", "", el.get("style") or "", flags=re.S)
    out = {}
    for decl in text.split(";"):
        if ":" in decl:
            k, v = decl.split(":", 1)
            if k.strip() in PRESENTATION:
                out[k.strip()] = v.strip()
    return out


def _attrs(el, parent, css=None):
    """An element's properties: inherited ones, its attributes, then CSS rules
    (normal), its style attribute, then CSS rules marked !important."""
    a = {k: v for k, v in parent.items() if k in INHERITED}
    for k, v in el.attrib.items():
        a[k] = v
    normal, important = css if css is not None else ({}, {})
    a.update(normal)
    a.update(_style(el))
    a.update(important)
    for k in ("mask", "filter", "marker-start", "marker-mid", "marker-end", "marker"):
        v = a.get(k)
        if v is not None and v.strip() not in ("none", "1", ""):
            raise SvgUnsupported(f"{k}={v}")
    for k, ok in (("mix-blend-mode", ("normal",)), ("isolation", ("auto",))):
        v = a.get(k)
        if v is not None and v.strip() not in ok:
            raise SvgUnsupported(f"{k}={v}")
    return a


def _stroke_first(text):
    """paint-order: the keywords given, the rest in the default order (fill,
    stroke, markers); a repeated or unknown keyword leaves the default."""
    if text is None:
        return False
    words = text.split()
    if words == ["normal"] or not words or len(set(words)) != len(words) or \
            any(w not in ("fill", "stroke", "markers") for w in words):
        return False
    order = words + [w for w in ("fill", "stroke", "markers") if w not in words]
    return order.index("stroke") < order.index("fill")


def _opacity(text):
    """An opacity value: a number or a percentage, clamped to [0, 1]; 1 when absent or invalid."""
    if text is None:
        return 1.0
    t = text.strip()
    try:
        v = float(t[:-1]) / 100.0 if t.endswith("%") else float(t)
    except ValueError:
        return 1.0
    return min(max(v, 0.0), 1.0)


def _gradient_stops(el, inherited_color="black"):
    """usvg's stop conversion: offsets (numbers or percentages) clamped to [0, 1]
    and non-decreasing, runs of three equal offsets cut to their ends, leading
    zeros and equal neighbours pulled apart by an epsilon; colour and opacity
    multiplied.  Returns [(offset, (r, g, b, alpha 0-1))], r, g, b in 0-255."""
    if el is None:
        return []
    stops, prev = [], 0.0
    for st in el:
        if _local(st.tag) != "stop":
            continue
        raw = _presentation(st, "offset")
        off = prev
        if raw is not None:
            mm = re.fullmatch(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)\s*(%?)\s*", raw)
            if mm:
                off = float(mm.group(1)) / (100.0 if mm.group(2) else 1.0)
        prev = off
        off = _f32(min(max(off, 0.0), 1.0))
        cv = _presentation(st, "stop-color")
        if cv is not None and cv.strip() == "inherit":  # from the gradient element
            cv = _presentation(el, "stop-color")
        current = _color(_presentation(st, "color") or inherited_color) or (0, 0, 0)
        rgb = (0, 0, 0) if cv is None or cv.strip() == "inherit" else _color(cv, current[:3])
        color_alpha = 1.0
        if rgb is None:
            rgb = (0, 0, 0)
        elif len(rgb) == 4:
            rgb, color_alpha = rgb[:3], rgb[3]
        alpha = _opacity(_presentation(st, "stop-opacity")) * color_alpha
        stops.append([off, rgb, alpha])
    eps = float(np.finfo(np.float32).eps)

    def close(x, y):
        return abs(np.float32(x).view(np.int32).astype(np.int64) - np.float32(y).view(np.int32).astype(np.int64)) <= 4

    i = 0
    while len(stops) >= 3 and i < len(stops) - 2:
        if close(stops[i][0], stops[i + 1][0]) and close(stops[i + 1][0], stops[i + 2][0]):
            del stops[i + 1]
        else:
            i += 1
    for i in range(len(stops) - 1):
        if close(stops[i][0], 0.0) and close(stops[i + 1][0], 0.0):
            stops[i + 1][0] = _f32(min(max(stops[i][0] + eps, 0.0), 1.0))
    for i in range(1, len(stops)):
        o1, o2 = stops[i - 1][0], stops[i][0]
        if o1 > o2 or close(o1, o2):
            stops[i - 1][0] = _f32(min(max(o1 - eps, 0.0), 1.0))
            stops[i][0] = o1
    return [(o, (c[0], c[1], c[2], a)) for o, c, a in stops]


_SIMPLE = re.compile(r"(\*|[A-Za-z][\w-]*)|\.([\w-]+)|#([\w-]+)|\[\s*([\w:-]+)\s*(?:=\s*['\"]?([^'\"\]]*)['\"]?\s*)?\]"
                     r"|:(first-child)")


def _parse_compound(text):
    """A compound selector as (type, [(kind, name, value)]), or None when it uses
    something this parser does not take (the rule is then dropped, as CSS does)."""
    pos, typ, conds = 0, None, []
    while pos < len(text):
        m = _SIMPLE.match(text, pos)
        if not m:
            return None
        if m.group(1):
            if pos:
                return None
            typ = m.group(1)
        elif m.group(2):
            conds.append(("class", m.group(2), None))
        elif m.group(3):
            conds.append(("id", m.group(3), None))
        elif m.group(4):
            conds.append(("attr", m.group(4), m.group(5)))
        else:
            conds.append(("first-child", None, None))
        pos = m.end()
    return typ, conds


def _compound_matches(el, compound, parent_of):
    typ, conds = compound
    if typ not in (None, "*") and _local(el.tag) != typ:
        return False
    for kind, name, value in conds:
        if kind == "class" and name not in (el.get("class") or "").split():
            return False
        if kind == "id" and el.get("id") != name:
            return False
        if kind == "attr" and (el.get(name) is None or (value is not None and el.get(name) != value)):
            return False
        if kind == "first-child":
            parent = parent_of.get(el)
            if parent is None or next((c for c in parent if isinstance(c.tag, str)), None) is not el:
                return False
    return True


def _css_rules(root):
    """The rules of the document's CSS style elements: [(selector, specificity, {property: (value, important)})],
    selector a list of (compound, combinator to its left)."""
    rules = []
    for st in root.iter():
        if _local(st.tag) != "style" or (st.get("type") or "text/css").strip() not in ("text/css", ""):
            continue
        text = re.sub(r"/\*.*?\*/", "", "".join(st.itertext()), flags=re.S)
        if "@" in text:
            raise SvgUnsupported("CSS at-rules")
        for block in text.split("}"):
            if "{" not in block:
                continue
            sel_text, body = block.split("{", 1)
            decls = {}
            for decl in body.split(";"):
                if ":" not in decl:
                    continue
                k, v = decl.split(":", 1)
                k, v = k.strip(), v.strip()
                imp = v.endswith("!important")
                if imp:
                    v = v[:-len("!important")].strip()
                if k in PRESENTATION:
                    decls[k] = (v, imp)
            for sel in sel_text.split(","):
                tokens = re.split(r"\s*(>)\s*|\s+", sel.strip())
                parts, comb, ok = [], None, True
                for tok in tokens:
                    if tok is None or tok == "":
                        continue
                    if tok == ">":
                        comb = ">"
                        continue
                    compound = _parse_compound(tok)
                    if compound is None:
                        ok = False
                        break
                    parts.append((compound, comb if parts else None))
                    comb = " "
                if not ok or not parts:
                    continue
                # combinator stored with the compound to its right
                sel_parts = [(parts[0][0], None)] + [(c, cb) for c, cb in parts[1:]]
                ids = sum(1 for c, _ in sel_parts for k, _, _ in c[1] if k == "id")
                classes = sum(1 for c, _ in sel_parts for k, _, _ in c[1] if k != "id")
                types = sum(1 for c, _ in sel_parts if c[0] not in (None, "*"))
                rules.append((sel_parts, (ids, classes, types), decls))
    return rules


def _presentation(el, name):
    """An element's own value of a presentation attribute (style wins), or None."""
    return _style(el).get(name, el.get(name))


def _local(tag):
    return tag[len(SVG_NS):] if tag.startswith(SVG_NS) else tag


class Document:
    """A parsed SVG: size (width, height) and the draw list in painter order:
    ("shape", path, transform, fill (rgba, rule) or None, stroke (rgba, style) or
    None), ("image", tk.Image, transform, sampling), ("clip", path, transform) and
    ("pop",).  Relative image hrefs resolve against base_dir."""

    def __init__(self, text, base_dir=None):
        self.base_dir = base_dir
        root = ET.fromstring(text)
        if _local(root.tag) != "svg":
            raise SvgUnsupported("not an svg document")
        vb = [float(v) for v in _NUM.findall(root.get("viewBox") or "")]
        w_attr, h_attr = root.get("width"), root.get("height")
        if len(vb) == 4 and vb[2] > 0 and vb[3] > 0:
            vw, vh = vb[2], vb[3]
        else:
            vb = None
        self.width = _length(w_attr) if w_attr and not w_attr.strip().endswith("%") else (vb[2] if vb else 100.0)
        self.height = _length(h_attr) if h_attr and not h_attr.strip().endswith("%") else (vb[3] if vb else 100.0)
        if root.get("background-color") is not None or "background-color" in (root.get("style") or ""):
            raise SvgUnsupported("background-color on the root")
        if not vb and not (w_attr and h_attr):
            raise SvgUnsupported("a root without a viewBox or a size")
        view = IDENTITY
        if vb:
            sx, sy = self.width / vb[2], self.height / vb[3]
            align, slice_ = _aspect(root.get("preserveAspectRatio"))
            if align == "none":
                view = (sx, 0.0, 0.0, sy, -vb[0] * sx, -vb[1] * sy)
            else:
                s = max(sx, sy) if slice_ else min(sx, sy)
                free_x, free_y = self.width - vb[2] * s, self.height - vb[3] * s
                tx = free_x * align[0] / 2.0 - vb[0] * s
                ty = free_y * align[1] / 2.0 - vb[1] * s
                view = (s, 0.0, 0.0, s, tx, ty)
        self.view_w, self.view_h = (vb[2], vb[3]) if vb else (self.width, self.height)
        self.clips = {el.get("id"): el for el in root.iter() if _local(el.tag) == "clipPath" and el.get("id")}
        self.gradients = {el.get("id"): el for el in root.iter()
                          if _local(el.tag) in ("linearGradient", "radialGradient") and el.get("id")}
        self.parent_of = {child: node for node in root.iter() for child in node}
        self.css = _css_rules(root)
        self.ids = {}
        for el in root.iter():  # the first element of an id wins
            if el.get("id") and el.get("id") not in self.ids:
                self.ids[el.get("id")] = el
        self.use_stack = []
        self.draws = []
        self._walk(root, {}, view, _to_f32(view))

    def _clip(self, value, a, tag, t):
        """The clip layer of a clip-path reference, as src/svg/clip.rs:172-186 lowers
        usvg's clipPath: one visible shape is a clip layer of its outline under the
        clip's transform ("clip", path, transform, rule); none clips everything
        ("empty"); anything else needs the reference's mask layers."""
        m = re.fullmatch(r"\s*url\(\s*#([^)\s]+)\s*\)\s*", value)
        if not m or m.group(1) not in self.clips:
            raise SvgUnsupported(f"clip-path={value}")
        clip = self.clips[m.group(1)]
        if _presentation(clip, "transform-origin") is not None:
            raise SvgUnsupported("transform-origin on a clipPath")
        # clip-rule is inherited through the clipPath's own ancestors in the document
        rule, node = "nonzero", clip
        while node is not None:
            v = _presentation(node, "clip-rule")
            if v is not None:
                rule = v
                break
            node = self.parent_of.get(node)
        ca = _attrs(clip, {"clip-rule": rule}, self._css(clip))
        if ca.get("clip-path", "none").strip() != "none":
            raise SvgUnsupported("nested clip-path")
        shapes = []
        for child in clip:
            ctag = _local(child.tag)
            if ctag in ("title", "desc", "metadata"):
                continue
            if ctag not in SHAPES:
                raise SvgUnsupported(f"clipPath child {ctag}")
            cattrs = _attrs(child, {"clip-rule": ca.get("clip-rule", "nonzero")}, self._css(child))
            if cattrs.get("clip-path", "none").strip() != "none":
                raise SvgUnsupported("clip-path inside clipPath")
            if cattrs.get("display") == "none" or cattrs.get("visibility") in ("hidden", "collapse"):
                continue
            path = self._shape(ctag, cattrs)
            if path is None:
                continue
            shapes.append((path, _transform(cattrs.get("transform")), cattrs.get("clip-rule", "nonzero")))
        if len(shapes) > 1:
            raise SvgUnsupported("clipPath of several shapes (a mask layer)")
        if not shapes:
            return ("empty",)
        ct = _mul(t, _transform(ca.get("transform")))
        if ca.get("clipPathUnits") == "objectBoundingBox":
            box = self._bbox(tag, a)
            if box is None:
                raise SvgUnsupported("objectBoundingBox of this element")
            x0, y0, x1, y1 = box
            if not (x1 > x0 and y1 > y0):
                return ("empty",)
            ct = _mul(ct, (x1 - x0, 0.0, 0.0, y1 - y0, x0, y0))
        path, child_t, rule = shapes[0]
        return ("clip", path, _mul(ct, child_t), tk.FillRule.EvenOdd if rule == "evenodd" else tk.FillRule.NonZero)

    def _paint(self, value, tag, a, prop="fill"):
        """A fill or stroke value: (r, g, b) in 0-255, a gradient (a dict with the
        user-space geometry, stops, spread and paint-server transform, as usvg's
        paint-server conversion leaves them), or None (no paint)."""
        value = value.strip()
        if "icc-color" in value:
            raise SvgUnsupported("icc-color")
        current = _color(a.get("color", "black")) or (0, 0, 0)
        if len(current) == 4:
            raise SvgUnsupported("a translucent currentColor")
        if value == "inherit":
            raise SvgUnsupported("paint inherit")
        if not value.startswith("url("):
            try:
                c = _color(value, current)
            except InvalidColor:  # the property keeps its initial value
                c = (0, 0, 0) if prop == "fill" else None
            return (c[:3], c[3]) if c is not None and len(c) == 4 else c
        m = re.fullmatch(r"url\(\s*#([^)\s]+)\s*\)\s*(.*)", value)
        if not m:
            raise SvgUnsupported(f"paint {value}")
        el = self.gradients.get(m.group(1))
        if el is None:  # a missing server: the fallback colour, or no paint
            return self._fallback(m.group(2), current)
        chain, seen = [el], {id(el)}
        while True:  # xlink:href: attributes and stops from the referenced gradients
            href = chain[-1].get(XLINK_HREF) or chain[-1].get("href")
            if not href or not href.startswith("#") or href[1:] not in self.gradients:
                break
            nxt = self.gradients[href[1:]]
            if id(nxt) in seen:
                break
            chain.append(nxt)
            seen.add(id(nxt))

        kind = _local(el.tag)
        if any(_presentation(g, "transform-origin") is not None for g in chain):
            raise SvgUnsupported("transform-origin on a gradient")

        def attr(name, same_kind=False):
            # geometry comes only from gradients of the same kind; units, spread,
            # transform and stops from any in the chain
            for g in chain:
                if same_kind and _local(g.tag) != kind:
                    continue
                v = _presentation(g, name) if name in ("gradientTransform",) else g.get(name)
                if v is not None:
                    return v
            return None

        holder = next((g for g in chain if any(_local(c.tag) == "stop" for c in g)), None)
        inherited, node = "black", holder
        while node is not None:  # `color` for currentColor stops, through the document's ancestors
            v = _presentation(node, "color")
            if v is not None:
                inherited = v
                break
            node = self.parent_of.get(node)
        stops = _gradient_stops(holder, inherited)
        if len(stops) < 2:
            if stops:
                return (stops[0][1][:3], stops[0][1][3])
            return self._fallback(m.group(2), current)  # no stops: the fallback, or no paint
        units = attr("gradientUnits") or "objectBoundingBox"
        bbox_units = units != "userSpaceOnUse"
        W, H = self.view_w, self.view_h
        diag = math.hypot(W, H) / math.sqrt(2)

        def length(name, default, ref):
            v = attr(name, same_kind=True)
            text = default if v is None else v
            mm = re.fullmatch(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)\s*(%?)\s*", text)
            if not mm:
                raise SvgUnsupported(f"gradient length {text!r}")
            num = float(mm.group(1))
            if mm.group(2):
                return num / 100.0 if bbox_units else num / 100.0 * ref
            return num if bbox_units else _length(text, ref)

        spread = {"reflect": 2, "repeat": 1}.get((attr("spreadMethod") or "pad").strip(), 0)
        ts = _transform(attr("gradientTransform"))
        if bbox_units:
            box = self._bbox(tag, a)
            if box is None:
                raise SvgUnsupported("objectBoundingBox of this element")
            x0, y0, x1, y1 = box
            if not (x1 > x0 and y1 > y0):  # usvg drops a bounding-box paint on a zero-sized box
                return self._fallback(m.group(2), current)
            ts = _mul((x1 - x0, 0.0, 0.0, y1 - y0, x0, y0), ts)
        if _local(el.tag) == "linearGradient" and _local(chain[0].tag) == "linearGradient":
            geom = ("linear", length("x1", "0%", W), length("y1", "0%", H), length("x2", "100%", W),
                    length("y2", "0%", H))
        elif _local(el.tag) == "radialGradient":
            cx, cy = length("cx", "50%", W), length("cy", "50%", H)
            r = length("r", "50%", diag)
            fx = length("fx", "50%", W) if attr("fx", True) is not None else cx
            fy = length("fy", "50%", H) if attr("fy", True) is not None else cy
            fr = length("fr", "0%", diag)
            if r < 0 or fr < 0:
                raise SvgUnsupported("negative gradient radius")  # undefined in SVG 2
            if not (r > 0):
                return (stops[-1][1][:3], stops[-1][1][3])  # r = 0: the last stop's colour
            geom = ("radial", fx, fy, fr, cx, cy, r)
        else:
            raise SvgUnsupported("gradient type")
        return {"geom": geom, "stops": stops, "spread": spread, "transform": ts}

    def _css(self, el):
        """(normal, important) declarations of the rules matching el, in cascade order."""
        if not self.css:
            return None
        hits = []
        for order, (selector, spec, decls) in enumerate(self.css):
            if self._matches(el, selector):
                hits.append((spec, order, decls))
        normal, important = {}, {}
        for _, _, decls in sorted(hits, key=lambda h: (h[0], h[1])):
            for k, (v, imp) in decls.items():
                (important if imp else normal)[k] = v
        return normal, important

    def _matches(self, el, selector):
        parts = list(selector)  # [(compound, combinator to the left)], rightmost last
        node = el
        compound, comb = parts[-1]
        if not _compound_matches(node, compound, self.parent_of):
            return False
        for compound_l, comb_l in reversed(parts[:-1]):
            if comb == ">":
                node = self.parent_of.get(node)
                if node is None or not _compound_matches(node, compound_l, self.parent_of):
                    return False
            else:  # descendant
                node = self.parent_of.get(node)
                while node is not None and not _compound_matches(node, compound_l, self.parent_of):
                    node = self.parent_of.get(node)
                if node is None:
                    return False
            comb = comb_l
        return True

    def _nested_svg(self, el, a, t, t32):
        """A nested svg: a new viewport at (x, y, width, height) of the parent's, its
        viewBox fitted by preserveAspectRatio, and (overflow not visible) a clip to
        the viewport, as usvg converts it into a clipped, transformed group."""
        fs = _length(a["font-size"]) if "font-size" in a else FONT_SIZE
        W, H = self.view_w, self.view_h
        x, y = _length(a.get("x", "0"), W, fs), _length(a.get("y", "0"), H, fs)
        w, h = _length(a.get("width", "100%"), W, fs), _length(a.get("height", "100%"), H, fs)
        if not (w > 0 and h > 0):
            return
        inner = (1.0, 0.0, 0.0, 1.0, x, y)
        vb = [float(v) for v in _NUM.findall(el.get("viewBox") or "")]
        view_w, view_h = w, h
        if len(vb) == 4:
            if not (vb[2] > 0 and vb[3] > 0):
                return
            sx, sy = w / vb[2], h / vb[3]
            align, slice_ = _aspect(el.get("preserveAspectRatio"))
            if align == "none":
                fit = (sx, 0.0, 0.0, sy, -vb[0] * sx, -vb[1] * sy)
            else:
                sc = max(sx, sy) if slice_ else min(sx, sy)
                fit = (sc, 0.0, 0.0, sc, (w - vb[2] * sc) * align[0] / 2.0 - vb[0] * sc,
                       (h - vb[3] * sc) * align[1] / 2.0 - vb[1] * sc)
            inner = _mul(inner, fit)
            view_w, view_h = vb[2], vb[3]
        clip = a.get("overflow", "hidden").strip() not in ("visible", "auto")
        if clip:
            self.draws.append(("clip", _rect_path(x, y, w, h, 0.0, 0.0), t, tk.FillRule.NonZero))
        saved = (self.view_w, self.view_h)
        self.view_w, self.view_h = view_w, view_h
        try:
            t2, t32b = _mul(t, inner), _concat_f32(t32, _to_f32(inner))
            for child in el:
                self._walk(child, a, t2, t32b)
        finally:
            self.view_w, self.view_h = saved
            if clip:
                self.draws.append(("pop",))

    def _dasharray(self, a, width):
        """usvg's dash list: lengths (percentages of the viewport diagonal), an odd list
        repeated; a negative entry or a zero sum leaves the stroke solid."""
        text = a.get("stroke-dasharray")
        if text is None or text.strip() in ("none", ""):
            return None
        diag = math.hypot(self.view_w, self.view_h) / math.sqrt(2)
        fs = _length(a["font-size"]) if "font-size" in a else FONT_SIZE
        try:
            dashes = [_length(w, diag, fs) for w in text.replace(",", " ").split()]
        except SvgUnsupported:
            return None
        if not dashes or any(d < 0 for d in dashes) or sum(dashes) <= 0:
            return None
        if any(d == 0 for d in dashes):
            raise SvgUnsupported("zero-length dashes")  # their caps are not modelled
        return dashes * 2 if len(dashes) % 2 else dashes

    def _origin(self, text):
        """transform-origin against the viewport (the view box of SVG content):
        keywords, lengths and percentages; a single value leaves the other centred."""
        words = text.replace(",", " ").split()
        if not words or len(words) > 2:
            raise SvgUnsupported(f"transform-origin {text!r}")
        kw = {"left": ("x", 0.0), "right": ("x", 1.0), "top": ("y", 0.0), "bottom": ("y", 1.0), "center": (None, 0.5)}
        if len(words) == 1:
            words = [words[0], "center"] if kw.get(words[0], ("x",))[0] != "y" else ["center", words[0]]
        if kw.get(words[0], (None,))[0] == "y" or kw.get(words[1], (None,))[0] == "x":
            words = [words[1], words[0]]

        def value(w, ref):
            if w in kw:
                return kw[w][1] * ref
            return _length(w, ref)

        return value(words[0], self.view_w), value(words[1], self.view_h)

    @staticmethod
    def _contains(outer, inner):
        return any(node is inner for node in outer.iter())

    @staticmethod
    def _fallback(text, current):
        if not text:
            return None
        c = _color(text, current)
        return (c[:3], c[3]) if c is not None and len(c) == 4 else c

    def _bbox(self, tag, a):
        """The object bounding box (geometry only) of a basic shape."""
        if tag not in SHAPES:
            return None
        path = self._shape(tag, a)
        if path is None:
            return None
        verbs, pts = path.arrays()
        pts = np.asarray(pts, np.float64).reshape(-1, 2)
        u = np.linspace(0.0, 1.0, 257)[:, None]
        out, i, last = [], 0, None
        for v in np.asarray(verbs):
            if v in (tk._abi.PATH_MOVE, tk._abi.PATH_LINE):
                last = pts[i]
                out.append(pts[i:i + 1])
                i += 1
            elif v == tk._abi.PATH_QUAD:  # the tight extremes, from dense samples
                p1, p2 = pts[i], pts[i + 1]
                out.append((1 - u) ** 2 * last + 2 * (1 - u) * u * p1 + u * u * p2)
                last, i = p2, i + 2
            elif v == tk._abi.PATH_CUBIC:
                p1, p2, p3 = pts[i], pts[i + 1], pts[i + 2]
                out.append((1 - u) ** 3 * last + 3 * (1 - u) ** 2 * u * p1 + 3 * (1 - u) * u * u * p2 + u ** 3 * p3)
                last, i = p3, i + 3
        if not out:
            return None
        pts = np.concatenate(out)
        return pts[:, 0].min(), pts[:, 1].min(), pts[:, 0].max(), pts[:, 1].max()

    def _walk(self, el, parent, transform, transform32):
        tag = _local(el.tag)
        if tag in ("title", "desc", "metadata", "defs", "clipPath", "linearGradient", "radialGradient", "style"):
            return
        if tag in UNSUPPORTED:
            raise SvgUnsupported(tag)
        a = _attrs(el, parent, self._css(el))
        if a.get("display") == "none":
            return
        langs = el.get("systemLanguage")
        if langs is not None:  # the system language is taken as en-US
            tags = [x.strip().lower() for x in langs.split(",")]
            if not any(x == "en" or x.startswith("en-") for x in tags):
                return
        has_t = "transform" in a  # an attribute, a CSS rule or the style attribute
        own = _transform(a.get("transform")) if has_t else IDENTITY
        origin = a.get("transform-origin")
        if has_t and origin is not None:  # translate(o) transform translate(-o)
            ox, oy = self._origin(origin)
            own = _mul((1.0, 0.0, 0.0, 1.0, ox, oy), _mul(own, (1.0, 0.0, 0.0, 1.0, -ox, -oy)))
        t = _mul(transform, own) if has_t else transform
        t32 = _concat_f32(transform32, _to_f32(own)) if has_t else transform32
        clip_ref = a.get("clip-path", "none").strip()
        if clip_ref != "none":
            layer = self._clip(clip_ref, a, tag, t)
            if layer[0] == "empty":
                return  # an empty clip hides the element
            self.draws.append(layer)
            try:
                a = dict(a)
                del a["clip-path"]
                self._walk_clipped(el, tag, a, t, t32)
            finally:
                self.draws.append(("pop",))
            return
        self._walk_clipped(el, tag, a, t, t32)

    def _walk_clipped(self, el, tag, a, t, t32):
        opacity = _opacity(a.get("opacity"))
        if opacity < 1.0 and tag not in SHAPES:
            raise SvgUnsupported(f"group opacity={a.get('opacity')}")
        if tag == "use":  # a clone of the referenced element under translate(x, y), inheriting from the use
            href = (el.get(XLINK_HREF) or el.get("href") or "").strip()
            target = self.ids.get(href[1:]) if href.startswith("#") else None
            if target is None:
                return
            if _local(target.tag) in ("svg", "symbol"):
                raise SvgUnsupported("use of an svg or symbol")
            if target in self.use_stack or any(target is p or self._contains(target, p) for p in self.use_stack):
                raise SvgUnsupported("recursive use")
            fs = _length(a["font-size"]) if "font-size" in a else FONT_SIZE
            x = _length(a.get("x", "0"), self.view_w, fs)
            y = _length(a.get("y", "0"), self.view_h, fs)
            shift = (1.0, 0.0, 0.0, 1.0, x, y)
            self.use_stack.append(el)
            try:
                inherited = {k: v for k, v in a.items() if k in INHERITED}
                self._walk(target, inherited, _mul(t, shift), _concat_f32(t32, _to_f32(shift)))
            finally:
                self.use_stack.pop()
            return
        if tag == "switch":  # the first child whose conditions pass (the system language is en-US)
            for child in el:
                if not isinstance(child.tag, str):
                    continue
                if child.get("requiredExtensions") is not None:
                    continue
                langs = child.get("systemLanguage")
                if langs is not None and not any(x.strip().lower() == "en" or x.strip().lower().startswith("en-")
                                                 for x in langs.split(",")):
                    continue
                self._walk(child, a, t, t32)
                break
            return
        if tag == "svg" and getattr(self, "_entered", False):
            self._nested_svg(el, a, t, t32)
            return
        if tag in ("svg", "g"):
            self._entered = True
            for child in el:
                self._walk(child, a, t, t32)
            return
        if tag == "image":
            self._image(el, a, t32)
            return
        path = self._shape(tag, a)
        if path is None or a.get("visibility") in ("hidden", "collapse"):
            return
        fill = self._paint(a.get("fill", "black"), tag, a)
        fill_op = _opacity(a.get("fill-opacity"))
        stroke = self._paint(a.get("stroke", "none"), tag, a, "stroke")
        stroke_op = _opacity(a.get("stroke-opacity"))
        if tag in ("line", "polyline"):
            fill = fill if tag == "polyline" else None
        if opacity < 1.0:
            # a shape's opacity fades its fill and stroke together as one layer; with one
            # paint that is the paint's own opacity times it
            if fill is not None and fill_op > 0 and stroke is not None and stroke_op > 0:
                raise SvgUnsupported("opacity on a filled and stroked shape")
            fill_op, stroke_op = fill_op * opacity, stroke_op * opacity
        rule = tk.FillRule.EvenOdd if a.get("fill-rule") == "evenodd" else tk.FillRule.NonZero
        first = len(self.draws)
        if fill is not None and fill_op > 0:
            self.draws.append(("shape", path, t, (fill, min(fill_op, 1.0), rule), None))
        if stroke is not None and stroke_op > 0:
            width = _length(a.get("stroke-width", "1"), math.hypot(self.view_w, self.view_h) / math.sqrt(2))
            if width > 0:
                jv = a.get("stroke-linejoin", "miter")
                if jv in ("miter-clip", "arcs"):
                    raise SvgUnsupported(f"stroke-linejoin {jv}")
                # an invalid value is ignored: the default (miter joins, butt caps)
                join = {"round": tk.Join.Round, "bevel": tk.Join.Bevel}.get(jv, tk.Join.Miter)
                cap = {"round": tk.Cap.Round, "square": tk.Cap.Square}.get(a.get("stroke-linecap", "butt"), tk.Cap.Butt)
                limit = max(_number(a.get("stroke-miterlimit"), 4.0), 1.0)
                dashes = self._dasharray(a, width)
                if dashes is not None:
                    fs = _length(a["font-size"]) if "font-size" in a else FONT_SIZE
                    offset = _length(a.get("stroke-dashoffset", "0"), math.hypot(self.view_w, self.view_h) / math.sqrt(2),
                                     fs)
                    path = _dash(path, dashes, offset)
                self.draws.append(("shape", path, t, None, (stroke, min(stroke_op, 1.0), width, join, cap, limit)))
        if len(self.draws) - first == 2 and _stroke_first(a.get("paint-order")):
            self.draws[first], self.draws[first + 1] = self.draws[first + 1], self.draws[first]

    def _image(self, el, a, t32):
        """usvg's image conversion (converter/image.rs) and the lowering of
        src/svg/scene_image.rs:25-83: the image's pixels on a rect path of its
        pixel size, under the node's f32 transform and placement."""
        if a.get("visibility") in ("hidden", "collapse"):
            return
        href = el.get(XLINK_HREF) or el.get("href")
        if not href:
            return
        data = _image_bytes(href, self.base_dir)
        if data is None:
            return
        image = decode_image(data)
        fs = _length(a["font-size"]) if "font-size" in a else FONT_SIZE
        W, H = self.view_w, self.view_h

        def given(k):
            v = a.get(k)
            if v is None or v.strip() == "auto":
                return None
            return _length(v, W if k in ("x", "width") else H, fs)

        x, y = given("x") or 0.0, given("y") or 0.0
        w, h = given("width"), given("height")
        f = np.float32
        iw, ih = f(image.width), f(image.height)
        if w is not None and h is None:
            h = float(ih * f(f(w) / iw))
        elif h is not None and w is None:
            w = float(iw * f(f(h) / ih))
        w = float(iw) if w is None else w
        h = float(ih) if h is None else h
        if not (w > 0 and h > 0):
            return  # usvg skips an image of invalid size
        aspect = _aspect(a.get("preserveAspectRatio"))
        ts, (vx, vy, vw, vh) = _image_placement((image.width, image.height), x, y, w, h, aspect)
        rendering = a.get("image-rendering", "optimizeQuality").strip()
        nearest = rendering in ("optimizeSpeed", "crisp-edges", "pixelated")
        if aspect[1]:  # a slice clips to the viewport rect, in the parent's space
            self.draws.append(("clip", _rect_path(vx, vy, vw, vh, 0.0, 0.0), t32))
        self.draws.append(("image", image, _concat_f32(t32, ts), tk.PATTERN_NEAREST if nearest else tk.PATTERN_BILINEAR))
        if aspect[1]:
            self.draws.append(("pop",))

    def _shape(self, tag, a):
        fs = _length(a["font-size"]) if "font-size" in a else FONT_SIZE
        L = lambda k, ref, default="0": _length(a.get(k, default), ref, fs)
        W, H = self.view_w, self.view_h
        if tag == "path":
            return parse_path(a.get("d"))
        if tag == "rect":
            x, y, w, h = L("x", W), L("y", H), L("width", W), L("height", H)
            if w <= 0 or h <= 0:
                return None
            rx, ry = a.get("rx"), a.get("ry")
            rx = L("rx", W) if rx not in (None, "auto") else None
            ry = L("ry", H) if ry not in (None, "auto") else None
            rx = None if rx is not None and rx < 0 else rx  # a negative radius is taken as auto
            ry = None if ry is not None and ry < 0 else ry
            if rx is None and ry is None:
                rx = ry = 0.0
            elif rx is None:
                rx = ry
            elif ry is None:
                ry = rx
            return _rect_path(x, y, w, h, min(max(rx, 0.0), w / 2), min(max(ry, 0.0), h / 2))
        if tag == "circle":
            r = L("r", math.hypot(W, H) / math.sqrt(2))
            return _ellipse_path(L("cx", W), L("cy", H), r, r) if r > 0 else None
        if tag == "ellipse":
            rx, ry = a.get("rx"), a.get("ry")
            rx = L("rx", W) if rx not in (None, "auto") else None
            ry = L("ry", H) if ry not in (None, "auto") else None
            if rx is None and ry is None:
                return None
            rx = None if rx is not None and rx < 0 else rx  # a negative radius is taken as auto
            ry = None if ry is not None and ry < 0 else ry
            if rx is None and ry is None:
                return None
            rx, ry = (ry if rx is None else rx), (rx if ry is None else ry)
            return _ellipse_path(L("cx", W), L("cy", H), rx, ry) if rx > 0 and ry > 0 else None
        if tag == "line":
            return tk.BezPath().move_to(L("x1", W), L("y1", H)).line_to(L("x2", W), L("y2", H))
        if tag in ("polyline", "polygon"):
            v = [float(x) for x in _NUM.findall(a.get("points", ""))]
            if len(v) < 4:
                return None
            p = tk.BezPath().move_to(v[0], v[1])
            for i in range(2, len(v) - 1, 2):
                p.line_to(v[i], v[i + 1])
            return p.close_path() if tag == "polygon" else p
        raise SvgUnsupported(tag)


def render_size(doc, target_width):
    """The canvas size the reference's fixture loader uses for a target width
    (examples/common/svg.rs:30-40: the SVG size scaled to the width, height rounded up)."""
    w, h = doc.width, doc.height
    iw, ih = max(1, int(math.ceil(w))), max(1, int(math.ceil(h)))
    nh = int(math.ceil(target_width * ih / iw))
    return target_width, max(nh, 1), target_width / w, nh / h


def load(text, target_width=300, base_dir=None):
    """Records an SVG document on a new Canvas at the target width; returns the
    canvas.  Relative image hrefs resolve against base_dir (none are read without it)."""
    doc = Document(text, base_dir)
    width, height, sx, sy = render_size(doc, target_width)
    canvas = tk.Canvas(width, height)
    scale = (sx, 0.0, 0.0, sy, 0.0, 0.0)
    for entry in doc.draws:
        if entry[0] == "image":
            _, image, abs32, sampling = entry
            m = _mul(scale, abs32)
            world_to_local = _inverse(m)
            if world_to_local is None:
                raise SvgUnsupported("non-invertible image transform")
            # PatternBrush::new_resource(scale(1 / w, 1 / h) * world_to_local, Pad, sampling, 255)
            brush_xf = _mul((1.0 / image.width, 0.0, 0.0, 1.0 / image.height, 0.0, 0.0), world_to_local)
            image_id = canvas.add_image(image)
            brush = tk.ImageBrush(image_id, tuple(_f32(v) for v in brush_xf), extend=0, sampling=sampling)
            canvas.push_path(_rect_path(0.0, 0.0, float(image.width), float(image.height), 0.0, 0.0), brush, m,
                             tk.FillRule.NonZero, TOLERANCE)
            continue
        if entry[0] == "clip":
            rule = entry[3] if len(entry) > 3 else tk.FillRule.NonZero
            canvas.push_clip_layer(entry[1], _mul(scale, entry[2]), rule, TOLERANCE)
            continue
        if entry[0] == "pop":
            canvas.pop_layer()
            continue
        _, path, t, fill, stroke = entry
        m = _mul(scale, t)
        if fill is not None and isinstance(fill[0], dict):
            canvas.push_path(path, _gradient_brush(fill[0], fill[1], _mul(m, fill[0]["transform"])), m, fill[2],
                             TOLERANCE)
        elif fill is not None and isinstance(fill[0], tuple) and len(fill[0]) == 2:  # a one-stop gradient
            (rgb, alpha), op, rule = fill
            canvas.push_path(path, tk.Color(rgb[0] / 255, rgb[1] / 255, rgb[2] / 255, alpha * op), m, rule, TOLERANCE)
        elif fill is not None:
            (r, g, b), op, rule = fill
            canvas.push_path(path, tk.Color(r / 255, g / 255, b / 255, op), m, rule, TOLERANCE)
        if stroke is not None:
            paint, op, width_, join, cap, limit = stroke
            if isinstance(paint, dict):
                brush = _gradient_brush(paint, op, _mul(m, paint["transform"]))
            elif len(paint) == 2:
                brush = tk.Color(paint[0][0] / 255, paint[0][1] / 255, paint[0][2] / 255, paint[1] * op)
            else:
                brush = tk.Color(paint[0] / 255, paint[1] / 255, paint[2] / 255, op)
            canvas.push_stroke(path, width_, brush, m, tk.FillRule.NonZero,
                               TOLERANCE, join=join, start_cap=cap, end_cap=cap, miter_limit=limit)
    return canvas


def _gradient_brush(g, opacity, server_to_canvas):
    """The reference's SVG gradient brush (src/svg/paint.rs:146-210): stops with
    their opacity times the paint's, interpolated in straight alpha, the brush
    transform the f32 inverse of (canvas transform x path transform x server
    transform)."""
    inv = _inverse(server_to_canvas)
    if inv is None:
        raise SvgUnsupported("non-invertible gradientTransform")
    stops = [(o, tk.Color(c[0] / 255, c[1] / 255, c[2] / 255, float(np.float32(c[3]) * np.float32(opacity))))
             for o, c in g["stops"]]
    xf = tuple(_f32(v) for v in inv)
    geom = g["geom"]
    if geom[0] == "linear":
        return tk.LinearGradient((geom[1], geom[2]), (geom[3], geom[4]), stops, g["spread"], transform=xf,
                                 unpremultiplied=True)
    _, fx, fy, fr, cx, cy, r = geom
    return tk.RadialGradient((fx, fy), fr, (cx, cy), r, stops, g["spread"], transform=xf, unpremultiplied=True)


def straight_rgba8(premultiplied):
    """Premultiplied RGBA8 words (H, W) as straight-alpha RGBA bytes (H, W, 4), with
    the rounding of the reference's PNG writer (src/shared/image.rs:198-214)."""
    px = np.ascontiguousarray(premultiplied, np.uint32).view(np.uint8).reshape(premultiplied.shape + (4,)).astype(
        np.uint32)
    a = px[..., 3:4]
    rgb = np.where(a == 0, 0, np.minimum((px[..., :3] * 255 + a // 2) // np.maximum(a, 1), 255))
    return np.concatenate([rgb, a], -1).astype(np.uint8)
