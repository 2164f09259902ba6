/*
 * oracle.c — CPU restatement of Tileink's compute-raster stages (test
 * infrastructure; see oracle.h).  Build: gcc -O2 -std=c11 -ffp-contract=off.
 *
 * Float semantics follow the reference shaders: HLSL `mad` is fmaf, every other
 * product and sum is rounded separately (no contraction), HLSL round() is
 * round-half-to-even (rintf), and float->int conversions saturate with NaN -> 0.
 * The CPU segment-capacity routine (scan_line.rs) uses Rust semantics instead:
 * round() is half-away-from-zero (roundf).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define NONE 0xFFFFFFFFu
#define TILE 16u

/* ---------- threads ----------------------------------------------------------
 * Stages whose work items are independent (lines of distinct paths, scan and
 * cumsum chunks, tiles) run over a pool of threads taking items in chunks from
 * a shared counter.  Integer results do not depend on the thread count: scan
 * counts add with atomics, and each path's lines stay with one thread, in
 * order, so emit claims its segment slots in index order exactly as one thread
 * does. */

static uint32_t g_threads = 1;

void or_set_threads(uint32_t n) { g_threads = n == 0u ? 1u : (n > 256u ? 256u : n); }
uint32_t or_get_threads(void) { return g_threads; }

typedef void (*RangeFn)(void *ctx, uint32_t begin, uint32_t end);
typedef struct {
    RangeFn fn;
    void *ctx;
    uint64_t n, grain, next;
} ParJob;

static void *par_worker(void *arg) {
    ParJob *j = (ParJob *)arg;
    for (;;) {
        uint64_t b = __atomic_fetch_add(&j->next, j->grain, __ATOMIC_RELAXED);
        if (b >= j->n) break;
        uint64_t e = b + j->grain < j->n ? b + j->grain : j->n;
        j->fn(j->ctx, (uint32_t)b, (uint32_t)e);
    }
    return NULL;
}

/* fn(ctx, begin, end) over [0, n) in chunks of `grain` items. */
static void par_for(uint32_t n, uint32_t grain, RangeFn fn, void *ctx) {
    if (n == 0u) return;
    if (grain == 0u) grain = 1u;
    if (g_threads <= 1u || n <= grain) {
        fn(ctx, 0u, n);
        return;
    }
    ParJob j = {fn, ctx, n, grain, 0u};
    pthread_t tid[256];
    uint32_t spawned = 0;
    for (uint32_t t = 1; t < g_threads; ++t)
        if (pthread_create(&tid[spawned], NULL, par_worker, &j) == 0) spawned++;
    par_worker(&j);
    for (uint32_t t = 0; t < spawned; ++t) pthread_join(tid[t], NULL);
}

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

static void add_i32(int32_t *p, int32_t v) { __atomic_fetch_add((uint32_t *)p, (uint32_t)v, __ATOMIC_RELAXED); }
static void add_u32(uint32_t *p, uint32_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }

/* ---------- scalar helpers --------------------------------------------------- */

static uint32_t f2u(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}
static int32_t f2i(float v) {
    if (v != v) return 0;
    if (v <= -2147483648.0f) return INT32_MIN;
    if (v >= 2147483648.0f) return INT32_MAX;
    return (int32_t)v;
}
static float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
static float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
static uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

/* ---------- scan geometry (src/shaders/hlsl/scan/geometry.hlsli:8-114) ---------- */

static const float SCAN_EPSILON = 1.0e-6f;
static const float TOP_TOUCH_EPSILON = 1.0e-12f;
static const float TILE_CLIP_NUDGE = 1.0e-3f;

typedef struct {
    uint32_t bbox[4]; /* x0 y0 x1 y1 */
    uint32_t data_offset;
    float pts[4];     /* oriented: pts[3] >= pts[1] */
    int down;
    uint32_t count;
    float a, b, sign, x0, y0;
    uint32_t imin, imax;
    int32_t ymin, ymax;
} Traversal;

static int is_tile_boundary_y(float v) { return fabsf(v - floorf(v)) <= TOP_TOUCH_EPSILON; }
static int32_t ceil_tile_boundary_y(float v) { return is_tile_boundary_y(v) ? f2i(floorf(v)) : f2i(ceilf(v)); }
static uint32_t span(float a, float b) {
    float hi = ceilf(a);
    if (b > a) hi = ceilf(b);
    float lo = floorf(a);
    if (b < a) lo = floorf(b);
    float value = hi - lo;
    if (value < 1.0f) value = 1.0f;
    return f2u(value);
}

/* scan_geometry (scan/geometry.hlsli:26-104). */
static int scan_geometry(const uint32_t *lines, const uint32_t *paths, uint32_t path_records, uint32_t line_index,
                         Traversal *s) {
    memset(s, 0, sizeof(*s));
    const uint32_t *line = lines + (size_t)line_index * 6u;
    uint32_t path_id = line[0];
    if (path_id >= path_records) return 0;
    const uint32_t *path = paths + (size_t)path_id * 19u;
    for (int i = 0; i < 4; ++i) s->bbox[i] = path[6 + i];
    s->data_offset = path[4];
    if (s->bbox[2] - s->bbox[0] == 0u || s->bbox[1] >= s->bbox[3]) return 0;
    float ta = bits_f(path[13]), tb = bits_f(path[14]), tc = bits_f(path[15]);
    float td = bits_f(path[16]), te = bits_f(path[17]), tf = bits_f(path[18]);
    float l0x = bits_f(line[2]), l0y = bits_f(line[3]), l1x = bits_f(line[4]), l1y = bits_f(line[5]);
    float p0x = ta * l0x + tc * l0y + te, p0y = tb * l0x + td * l0y + tf;
    float p1x = ta * l1x + tc * l1y + te, p1y = tb * l1x + td * l1y + tf;
    s->down = p1y >= p0y;
    if (s->down) {
        s->pts[0] = p0x; s->pts[1] = p0y; s->pts[2] = p1x; s->pts[3] = p1y;
    } else {
        s->pts[0] = p1x; s->pts[1] = p1y; s->pts[2] = p0x; s->pts[3] = p0y;
    }
    const float scale = 1.0f / (float)TILE;
    float sx = s->pts[0] * scale, sy = s->pts[1] * scale, sz = s->pts[2] * scale, sw = s->pts[3] * scale;
    uint32_t count_x = span(sx, sz) - 1u;
    s->count = count_x + span(sy, sw);
    float dx = fabsf(sz - sx), dy = sw - sy;
    if (dx + dy == 0.0f || (dy == 0.0f && floorf(sy) == sy)) return 0;
    float reciprocal = 1.0f / (dx + dy);
    s->a = dx * reciprocal;
    int positive = sz >= sx;
    s->sign = positive ? 1.0f : -1.0f;
    float xt0 = floorf(sx * s->sign);
    float c = sx * s->sign - xt0;
    s->y0 = floorf(sy);
    float ytop = s->y0 + 1.0f;
    if (sy == sw) ytop = ceilf(sy);
    s->b = fminf((dy * c + dx * (ytop - sy)) * reciprocal, 0.99999994f);
    float error = floorf(s->a * ((float)s->count - 1.0f) + s->b) - (float)count_x;
    if (error != 0.0f) {
        if (error > 0.0f) s->a -= 0.0000002f;
        else s->a += 0.0000002f;
    }
    s->x0 = xt0 * s->sign - 1.0f;
    if (positive) s->x0 = xt0 * s->sign;
    float xmin = fminf(sx, sz);
    float bx = (float)s->bbox[0], by = (float)s->bbox[1], bz = (float)s->bbox[2], bw = (float)s->bbox[3];
    if (sy >= bw || sw <= by + TOP_TOUCH_EPSILON || xmin >= bz) return 0;
    s->imin = 0u;
    if (sy < by) {
        float first = rintf((by - s->y0 + s->b - s->a) / (1.0f - s->a)) - 1.0f;
        if (s->y0 + first - floorf(s->a * first + s->b) < by) first += 1.0f;
        s->imin = f2u(first);
    }
    s->imax = s->count;
    if (sw > bw) {
        float last = rintf((bw - s->y0 + s->b - s->a) / (1.0f - s->a)) - 1.0f;
        if (s->y0 + last - floorf(s->a * last + s->b) < bw) last += 1.0f;
        s->imax = f2u(last);
    }
    s->ymin = 0;
    s->ymax = 0;
    if (fmaxf(sx, sz) < bx) {
        s->ymin = ceil_tile_boundary_y(sy);
        s->ymax = ceil_tile_boundary_y(sw);
        s->imax = s->imin;
    } else {
        float fudge = positive ? 0.0f : 1.0f;
        if (xmin < bx) {
            float cut = rintf((s->sign * (bx - s->x0) - s->b + fudge) / s->a);
            if ((s->x0 + s->sign * floorf(s->a * cut + s->b) < bx) == positive) cut += 1.0f;
            int32_t ynext = f2i(s->y0 + cut - floorf(s->a * cut + s->b) + 1.0f);
            if (positive) {
                if (f2u(cut) > s->imin) {
                    float ystart = s->y0 + 1.0f;
                    if (is_tile_boundary_y(sy)) ystart = s->y0;
                    s->ymin = f2i(ystart);
                    s->ymax = ynext;
                    s->imin = f2u(cut);
                }
            } else if (f2u(cut) < s->imax) {
                s->ymin = ynext;
                s->ymax = ceil_tile_boundary_y(sw);
                s->imax = f2u(cut);
            }
        }
        if (fmaxf(sx, sz) > bz) {
            float cut = rintf((s->sign * (bz - s->x0) - s->b + fudge) / s->a);
            if ((s->x0 + s->sign * floorf(s->a * cut + s->b) < bz) == positive) cut += 1.0f;
            if (positive) s->imax = umin(s->imax, f2u(cut));
            else s->imin = umax(s->imin, f2u(cut));
        }
    }
    s->imax = umax(s->imin, s->imax);
    if (s->ymin < (int32_t)s->bbox[1]) s->ymin = (int32_t)s->bbox[1];
    if (s->ymax > (int32_t)s->bbox[3]) s->ymax = (int32_t)s->bbox[3];
    return 1;
}

static void scan_tile(const Traversal *s, uint32_t index, float z, int32_t *tx, int32_t *ty) {
    *tx = f2i(s->x0 + s->sign * z);
    *ty = f2i(s->y0 + (float)index - z);
}
static int scan_tile_inside(const Traversal *s, int32_t tx, int32_t ty) {
    return ty >= (int32_t)s->bbox[1] && ty < (int32_t)s->bbox[3] && tx >= (int32_t)s->bbox[0] &&
           tx < (int32_t)s->bbox[2];
}
static uint32_t scan_tile_offset(const Traversal *s, int32_t tx, int32_t ty) {
    return s->data_offset + ((uint32_t)ty - s->bbox[1]) * (s->bbox[2] - s->bbox[0]) + (uint32_t)tx - s->bbox[0];
}

/* The scan's work items: lines [begin, end) of `lines`. */
typedef struct {
    const or_scan_config *cfg;
    const uint32_t *lines, *paths;
    uint32_t path_record_count;
    int32_t *backdrops;
    uint32_t *counts, *cursors, *segments;
    const uint32_t *bounds; /* chunk c is lines [bounds[c], bounds[c + 1]) */
} ScanJob;

/* Line-range chunks that never split a run of lines of one path. */
static uint32_t *path_aligned_bounds(const uint32_t *lines, uint32_t line_count, uint32_t *chunk_count) {
    uint32_t k = line_count / 4096u + 1u;
    uint32_t *b = (uint32_t *)malloc(sizeof(uint32_t) * (k + 1u));
    uint32_t n = 0;
    b[n++] = 0u;
    for (uint32_t c = 1; c < k; ++c) {
        uint32_t at = (uint32_t)((uint64_t)line_count * c / k);
        if (at <= b[n - 1]) continue;
        while (at < line_count && lines[(size_t)at * 6u] == lines[(size_t)(at - 1u) * 6u]) ++at;
        if (at < line_count && at > b[n - 1]) b[n++] = at;
    }
    b[n] = line_count;
    *chunk_count = n;
    return b;
}

static void scan_count_lines(void *ctx, uint32_t c0, uint32_t c1) {
    const ScanJob *j = (const ScanJob *)ctx;
    int32_t *backdrops = j->backdrops;
    for (uint32_t id = j->bounds[c0]; id < j->bounds[c1]; ++id) {
        Traversal s;
        if (!scan_geometry(j->lines, j->paths, j->path_record_count, id, &s)) continue;
        int32_t delta = s.down ? -1 : 1;
        uint32_t stride = s.bbox[2] - s.bbox[0];
        for (int32_t y = s.ymin; y < s.ymax; y++) {
            uint32_t local = (uint32_t)(y - (int32_t)s.bbox[1]) * stride;
            add_i32(&backdrops[s.data_offset + local], delta);
        }
        float last_z = floorf(s.a * ((float)s.imin - 1.0f) + s.b);
        for (uint32_t i = s.imin; i < s.imax; i++) {
            float z = floorf(s.a * (float)i + s.b);
            int32_t tx, ty;
            scan_tile(&s, i, z, &tx, &ty);
            if (scan_tile_inside(&s, tx, ty)) {
                int top = last_z == z;
                if (i == 0u) top = fabsf(s.y0 - s.pts[1] * (1.0f / (float)TILE)) <= SCAN_EPSILON;
                if (top && tx + 1 < (int32_t)s.bbox[2]) {
                    int32_t x = tx + 1 > (int32_t)s.bbox[0] ? tx + 1 : (int32_t)s.bbox[0];
                    uint32_t local = (uint32_t)(ty - (int32_t)s.bbox[1]) * stride + (uint32_t)(x - (int32_t)s.bbox[0]);
                    add_i32(&backdrops[s.data_offset + local], delta);
                }
                add_u32(&j->counts[scan_tile_offset(&s, tx, ty)], 1u);
            }
            last_z = z;
        }
    }
}

/* scan_count (scan/count.hlsl:13-40). */
void or_scan_count(const or_scan_config *cfg, const uint32_t *lines, const uint32_t *paths,
                   uint32_t path_record_count, int32_t *backdrops, uint32_t *counts) {
    uint32_t chunks;
    uint32_t *bounds = path_aligned_bounds(lines, cfg->line_count, &chunks);
    ScanJob j = {cfg, lines, paths, path_record_count, backdrops, counts, NULL, NULL, bounds};
    par_for(chunks, 1u, scan_count_lines, &j);
    free(bounds);
}

/* scan_prefix_chunks (scan/prefix_chunks.hlsl:16-38). chunks: 4 words {path, backdrop_offset, segment_start, len}. */
typedef struct {
    const or_scan_config *cfg;
    const uint32_t *chunks, *counts, *paths, *chunk_ranges;
    uint32_t *ranges, *chunk_totals, *bumps, *chunk_offsets, *cursors;
} PrefixJob;
static void prefix_chunk_range(void *ctx, uint32_t k0, uint32_t k1);
static void chunk_offset_range(void *ctx, uint32_t p0, uint32_t p1);
static void apply_chunk_range(void *ctx, uint32_t k0, uint32_t k1);

void or_scan_prefix_chunks(const or_scan_config *cfg, const uint32_t *chunks, const uint32_t *counts,
                           uint32_t *ranges, uint32_t *chunk_totals) {
    PrefixJob j = {cfg, chunks, counts, NULL, NULL, ranges, chunk_totals, NULL, NULL, NULL};
    par_for(cfg->scan_chunk_count, 256u, prefix_chunk_range, &j);
}
static void prefix_chunk_range(void *ctx, uint32_t k0, uint32_t k1) {
    const PrefixJob *j = (const PrefixJob *)ctx;
    const uint32_t *chunks = j->chunks, *counts = j->counts;
    uint32_t *ranges = j->ranges, *chunk_totals = j->chunk_totals;
    for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t *chunk = chunks + (size_t)k * 4u;
        uint32_t inclusive = 0u;
        for (uint32_t lane = 0; lane < 256u; ++lane) {
            uint32_t count = lane < chunk[3] ? counts[chunk[1] + lane] : 0u;
            inclusive += count;
            if (lane < chunk[3]) {
                ranges[(size_t)(chunk[1] + lane) * 2u] = inclusive - count;
                ranges[(size_t)(chunk[1] + lane) * 2u + 1u] = inclusive;
            }
            if (lane + 1u == chunk[3]) chunk_totals[k] = inclusive;
        }
    }
}

/* scan_chunk_offsets (scan/chunk_offsets.hlsl:14-26). */
void or_scan_chunk_offsets(const or_scan_config *cfg, const uint32_t *paths, const uint32_t *chunk_ranges,
                           uint32_t *bumps, const uint32_t *chunk_totals, uint32_t *chunk_offsets) {
    PrefixJob j = {cfg, NULL, NULL, paths, chunk_ranges, NULL, (uint32_t *)chunk_totals, bumps, chunk_offsets, NULL};
    par_for(cfg->path_count, 1024u, chunk_offset_range, &j);
}
static void chunk_offset_range(void *ctx, uint32_t p0, uint32_t p1) {
    const PrefixJob *j = (const PrefixJob *)ctx;
    const uint32_t *paths = j->paths, *chunk_ranges = j->chunk_ranges, *chunk_totals = j->chunk_totals;
    uint32_t *bumps = j->bumps, *chunk_offsets = j->chunk_offsets;
    for (uint32_t p = p0; p < p1; ++p) {
        uint32_t start = paths[(size_t)p * 19u + 10u];
        uint32_t next = start;
        for (uint32_t c = chunk_ranges[2 * p]; c < chunk_ranges[2 * p + 1]; ++c) {
            chunk_offsets[c] = next;
            next += chunk_totals[c];
        }
        bumps[p] = next - start;
    }
}

/* scan_apply_chunk_offsets (scan/apply_chunk_offsets.hlsl:15-27). */
void or_scan_apply_chunk_offsets(const or_scan_config *cfg, const uint32_t *chunks, uint32_t *ranges,
                                 uint32_t *cursors, const uint32_t *chunk_offsets) {
    PrefixJob j = {cfg, chunks, NULL, NULL, NULL, ranges, NULL, NULL, (uint32_t *)chunk_offsets, cursors};
    par_for(cfg->scan_chunk_count, 256u, apply_chunk_range, &j);
}
static void apply_chunk_range(void *ctx, uint32_t k0, uint32_t k1) {
    const PrefixJob *j = (const PrefixJob *)ctx;
    const uint32_t *chunks = j->chunks, *chunk_offsets = j->chunk_offsets;
    uint32_t *ranges = j->ranges, *cursors = j->cursors;
    for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t *chunk = chunks + (size_t)k * 4u;
        uint32_t base = chunk_offsets[k];
        for (uint32_t lane = 0; lane < 256u && lane < chunk[3]; ++lane) {
            size_t index = chunk[1] + lane;
            ranges[index * 2u] += base;
            ranges[index * 2u + 1u] += base;
            cursors[index] = ranges[index * 2u];
        }
    }
}

/* x_at_y / y_at_x / clip_y_at_x / write_clipped_segment (scan/clip.hlsli:10-75). */
static float x_at_y(const float *p, float y) { return fmaf(p[2] - p[0], (y - p[1]) / (p[3] - p[1]), p[0]); }
static float y_at_x(const float *p, float x) { return fmaf(p[3] - p[1], (x - p[0]) / (p[2] - p[0]), p[1]); }
static float clip_y_at_x(const float *p, float x, float lo, float hi) {
    float y = y_at_x(p, x);
    if (y <= lo + SCAN_EPSILON) {
        float top_x = x_at_y(p, lo);
        if (fabsf(top_x - x) > TILE_CLIP_NUDGE) return lo;
        return lo + TILE_CLIP_NUDGE;
    }
    return clampf(y, lo + TILE_CLIP_NUDGE, hi);
}

static void write_clipped_segment(uint32_t *out, const Traversal *s, uint32_t sub, float z, int32_t tx, int32_t ty) {
    const float size = (float)TILE;
    float minx = (float)tx * size, miny = (float)ty * size;
    float maxx = minx + size, maxy = miny + size;
    float q[4] = {s->pts[0], s->pts[1], s->pts[2], s->pts[3]};
    if (sub > 0u) {
        float previous = floorf(s->a * ((float)sub - 1.0f) + s->b);
        if (z == previous) {
            float x = clampf(x_at_y(s->pts, miny), minx + TILE_CLIP_NUDGE, maxx);
            q[0] = x;
            q[1] = miny;
        } else {
            float xc = maxx;
            if (s->sign > 0.0f) xc = minx;
            q[0] = xc;
            q[1] = clip_y_at_x(s->pts, xc, miny, maxy);
        }
    }
    if (sub < s->count - 1u) {
        float next = floorf(s->a * ((float)sub + 1.0f) + s->b);
        if (z == next) {
            float x = clampf(x_at_y(s->pts, maxy), minx + TILE_CLIP_NUDGE, maxx);
            q[2] = x;
            q[3] = maxy;
        } else {
            float xc = minx;
            if (s->sign > 0.0f) xc = maxx;
            q[2] = xc;
            q[3] = clip_y_at_x(s->pts, xc, miny, maxy);
        }
    }
    float edge = 1000000000.0f;
    float p0x = clampf(q[0] - minx, 0.0f, size), p0y = clampf(q[1] - miny, 0.0f, size);
    float p1x = clampf(q[2] - minx, 0.0f, size), p1y = clampf(q[3] - miny, 0.0f, size);
    if (p0x <= SCAN_EPSILON) p0x = 0.0f; else if (size - p0x <= SCAN_EPSILON) p0x = size;
    if (p0y <= SCAN_EPSILON) p0y = 0.0f; else if (size - p0y <= SCAN_EPSILON) p0y = size;
    if (p1x <= SCAN_EPSILON) p1x = 0.0f; else if (size - p1x <= SCAN_EPSILON) p1x = size;
    if (p1y <= SCAN_EPSILON) p1y = 0.0f; else if (size - p1y <= SCAN_EPSILON) p1y = size;
    if (p0x == 0.0f) {
        if (p1x == 0.0f) {
            p0x = SCAN_EPSILON;
            if (p0y == 0.0f) { p1x = SCAN_EPSILON; p1y = size; }
            else { p1x = 2.0f * SCAN_EPSILON; p1y = p0y; }
        } else if (p0y == 0.0f) {
            p0x = SCAN_EPSILON;
        } else {
            edge = p0y;
        }
    } else if (p1x == 0.0f) {
        if (p1y == 0.0f) p1x = SCAN_EPSILON;
        else edge = p1y;
    }
    if (floorf(p0x) == p0x && p0x != 0.0f) p0x -= SCAN_EPSILON;
    if (floorf(p1x) == p1x && p1x != 0.0f) p1x -= SCAN_EPSILON;
    if (!s->down) {
        float x = p0x, y = p0y;
        p0x = p1x; p0y = p1y; p1x = x; p1y = y;
    }
    out[0] = f_bits(p0x);
    out[1] = f_bits(p0y);
    out[2] = f_bits(p1x);
    out[3] = f_bits(p1y);
    out[4] = f_bits(edge);
}

static void scan_emit_lines(void *ctx, uint32_t c0, uint32_t c1) {
    const ScanJob *j = (const ScanJob *)ctx;
    for (uint32_t id = j->bounds[c0]; id < j->bounds[c1]; ++id) {
        Traversal s;
        if (!scan_geometry(j->lines, j->paths, j->path_record_count, id, &s)) continue;
        for (uint32_t i = s.imin; i < s.imax; i++) {
            float z = floorf(s.a * (float)i + s.b);
            int32_t tx, ty;
            scan_tile(&s, i, z, &tx, &ty);
            if (scan_tile_inside(&s, tx, ty)) {
                uint32_t dst = __atomic_fetch_add(&j->cursors[scan_tile_offset(&s, tx, ty)], 1u, __ATOMIC_RELAXED);
                if (dst < j->cfg->segment_capacity)
                    write_clipped_segment(j->segments + (size_t)dst * 5u, &s, i, z, tx, ty);
            }
        }
    }
}

/* scan_emit (scan/emit.hlsl:14-28), each path's lines in index order. */
void or_scan_emit(const or_scan_config *cfg, const uint32_t *lines, const uint32_t *paths,
                  uint32_t path_record_count, uint32_t *cursors, uint32_t *segments) {
    uint32_t chunks;
    uint32_t *bounds = path_aligned_bounds(lines, cfg->line_count, &chunks);
    ScanJob j = {cfg, lines, paths, path_record_count, NULL, NULL, cursors, segments, bounds};
    par_for(chunks, 1u, scan_emit_lines, &j);
    free(bounds);
}

/* cumsum_prefix_chunks, cumsum_chunk_offsets, cumsum_apply_chunk_offsets (cumsum.hlsl:16-71). */
typedef struct {
    const uint32_t *offs, *lens;
    uint32_t *totals, *carries;
    const uint32_t *row_starts, *row_ends;
    int32_t *backdrops;
} CumsumJob;
static void cumsum_prefix_range(void *ctx, uint32_t c0, uint32_t c1) {
    const CumsumJob *j = (const CumsumJob *)ctx;
    for (uint32_t c = c0; c < c1; ++c) {
        uint32_t sum = 0u;
        for (uint32_t lane = 0; lane < 256u; ++lane) {
            uint32_t v = lane < j->lens[c] ? (uint32_t)j->backdrops[j->offs[c] + lane] : 0u;
            sum += v;
            if (lane < j->lens[c]) j->backdrops[j->offs[c] + lane] = (int32_t)sum;
        }
        j->totals[c] = sum;
    }
}
static void cumsum_carry_range(void *ctx, uint32_t r0, uint32_t r1) {
    const CumsumJob *j = (const CumsumJob *)ctx;
    for (uint32_t r = r0; r < r1; ++r) {
        uint32_t carry = 0u;
        for (uint32_t c = j->row_starts[r]; c < j->row_ends[r]; ++c) {
            j->carries[c] = carry;
            carry += j->totals[c];
        }
    }
}
static void cumsum_apply_range(void *ctx, uint32_t c0, uint32_t c1) {
    const CumsumJob *j = (const CumsumJob *)ctx;
    for (uint32_t c = c0; c < c1; ++c)
        for (uint32_t lane = 0; lane < j->lens[c] && lane < 256u; ++lane)
            j->backdrops[j->offs[c] + lane] = (int32_t)((uint32_t)j->backdrops[j->offs[c] + lane] + j->carries[c]);
}

void or_cumsum(uint32_t row_count, uint32_t chunk_count, const uint32_t *offs, const uint32_t *lens,
               const uint32_t *row_starts, const uint32_t *row_ends, int32_t *backdrops) {
    uint32_t *totals = (uint32_t *)calloc(chunk_count ? chunk_count : 1, 4);
    uint32_t *carries = (uint32_t *)calloc(chunk_count ? chunk_count : 1, 4);
    CumsumJob j = {offs, lens, totals, carries, row_starts, row_ends, backdrops};
    par_for(chunk_count, 1024u, cumsum_prefix_range, &j);
    par_for(row_count, 1024u, cumsum_carry_range, &j);
    par_for(chunk_count, 1024u, cumsum_apply_range, &j);
    free(totals);
    free(carries);
}

/* line_scanned_tile_count / plan_scan_line (src/shared/scan_line.rs:17-151). */
static uint32_t rs_u32(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967295.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}
static int32_t rs_i32(float v) {
    if (v != v) return 0;
    if (v <= -2147483648.0f) return INT32_MIN;
    if (v >= 2147483647.0f) return INT32_MAX;
    return (int32_t)v;
}
static uint32_t rs_span(float a, float b) {
    return rs_u32(fmaxf(ceilf(fmaxf(a, b)) - floorf(fminf(a, b)), 1.0f));
}
uint32_t or_line_scanned_tile_count(const float p0[2], const float p1[2], uint32_t bx0, uint32_t by0, uint32_t bx1,
                                    uint32_t by1) {
    int is_down = p1[1] >= p0[1];
    const float *xy0 = is_down ? p0 : p1, *xy1 = is_down ? p1 : p0;
    const float scale = 1.0f / 16.0f;
    float s0x = xy0[0] * scale, s0y = xy0[1] * scale, s1x = xy1[0] * scale, s1y = xy1[1] * scale;
    uint32_t count_x = rs_span(s0x, s1x) - 1u;
    uint32_t count = count_x + rs_span(s0y, s1y);
    float dx = fabsf(s1x - s0x), dy = s1y - s0y;
    if (dx + dy == 0.0f) return 0;
    if (dy == 0.0f && floorf(s0y) == s0y) return 0;
    float idxdy = 1.0f / (dx + dy);
    float a = dx * idxdy;
    int pos = s1x >= s0x;
    float sign = pos ? 1.0f : -1.0f;
    float xt0 = floorf(s0x * sign);
    float c = s0x * sign - xt0;
    float y0 = floorf(s0y);
    float ytop = s0y == s1y ? ceilf(s0y) : y0 + 1.0f;
    float b = fminf((dy * c + dx * (ytop - s0y)) * idxdy, 0.99999994f);
    float err = floorf(a * ((float)count - 1.0f) + b) - (float)count_x;
    if (err != 0.0f) a -= copysignf(2e-7f, err);
    float x0 = xt0 * sign + (pos ? 0.0f : -1.0f);
    float xmin = fminf(s0x, s1x);
    if (s0y >= (float)by1 || s1y <= (float)by0 + 1.0e-12f || xmin >= (float)bx1) return 0;
    uint32_t imin = 0, imax = count;
    if (s0y < (float)by0) {
        float f = roundf(((float)by0 - y0 + b - a) / (1.0f - a)) - 1.0f;
        if (y0 + f - floorf(a * f + b) < (float)by0) f += 1.0f;
        imin = rs_u32(f);
    }
    if (s1y > (float)by1) {
        float f = roundf(((float)by1 - y0 + b - a) / (1.0f - a)) - 1.0f;
        if (y0 + f - floorf(a * f + b) < (float)by1) f += 1.0f;
        imax = rs_u32(f);
    }
    if (fmaxf(s0x, s1x) < (float)bx0) {
        imax = imin;
    } else {
        float fudge = pos ? 0.0f : 1.0f;
        if (xmin < (float)bx0) {
            float f = roundf((sign * ((float)bx0 - x0) - b + fudge) / a);
            if ((x0 + sign * floorf(a * f + b) < (float)bx0) == pos) f += 1.0f;
            if (pos) {
                if (rs_u32(f) > imin) imin = rs_u32(f);
            } else if (rs_u32(f) < imax) {
                imax = rs_u32(f);
            }
        }
        if (fmaxf(s0x, s1x) > (float)bx1) {
            float f = roundf((sign * ((float)bx1 - x0) - b + fudge) / a);
            if ((x0 + sign * floorf(a * f + b) < (float)bx1) == pos) f += 1.0f;
            if (pos) imax = umin(imax, rs_u32(f));
            else imin = umax(imin, rs_u32(f));
        }
    }
    imax = umax(imin, imax);
    uint32_t total = 0;
    for (uint32_t i = imin; i < imax; ++i) {
        float z = floorf(a * (float)i + b);
        int32_t y = rs_i32(y0 + (float)i - z), x = rs_i32(x0 + sign * z);
        total += (y >= (int32_t)by0 && y < (int32_t)by1 && x >= (int32_t)bx0 && x < (int32_t)bx1);
    }
    return total;
}

/* ---------- pixel math (shared/pixel.hlsli:7-66) ----------------------------- */

static const float CHANNEL_SCALE = 1.0f / 255.0f;
static uint32_t coverage_to_u8(float c) { return f2u(clampf(c, 0.0f, 1.0f) * 255.0f + 0.5f); }
static uint32_t coverage_to_alpha(float v, uint32_t rule) {
    float alpha = fminf(fabsf(v), 1.0f);
    if (rule == 1u) alpha = fabsf(v - 2.0f * rintf(0.5f * v));
    return coverage_to_u8(alpha);
}
static float signum_f32(float v) { return v < 0.0f ? -1.0f : 1.0f; }
static uint32_t pack4(uint32_t r, uint32_t g, uint32_t b, uint32_t a) { return r | (g << 8) | (b << 16) | (a << 24); }
typedef struct { float c[4]; } Px;
static uint32_t unorm_to_rgba8(Px p) {
    uint32_t ch[4];
    for (int i = 0; i < 4; ++i) ch[i] = f2u(clampf(p.c[i] * 255.0f + 0.5f, 0.0f, 255.0f));
    return pack4(ch[0], ch[1], ch[2], ch[3]);
}
static Px rgba8_to_unorm(uint32_t p) {
    Px r;
    for (int i = 0; i < 4; ++i) r.c[i] = (float)((p >> (8 * i)) & 255u) * CHANNEL_SCALE;
    return r;
}
static uint32_t mul_div255(uint32_t a, uint32_t b) {
    uint32_t product = a * b + 128u;
    return (product + (product >> 8)) >> 8;
}
static uint32_t combine_alpha(uint32_t a, uint32_t b) { return mul_div255(a, b); }
static float straight_channel(uint32_t premul, uint32_t alpha) { return alpha == 0u ? 0.0f : (float)premul / (float)alpha; }
static uint32_t scale_premul_u8(uint32_t s, uint32_t f) {
    if (f == 0u) return 0u;
    if (f == 255u) return s;
    return pack4(mul_div255(s & 255u, f), mul_div255((s >> 8) & 255u, f), mul_div255((s >> 16) & 255u, f),
                 mul_div255(s >> 24, f));
}
static Px scale_premul_u8_to_unorm(uint32_t s, uint32_t f) {
    Px r = {{0, 0, 0, 0}};
    if (f == 0u || (s >> 24) == 0u) return r;
    float scale = (float)f * (1.0f / 65025.0f);
    for (int i = 0; i < 4; ++i) r.c[i] = (float)((s >> (8 * i)) & 255u) * scale;
    return r;
}
static uint32_t src_over_premul_u8(uint32_t d, uint32_t s) {
    uint32_t alpha = s >> 24;
    if (alpha == 0u) return d;
    if (alpha == 255u) return s;
    uint32_t inv = 255u - alpha;
    return pack4((s & 255u) + mul_div255(d & 255u, inv), ((s >> 8) & 255u) + mul_div255((d >> 8) & 255u, inv),
                 ((s >> 16) & 255u) + mul_div255((d >> 16) & 255u, inv), alpha + mul_div255(d >> 24, inv));
}
static Px src_over_premul_unorm(Px d, Px s) {
    if (s.c[3] <= 0.0f) return d;
    if (s.c[3] >= 1.0f) return s;
    Px r;
    for (int i = 0; i < 4; ++i) r.c[i] = s.c[i] + d.c[i] * (1.0f - s.c[3]);
    return r;
}
static uint32_t pack_premul_rgba8(float r, float g, float b, float a) {
    return pack4(f2u(clampf(r, 0, 1) * 255.0f + 0.5f), f2u(clampf(g, 0, 1) * 255.0f + 0.5f),
                 f2u(clampf(b, 0, 1) * 255.0f + 0.5f), f2u(clampf(a, 0, 1) * 255.0f + 0.5f));
}
static uint32_t lerp_premul_u8(uint32_t a, uint32_t b, float t) {
    uint32_t v[4];
    for (int i = 0; i < 4; ++i) {
        float l = (float)((a >> (8 * i)) & 255u), r = (float)((b >> (8 * i)) & 255u);
        v[i] = f2u(clampf(fmaf(r - l, t, l) + 0.5f, 0.0f, 255.0f));
    }
    return pack4(v[0], v[1], v[2], v[3]);
}
static float rem_euclid_f32(float v, float m) { return fmaf(-floorf(v / m), m, v); }

/* ---------- text compositing of path glyphs ---------------------------------------
 * srgb_to_linear / linear_to_srgb (fine/text/gamma.hlsli:5-21), text_foreground and
 * text_correcting_coverage (fine/text/auto.hlsli:6-26), src_over_subpixel_mask_f32
 * with one coverage for the three channels (fine/text/basic.hlsli:101-115) and
 * src_over_mask_linear_auto_u8 (auto.hlsli:28-33).  lerp(a, b, t) is
 * fmaf(b - a, t, a); dot() sums its products in order. */
static float srgb_to_linear(float value) {
    float v = clampf(value, 0.0f, 1.0f);
    float result = v / 12.92f;
    if (v > 0.04045f) result = powf((v + 0.055f) / 1.055f, 2.4f);
    return result;
}
static float linear_to_srgb(float value) {
    float v = clampf(value, 0.0f, 1.0f);
    float result = v * 12.92f;
    if (v > 0.0031308f) result = 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
    return result;
}
static float text_foreground_luma(uint32_t source) {
    float alpha = (float)(source >> 24);
    if (!(alpha > 0.0f)) return 0.0f;
    float r = clampf((float)(source & 255u) / alpha, 0.0f, 1.0f);
    float g = clampf((float)((source >> 8) & 255u) / alpha, 0.0f, 1.0f);
    float b = clampf((float)((source >> 16) & 255u) / alpha, 0.0f, 1.0f);
    return r * 0.2126f + g * 0.7152f + b * 0.0722f;
}
static float text_correcting_coverage(uint32_t coverage, float foreground) {
    if (coverage == 0u || coverage == 255u) return (float)coverage * (1.0f / 255.0f);
    float raw = (float)coverage * (1.0f / 255.0f);
    float background = 1.0f - foreground;
    float bg_linear = srgb_to_linear(background);
    float enhanced = raw + raw * (1.0f - raw) * bg_linear;
    float difference = foreground - background;
    if (fabsf(difference) < (1.0f / 256.0f)) return enhanced;
    float fg_linear = srgb_to_linear(foreground);
    float output = linear_to_srgb(fmaf(fg_linear - bg_linear, enhanced, bg_linear));
    return clampf((output - background) / difference, 0.0f, 1.0f);
}
static uint32_t src_over_mask_f32(uint32_t destination, uint32_t source, float coverage) {
    if ((source >> 24) == 0u || coverage == 0.0f) return destination;
    float cov = coverage;
    if ((destination >> 24) != 255u) cov = (coverage + coverage + coverage) / 3.0f;
    float alpha = (float)(source >> 24) * (1.0f / 255.0f);
    float amount = alpha * cov;
    uint32_t out = 0u;
    for (int k = 0; k < 3; ++k) {
        float sv = (float)((source >> (8 * k)) & 255u), dv = (float)((destination >> (8 * k)) & 255u);
        out |= f2u(sv * cov + dv * (1.0f - amount) + 0.5f) << (8 * k);
    }
    float out_alpha = amount * 255.0f + (float)(destination >> 24) * (1.0f - amount);
    return out | (f2u(out_alpha + 0.5f) << 24);
}
static uint32_t src_over_mask_linear_auto_u8(uint32_t destination, uint32_t source, uint32_t coverage) {
    if ((source >> 24) == 0u || coverage == 0u) return destination;
    return src_over_mask_f32(destination, source, text_correcting_coverage(coverage, text_foreground_luma(source)));
}

/* Per request {destination, source, coverage, foreground bits}: the composite, the
 * corrected coverage, srgb_to_linear(foreground) and linear_to_srgb(foreground). */
void or_text_words(uint32_t count, const uint32_t *requests, uint32_t *out) {
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t *r = requests + (size_t)i * 4u;
        uint32_t *o = out + (size_t)i * 4u;
        float fg = bits_f(r[3]);
        o[0] = src_over_mask_linear_auto_u8(r[0], r[1], r[2]);
        o[1] = f_bits(text_correcting_coverage(r[2], fg));
        o[2] = f_bits(srgb_to_linear(fg));
        o[3] = f_bits(linear_to_srgb(fg));
    }
}
static int premul_opaque(uint32_t c) { return (c >> 24) == 255u; }

/* ---------- row-sweep coverage (shared/coverage.hlsli:7-51) --------------------- */

typedef struct { float edge, dy, xmin, xmax; } RowParts;
static RowParts segment_row_parts(float p0x, float p0y, float p1x, float p1y, float y_edge, uint32_t y) {
    RowParts r;
    float delta_x = p1x - p0x, delta_y = p1y - p0y;
    float row_y = (float)y, local_y = p0y - row_y;
    float y0 = clampf(local_y, 0.0f, 1.0f), y1 = clampf(p1y - row_y, 0.0f, 1.0f);
    float dy = y0 - y1;
    r.edge = signum_f32(delta_x) * clampf(row_y - y_edge + 1.0f, 0.0f, 1.0f);
    r.dy = dy;
    r.xmin = r.xmax = 0.0f;
    if (dy == 0.0f) return r;
    float slope = delta_x / delta_y;
    int nearer_p0 = fabsf(local_y - 0.5f) <= fabsf(p1y - row_y - 0.5f);
    float ax = nearer_p0 ? p0x : p1x;
    float ay = nearer_p0 ? local_y : p1y - row_y;
    float sx0 = fmaf(y0 - ay, slope, ax), sx1 = fmaf(y1 - ay, slope, ax);
    r.xmin = fminf(sx0, sx1);
    r.xmax = fmaxf(sx0, sx1);
    return r;
}
static float segment_area_at(float xmin_abs, float xmax_abs, uint32_t x) {
    float xmin = xmin_abs - (float)x, xmax = xmax_abs - (float)x;
    float width = xmax - xmin;
    if (width == 0.0f) return clampf(1.0f - xmin, 0.0f, 1.0f);
    if (xmin >= 0.0f && xmax <= 1.0f) return fmaf(-0.5f, xmin + xmax, 1.0f);
    float left = clampf(xmin, 0.0f, 1.0f), right = clampf(xmax, 0.0f, 1.0f);
    float full = clampf(-xmin, 0.0f, width);
    return fmaf(right - left, fmaf(-0.5f, left + right, 1.0f), full) / width;
}
/* fill_alpha_at's FP32 sums taken literally, in stored segment order
 * (shared/coverage.hlsli:34-51).  The reference's stored order comes from emit's
 * atomic slots, so this value can change from run to run in its last bits; it is
 * kept only to bound the difference of the exact form below. */
static float fill_value_sequential(const uint32_t *segs, int32_t backdrop, uint32_t start, uint32_t end, uint32_t x,
                                   uint32_t y) {
    float base = (float)backdrop, running = 0.0f, partial = 0.0f;
    for (uint32_t i = start; i < end; ++i) {
        const uint32_t *s = segs + (size_t)i * 5u;
        RowParts p = segment_row_parts(bits_f(s[0]), bits_f(s[1]), bits_f(s[2]), bits_f(s[3]), bits_f(s[4]), y);
        base += p.edge;
        if (p.dy != 0.0f) {
            int full_start = clampi(f2i(ceilf(p.xmax)), 0, (int)TILE);
            if (full_start < (int)TILE && (int)x >= full_start) running += p.dy;
            int ps = clampi(f2i(floorf(p.xmin)), 0, (int)TILE), pe = clampi(f2i(ceilf(p.xmax)), 0, (int)TILE);
            if ((int)x >= ps && (int)x < pe) partial += segment_area_at(p.xmin, p.xmax, x) * p.dy;
        }
    }
    return base + running + partial;
}

/* The same three sums, exact: every FP32 term (in (-2, 2) by construction) is
 * truncated to a multiple of 2^-30 (trunc(t * 2^30) as a saturating int32
 * conversion, NaN -> 0) and summed as a 64-bit integer; the three sums are
 * added exactly and the total converted back once (floored to 2^-24, rounded
 * once), where fill_alpha_at rounds each sum and then (base + running) +
 * partial.  No segment order enters the result. */
static int64_t coverage_fixed(float t) { return (int64_t)f2i(t * 1073741824.0f); }
/* Back to FP32: the sum floored to a multiple of 2^-24, then one rounding
 * (floor(v / 2^30) and the top 24 fraction bits convert exactly, the 2^-24
 * scaling is exact, the add rounds). */
static float fixed_to_float(int64_t v) {
    const float hi = (float)(int32_t)(v >> 30);
    const float lo = (float)(((uint32_t)v & 0x3FFFFFFFu) >> 6) * 5.9604644775390625e-8f;
    return hi + lo;
}

static float fill_value_at(const uint32_t *segs, int32_t backdrop, uint32_t start, uint32_t end, uint32_t x,
                           uint32_t y) {
    int64_t base = (int64_t)backdrop * 1073741824LL, running = 0, partial = 0;
    for (uint32_t i = start; i < end; ++i) {
        const uint32_t *s = segs + (size_t)i * 5u;
        RowParts p = segment_row_parts(bits_f(s[0]), bits_f(s[1]), bits_f(s[2]), bits_f(s[3]), bits_f(s[4]), y);
        base += coverage_fixed(p.edge);
        if (p.dy != 0.0f) {
            int full_start = clampi(f2i(ceilf(p.xmax)), 0, (int)TILE);
            if (full_start < (int)TILE && (int)x >= full_start) running += coverage_fixed(p.dy);
            int ps = clampi(f2i(floorf(p.xmin)), 0, (int)TILE), pe = clampi(f2i(ceilf(p.xmax)), 0, (int)TILE);
            if ((int)x >= ps && (int)x < pe) partial += coverage_fixed(segment_area_at(p.xmin, p.xmax, x) * p.dy);
        }
    }
    return fixed_to_float(base + running + partial);
}

/* ---------- SDF + brush (shared/sdf/ and shared/brush/ headers) -------------------- */

static uint32_t blob_word(const uint32_t *paint, uint32_t count, uint32_t base, uint32_t index) {
    if (base < count && index < count - base) return paint[base + index];
    return 0u;
}
static float blob_float(const uint32_t *paint, uint32_t count, uint32_t base, uint32_t index) {
    return bits_f(blob_word(paint, count, base, index));
}
typedef struct { float a, b, c, d, e, f; } Aff;
typedef struct { float distance, nx, ny; } Sdf;
static float vlen(float x, float y) { return sqrtf(x * x + y * y); }
static float sdf_length(float x, float y) {
    float m = vlen(x, y);
    if (m > 0.0f && isfinite(m)) return m;
    float axis = fmaxf(fabsf(x), fabsf(y));
    return axis == 0.0f ? 0.0f : axis * vlen(x / axis, y / axis);
}
static Sdf rounded_box(float px, float py, float hx, float hy, float radius) {
    float qx = fabsf(px) - hx + radius, qy = fabsf(py) - hy + radius;
    float ox = fmaxf(qx, 0.0f), oy = fmaxf(qy, 0.0f);
    float ol = sdf_length(ox, oy);
    Sdf s;
    s.distance = fminf(fmaxf(qx, qy), 0.0f) + ol - radius;
    float sx = px >= 0.0f ? 1.0f : -1.0f, sy = py >= 0.0f ? 1.0f : -1.0f;
    if (ol > 0.0f) { s.nx = ox / ol * sx; s.ny = oy / ol * sy; }
    else if (qx > qy) { s.nx = sx; s.ny = 0.0f; }
    else { s.nx = 0.0f; s.ny = sy; }
    return s;
}
static Sdf rect_sdf(float x, float y, float x0r, float y0r, float x1r, float y1r, float tl, float tr, float bl, float br) {
    float x0 = fminf(x0r, x1r), y0 = fminf(y0r, y1r), x1 = fmaxf(x0r, x1r), y1 = fmaxf(y0r, y1r);
    float cx = (x0 + x1) * 0.5f, cy = (y0 + y1) * 0.5f, hx = (x1 - x0) * 0.5f, hy = (y1 - y0) * 0.5f;
    float px = x - cx, py = y - cy;
    float radius = tl;
    if (px >= 0.0f) radius = py <= 0.0f ? tr : br;
    else if (py > 0.0f) radius = bl;
    float r = fmaxf(fminf(fminf(radius, hx), hy), 0.0f);
    return rounded_box(px, py, hx, hy, r);
}
static float sdf_cov(Sdf s, const Aff *inv) {
    float gx = fmaf(inv->a, s.nx, inv->b * s.ny), gy = fmaf(inv->c, s.nx, inv->d * s.ny);
    float fp = sdf_length(gx, gy);
    if (fp == 0.0f) return s.distance < 0.0f ? 1.0f : (s.distance > 0.0f ? 0.0f : 0.5f);
    return clampf(0.5f - s.distance / fp, 0.0f, 1.0f);
}
static float sdf_coverage_from_blob(const uint32_t *paint, uint32_t n, uint32_t base, float x, float y, const Aff *inv) {
    uint32_t kind = blob_word(paint, n, base, 0);
    float v[17];
    for (uint32_t i = 1; i < 17; ++i) v[i] = blob_float(paint, n, base, i);
    if (kind == 1u) return sdf_cov(rect_sdf(x, y, v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]), inv);
    if (kind == 3u) {
        float ht = fmaxf(v[9], 0.0f), hr = fmaxf(v[10], 0.0f), hb = fmaxf(v[11], 0.0f), hl = fmaxf(v[12], 0.0f);
        float rx0 = fminf(v[1], v[3]), ry0 = fminf(v[2], v[4]), rx1 = fmaxf(v[1], v[3]), ry1 = fmaxf(v[2], v[4]);
        float outer = sdf_cov(rect_sdf(x, y, rx0 - hl, ry0 - ht, rx1 + hr, ry1 + hb, v[5] + fmaxf(ht, hl),
                                       v[6] + fmaxf(ht, hr), v[7] + fmaxf(hb, hl), v[8] + fmaxf(hb, hr)), inv);
        float ix0 = rx0 + hl, iy0 = ry0 + ht, ix1 = rx1 - hr, iy1 = ry1 - hb, inner = 0.0f;
        if (ix0 < ix1 && iy0 < iy1)
            inner = sdf_cov(rect_sdf(x, y, ix0, iy0, ix1, iy1, fmaxf(v[5] - fmaxf(ht, hl), 0.0f),
                                     fmaxf(v[6] - fmaxf(ht, hr), 0.0f), fmaxf(v[7] - fmaxf(hb, hl), 0.0f),
                                     fmaxf(v[8] - fmaxf(hb, hr), 0.0f)), inv);
        return clampf(outer - inner, 0.0f, 1.0f);
    }
    return 0.0f;
}
static float apply_extend(float t, uint32_t extend) {
    if (extend == 1u) return rem_euclid_f32(t, 1.0f);
    if (extend == 2u) {
        float v = rem_euclid_f32(t, 2.0f);
        return v <= 1.0f ? v : 2.0f - v;
    }
    return clampf(t, 0.0f, 1.0f);
}
static uint32_t sample_ramp(const uint32_t *paint, uint32_t n, uint32_t bb, uint32_t payload, uint32_t len, float t,
                            uint32_t extend) {
    if (len == 0u) return 0u;
    uint32_t last = len - 1u;
    float e = apply_extend(t, extend);
    float position = e * (float)last;
    uint32_t li = f2u(floorf(position)), ri = umin(li + 1u, last);
    float fraction = fmaf(e, (float)last, -(float)li);
    uint32_t left = blob_word(paint, n, bb, payload + li), right = blob_word(paint, n, bb, payload + ri);
    if (fraction <= 0.00000011920929f || li == ri) return left;
    return lerp_premul_u8(left, right, fraction);
}
/* Image brushes: sample_resource_pattern_atlas and sample_atlas_brush
 * (shared/brush/pattern.hlsli:9-55) over the scene's atlas, with
 * euclidean_remainder_i32 (shared/integer.hlsli:5-9) and
 * pattern_transform_component (shared/pattern_transform.hlsli:7-11).  The padded
 * bilinear case is the reference's hardware sampler (SampleLevel): position
 * clamped to the image, neighbours clamped to it, the position resolved to 8
 * fractional bits (weights k / 256, the subtexel precision Direct3D requires) and
 * each 0-255 channel blended in integers and rounded once. */
typedef struct { const uint32_t *texels; uint32_t width, height, pages; } Atlas;
static uint32_t euclid_rem_i32(int32_t value, uint32_t period) {
    uint32_t magnitude = value < 0 ? 0u - (uint32_t)value : (uint32_t)value;
    uint32_t remainder = magnitude % period;
    return value < 0 && remainder != 0u ? period - remainder : remainder;
}
static uint32_t extend_coord(int32_t value, uint32_t size, uint32_t extend) {
    if (size <= 1u) return 0u;
    if (extend == 1u || extend == 2u) {
        uint32_t period = size * (extend == 2u ? 2u : 1u);
        uint32_t coord = euclid_rem_i32(value, period);
        return coord < size ? coord : period - coord - 1u;
    }
    int32_t hi = (int32_t)size - 1;
    return (uint32_t)(value < 0 ? 0 : (value > hi ? hi : value));
}
static uint32_t atlas_texel(const Atlas *a, uint32_t page, uint32_t x, uint32_t y) {
    if (!a || !a->texels || page >= a->pages || x >= a->width || y >= a->height) return 0u;
    return a->texels[((size_t)page * a->height + y) * a->width + x];
}
static uint32_t atlas_pixel(const Atlas *a, uint32_t page, uint32_t ox, uint32_t oy, uint32_t sx, uint32_t sy,
                            uint32_t extend, int32_t px, int32_t py) {
    return atlas_texel(a, page, ox + extend_coord(px, sx, extend), oy + extend_coord(py, sy, extend));
}
static float pattern_component(float a, float b, float offset, float x, float y) {
    float by = b * y;
    float dot = fmaf(a, x, by) + fmaf(b, y, -by);
    return dot + offset;
}
/* The texture unit's 8-bit subtexel precision: position q split into a texel
 * index and a weight k / 256, the fraction rounded to the nearest 1/256. */
static void subtexel_position(float q, int32_t *index, uint32_t *weight) {
    float base = floorf(q);
    *weight = (uint32_t)rintf((q - base) * 256.0f);
    *index = f2i(base);
    if (*weight == 256u) {
        *index += 1;
        *weight = 0u;
    }
}
static uint32_t sample_image(const uint32_t *paint, uint32_t n, uint32_t bb, uint32_t db, const Atlas *atlas, float x,
                             float y) {
    uint32_t sx = blob_word(paint, n, bb, db + 5u), sy = blob_word(paint, n, bb, db + 6u);
    if (sx == 0u || sy == 0u) return 0u;
    uint32_t page = blob_word(paint, n, bb, db + 4u);
    if (page & 0x80000000u) return 0u; /* a texture-table image: only the atlas is modelled */
    uint32_t ox = blob_word(paint, n, bb, db + 2u), oy = blob_word(paint, n, bb, db + 3u);
    uint32_t opacity = blob_word(paint, n, bb, db + 7u), extend = blob_word(paint, n, bb, db + 1u);
    uint32_t sampling = blob_word(paint, n, bb, db + 8u);
    float p[6];
    for (uint32_t i = 0; i < 6u; ++i) p[i] = blob_float(paint, n, bb, db + 9u + i);
    float px = pattern_component(p[0], p[2], p[4], x, y) * (float)sx;
    float py = pattern_component(p[1], p[3], p[5], x, y) * (float)sy;
    uint32_t color;
    if (sampling == 1u && extend == 0u) {
        int32_t ix, iy;
        uint32_t wx, wy;
        subtexel_position(clampf(px, 0.0f, (float)sx) - 0.5f, &ix, &wx);
        subtexel_position(clampf(py, 0.0f, (float)sy) - 0.5f, &iy, &wy);
        uint32_t t[4] = {atlas_pixel(atlas, page, ox, oy, sx, sy, 0u, ix, iy),
                         atlas_pixel(atlas, page, ox, oy, sx, sy, 0u, ix + 1, iy),
                         atlas_pixel(atlas, page, ox, oy, sx, sy, 0u, ix, iy + 1),
                         atlas_pixel(atlas, page, ox, oy, sx, sy, 0u, ix + 1, iy + 1)};
        uint32_t w[4] = {(256u - wx) * (256u - wy), wx * (256u - wy), (256u - wx) * wy, wx * wy};
        color = 0u;
        for (int k = 0; k < 4; ++k) {
            uint32_t sum = 32768u;
            for (int n = 0; n < 4; ++n) sum += w[n] * ((t[n] >> (8 * k)) & 255u);
            color |= (sum >> 16) << (8 * k);
        }
    } else if (sampling == 1u) {
        float qx = px - 0.5f, qy = py - 0.5f;
        float bx = floorf(qx), by = floorf(qy), fx = qx - bx, fy = qy - by;
        int32_t ix = f2i(bx), iy = f2i(by);
        uint32_t tl = atlas_pixel(atlas, page, ox, oy, sx, sy, extend, ix, iy);
        uint32_t tr = atlas_pixel(atlas, page, ox, oy, sx, sy, extend, ix + 1, iy);
        uint32_t bl = atlas_pixel(atlas, page, ox, oy, sx, sy, extend, ix, iy + 1);
        uint32_t br = atlas_pixel(atlas, page, ox, oy, sx, sy, extend, ix + 1, iy + 1);
        color = lerp_premul_u8(lerp_premul_u8(tl, tr, fx), lerp_premul_u8(bl, br, fx), fy);
    } else {
        color = atlas_pixel(atlas, page, ox, oy, sx, sy, extend, f2i(floorf(px)), f2i(floorf(py)));
    }
    return scale_premul_u8(color, opacity);
}

static uint32_t sample_brush_atlas(const uint32_t *paint, uint32_t n, uint32_t bb, uint32_t db, float x, float y,
                                   const Atlas *atlas);
static uint32_t sample_brush(const uint32_t *paint, uint32_t n, uint32_t bb, uint32_t db, float x, float y) {
    return sample_brush_atlas(paint, n, bb, db, x, y, NULL);
}
static uint32_t sample_brush_atlas(const uint32_t *paint, uint32_t n, uint32_t bb, uint32_t db, float x, float y,
                                   const Atlas *atlas) {
    uint32_t kind = blob_word(paint, n, bb, db);
    uint32_t extend = blob_word(paint, n, bb, db + 1u);
    uint32_t payload = db + blob_word(paint, n, bb, db + 2u);
    uint32_t len = blob_word(paint, n, bb, db + 3u);
    uint32_t base = db + 9u;
    uint32_t color = blob_word(paint, n, bb, db + 4u);
    if (kind == 2u) {
        float p[10];
        for (uint32_t i = 0; i < 10; ++i) p[i] = blob_float(paint, n, bb, base + i);
        float tx = fmaf(p[4], x, fmaf(p[6], y, p[8])), ty = fmaf(p[5], x, fmaf(p[7], y, p[9]));
        float dx = p[2] - p[0], dy = p[3] - p[1];
        float den = dx * dx + dy * dy, t = 0.0f;
        if (den > 0.00000011920929f) t = fmaf(tx - p[0], dx, (ty - p[1]) * dy) / den;
        return sample_ramp(paint, n, bb, payload, len, t, extend);
    }
    if (kind == 3u) { /* sample_radial (shared/brush/radial.hlsli:5-33) */
        float p[12];
        for (uint32_t i = 0; i < 12; ++i) p[i] = blob_float(paint, n, bb, base + i);
        float tx = fmaf(p[6], x, fmaf(p[8], y, p[10])), ty = fmaf(p[7], x, fmaf(p[9], y, p[11]));
        float r0 = p[4], qx = tx - p[0], qy = ty - p[1], dcx = p[2] - p[0], dcy = p[3] - p[1], dr = p[5] - r0;
        float a = fmaf(dcx, dcx, fmaf(dcy, dcy, -(dr * dr)));
        float b = -2.0f * fmaf(qx, dcx, fmaf(qy, dcy, r0 * dr));
        float c = fmaf(qx, qx, fmaf(qy, qy, -(r0 * r0)));
        int has_t = 0;
        float t = 0.0f;
        if (fabsf(a) <= 0.000001f) {
            if (fabsf(b) > 0.000001f) {
                float candidate = -c / b;
                if (fmaf(candidate, dr, r0) >= 0.0f) { has_t = 1; t = candidate; }
            }
        } else {
            float disc = fmaf(b, b, (-4.0f * a) * c);
            if (disc >= 0.0f) {
                float root = sqrtf(disc);
                float t0 = (-b - root) / (2.0f * a), t1 = (-b + root) / (2.0f * a);
                int v0 = fmaf(t0, dr, r0) >= 0.0f, v1 = fmaf(t1, dr, r0) >= 0.0f;
                if (v0) { has_t = 1; t = v1 ? fmaxf(t0, t1) : t0; }
                else if (v1) { has_t = 1; t = t1; }
            }
        }
        return has_t ? sample_ramp(paint, n, bb, payload, len, t, extend) : 0u;
    }
    if (kind == 4u) { /* sample_sweep (shared/brush/sweep.hlsli:5-20) */
        float cx = blob_float(paint, n, bb, base), cy = blob_float(paint, n, bb, base + 1u);
        float start = blob_float(paint, n, bb, base + 2u), span = blob_float(paint, n, bb, base + 3u), t = 0.0f;
        if (fabsf(span) > 0.00000011920929f) {
            float angle = 0.0f;
            if (x != cx || y != cy) angle = atan2f(y - cy, x - cx);
            float distance = angle - start;
            if (span > 0.0f) { if (distance < 0.0f) distance += 6.2831855f; }
            else if (distance > 0.0f) distance -= 6.2831855f;
            t = distance / span;
        }
        return sample_ramp(paint, n, bb, payload, len, t, extend);
    }
    if (kind == 5u) { /* sample_four_corner (shared/brush/four_corner.hlsli:5-15) */
        float x0 = blob_float(paint, n, bb, base), y0 = blob_float(paint, n, bb, base + 1u);
        float x1 = blob_float(paint, n, bb, base + 2u), y1 = blob_float(paint, n, bb, base + 3u);
        float w = x1 - x0, h = y1 - y0, u = 0.0f, v = 0.0f;
        if (fabsf(w) > 0.00000011920929f) u = clampf((x - x0) / w, 0.0f, 1.0f);
        if (fabsf(h) > 0.00000011920929f) v = clampf((y - y0) / h, 0.0f, 1.0f);
        uint32_t tl = blob_word(paint, n, bb, payload), tr = blob_word(paint, n, bb, payload + 1u);
        uint32_t br = blob_word(paint, n, bb, payload + 2u), bl = blob_word(paint, n, bb, payload + 3u);
        return lerp_premul_u8(lerp_premul_u8(tl, tr, u), lerp_premul_u8(bl, br, u), v);
    }
    if (kind == 6u) return 0u;
    if (kind == 7u) return sample_image(paint, n, bb, db, atlas, x, y);
    return color;
}

/* brush_words (validation/brush.hlsl:14-18). */
void or_brush_words(const uint32_t *paint, uint32_t paint_words, uint32_t paint_brush_base, uint32_t count,
                    const uint32_t *requests, uint32_t *out) {
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t *q = requests + (size_t)i * 4u;
        out[i] = sample_brush(paint, paint_words, paint_brush_base, q[0], bits_f(q[1]), bits_f(q[2]));
    }
}

/* ---------- coarse (src/shaders/hlsl/coarse/, shared/draw.hlsli) ------------------ */

typedef struct {
    uint32_t path_id, glyph_run_id, sdf_offset, sdf_len, shadow_offset, tag, fill_rule, brush_offset, solid_rect;
    int32_t pb[4];
    float lin[4], tr[2];
} Draw;

static Draw load_draw(const uint32_t *draws, uint32_t count, uint32_t index) {
    Draw d;
    memset(&d, 0, sizeof(d));
    if (index >= count) return d;
    const uint32_t *r = draws + (size_t)index * 31u;
    d.path_id = r[0]; d.glyph_run_id = r[1]; d.sdf_offset = r[2]; d.sdf_len = r[3]; d.shadow_offset = r[4];
    d.brush_offset = r[6]; d.tag = r[8]; d.fill_rule = r[9]; d.solid_rect = r[18];
    for (int i = 0; i < 4; ++i) d.pb[i] = (int32_t)r[10 + i];
    for (int i = 0; i < 4; ++i) d.lin[i] = bits_f(r[19 + i]);
    d.tr[0] = bits_f(r[23]);
    d.tr[1] = bits_f(r[24]);
    return d;
}
static int has_sdf(const Draw *d) { return d->sdf_offset != NONE || d->shadow_offset != NONE; }
static uint32_t tmin_of(int32_t p, uint32_t lim) { return p > 0 ? umin((uint32_t)p / TILE, lim) : 0u; }
static uint32_t tmax_of(int32_t p, uint32_t lim) { return p > 0 ? umin(((uint32_t)p + TILE - 1u) / TILE, lim) : 0u; }
static int draw_hits_tile(const Draw *d, uint32_t tx, uint32_t ty, uint32_t tw, uint32_t th) {
    return tx >= tmin_of(d->pb[0], tw) && ty >= tmin_of(d->pb[1], th) && tx < tmax_of(d->pb[2], tw) &&
           ty < tmax_of(d->pb[3], th);
}
static uint32_t draw_backdrop(const or_coarse_config *cfg, const uint32_t *paths, const Draw *d, uint32_t tx,
                              uint32_t ty) {
    if (d->path_id == NONE || d->tag > 5u || d->path_id >= cfg->path_count ||
        !draw_hits_tile(d, tx, ty, cfg->tiles_width, cfg->tiles_height))
        return NONE;
    const uint32_t *p = paths + (size_t)d->path_id * 19u;
    if (p[8] <= p[6] || tx < p[6] || ty < p[7] || tx >= p[8] || ty >= p[9]) return NONE;
    return p[4] + (ty - p[7]) * (p[8] - p[6]) + tx - p[6];
}
static uint32_t pw(const or_coarse_config *cfg, const or_coarse_inputs *in, uint32_t i) {
    return i < cfg->paint_words ? in->paint[i] : 0u;
}
static int corner_covers(float px, float py, float cx, float cy, float ox, float oy, float r, float inset) {
    int in_square = r > 0.0f && fabsf(px - cx) < r && fabsf(py - cy) < r;
    float inner = r - inset, dx = px - ox, dy = py - oy;
    return !in_square || (inner > 0.0f && dx * dx + dy * dy <= inner * inner);
}
/* rect_covers_tile (coarse/sdf_clip.hlsli:15-32). */
static int rect_covers_tile(const float *rect, const float *radii_in, const Draw *d, uint32_t tx, uint32_t ty,
                            float inset) {
    float lx = fminf(rect[0], rect[2]), ly = fminf(rect[1], rect[3]);
    float ux = fmaxf(rect[0], rect[2]), uy = fmaxf(rect[1], rect[3]);
    float tlx = (float)(tx * TILE) + 0.5f - d->tr[0], tly = (float)(ty * TILE) + 0.5f - d->tr[1];
    float tux = tlx + (float)(TILE - 1u), tuy = tly + (float)(TILE - 1u);
    float sx = ux - lx, sy = uy - ly, need = (float)(TILE - 1u) + 2.0f * inset;
    if (sx < need || sy < need || tlx < lx + inset || tly < ly + inset || tux > ux - inset || tuy > uy - inset)
        return 0;
    float limit = fminf(sx, sy) * 0.5f;
    for (int i = 0; i < 4; ++i)
        if (!(fabsf(rect[i]) <= 65536.0f) || !(fabsf(radii_in[i]) <= 65536.0f)) return 0;
    if (!(fabsf(d->tr[0]) <= 65536.0f) || !(fabsf(d->tr[1]) <= 65536.0f)) return 0;
    float r[4];
    for (int i = 0; i < 4; ++i) r[i] = fminf(fmaxf(radii_in[i], 0.0f), limit);
    return corner_covers(tlx, tly, lx, ly, lx + r[0], ly + r[0], r[0], inset) &&
           corner_covers(tux, tly, ux, ly, ux - r[1], ly + r[1], r[1], inset) &&
           corner_covers(tlx, tuy, lx, uy, lx + r[2], uy - r[2], r[2], inset) &&
           corner_covers(tux, tuy, ux, uy, ux - r[3], uy - r[3], r[3], inset);
}
static int identity_linear(const Draw *d) {
    return d->lin[0] == 1.0f && d->lin[1] == 0.0f && d->lin[2] == 0.0f && d->lin[3] == 1.0f;
}
/* sdf_clip_covers (coarse/sdf_clip.hlsli:34-39). */
static int sdf_clip_covers(const or_coarse_config *cfg, const or_coarse_inputs *in, const Draw *d, uint32_t tx,
                           uint32_t ty) {
    if (d->sdf_offset == NONE || d->shadow_offset != NONE || d->sdf_len < 9u) return 0;
    uint32_t base = d->sdf_offset;
    if (pw(cfg, in, base) != 1u || !identity_linear(d)) return 0;
    float rect[4], radii[4];
    for (int i = 0; i < 4; ++i) {
        rect[i] = bits_f(pw(cfg, in, base + 1u + (uint32_t)i));
        radii[i] = bits_f(pw(cfg, in, base + 5u + (uint32_t)i));
    }
    return rect_covers_tile(rect, radii, d, tx, ty, 0.5f);
}
/* sdf_stroke_empty (coarse/sdf_clip.hlsli:40-59). */
static int sdf_stroke_empty(const or_coarse_config *cfg, const or_coarse_inputs *in, const Draw *d, uint32_t tx,
                            uint32_t ty) {
    if (d->sdf_offset == NONE || d->shadow_offset != NONE || d->sdf_len < 13u || !identity_linear(d)) return 0;
    uint32_t base = d->sdf_offset;
    if (pw(cfg, in, base) != 3u) return 0;
    float rect[4], radii[4], stroke[4];
    for (int i = 0; i < 4; ++i) {
        rect[i] = bits_f(pw(cfg, in, base + 1u + (uint32_t)i));
        radii[i] = bits_f(pw(cfg, in, base + 5u + (uint32_t)i));
        stroke[i] = bits_f(pw(cfg, in, base + 9u + (uint32_t)i));
    }
    for (int i = 0; i < 4; ++i)
        if (!(fabsf(rect[i]) <= 65536.0f) || !(fabsf(radii[i]) <= 65536.0f) || !(fabsf(stroke[i]) <= 65536.0f))
            return 0;
    float hw[4];
    for (int i = 0; i < 4; ++i) hw[i] = fmaxf(stroke[i], 0.0f);
    float lx = fminf(rect[0], rect[2]), ly = fminf(rect[1], rect[3]), ux = fmaxf(rect[0], rect[2]),
          uy = fmaxf(rect[1], rect[3]);
    float inner[4] = {lx + hw[3], ly + hw[0], ux - hw[1], uy - hw[2]};
    if (inner[0] >= inner[2] || inner[1] >= inner[3]) return 0;
    float corner[4] = {fmaxf(hw[0], hw[3]), fmaxf(hw[0], hw[1]), fmaxf(hw[2], hw[3]), fmaxf(hw[2], hw[1])};
    float r[4];
    for (int i = 0; i < 4; ++i) r[i] = fmaxf(radii[i] - corner[i], 0.0f);
    return rect_covers_tile(inner, r, d, tx, ty, 1.0f);
}
static int nontransparent_solid(const or_coarse_config *cfg, const or_coarse_inputs *in, const Draw *d) {
    if (d->brush_offset == NONE) return 0;
    uint32_t b = cfg->paint_brush_base + d->brush_offset;
    return pw(cfg, in, b) == 1u && pw(cfg, in, b + 4u) != 0u;
}
static uint32_t solid_color(const or_coarse_config *cfg, const or_coarse_inputs *in, const Draw *d) {
    return pw(cfg, in, cfg->paint_brush_base + d->brush_offset + 4u);
}

/* stack_wrapper_count (coarse/classify.hlsli:11-33). */
static uint32_t stack_wrapper_count(const or_coarse_config *cfg, const or_coarse_inputs *in, uint32_t tx, uint32_t ty) {
    uint32_t count = 0;
    for (uint32_t i = cfg->layer_stack_start; i < cfg->layer_stack_end; ++i) {
        const uint32_t *layer = in->layers + (size_t)i * 3u;
        Draw d = load_draw(in->draws, cfg->draw_count, layer[1]);
        if (has_sdf(&d)) {
            if (sdf_clip_covers(cfg, in, &d, tx, ty)) continue;
            if (!draw_hits_tile(&d, tx, ty, cfg->tiles_width, cfg->tiles_height)) return NONE;
        } else {
            uint32_t bd = draw_backdrop(cfg, in->paths, &d, tx, ty);
            if (bd == NONE) return NONE;
            uint32_t r0 = in->segment_ranges[2 * (size_t)bd], r1 = in->segment_ranges[2 * (size_t)bd + 1];
            int32_t w = in->backdrops[bd];
            int full = d.fill_rule == 1u ? (((uint32_t)w & 1u) != 0u) : w != 0;
            if (layer[0] == 0u && r0 == r1 && full) continue;
            if (r0 == r1 && w == 0) return NONE;
        }
        count++;
    }
    return count;
}

typedef struct { int valid; uint32_t tag, winding, fill_rule, s0, s1, color; } Particle;

/* draw_particle (coarse/emit_draw.hlsli:10-52); draw_particle_count (classify.hlsli:34-49) is its validity. */
static Particle draw_particle(const or_coarse_config *cfg, const or_coarse_inputs *in, uint32_t di, uint32_t tx,
                              uint32_t ty) {
    Particle p = {0, 1u, 0, 0, 0, 0, 0};
    if (di >= cfg->draw_count || in->draw_batch_ids[di] != cfg->batch_id) return p;
    Draw d = load_draw(in->draws, cfg->draw_count, di);
    if (has_sdf(&d)) {
        if (d.tag == 0u && !sdf_stroke_empty(cfg, in, &d, tx, ty)) {
            p.valid = 1;
            uint32_t color = nontransparent_solid(cfg, in, &d) && sdf_clip_covers(cfg, in, &d, tx, ty)
                                 ? solid_color(cfg, in, &d) : 0u;
            uint32_t bb = cfg->paint_brush_base + d.brush_offset;
            int image = d.brush_offset != NONE && pw(cfg, in, bb) == 7u && pw(cfg, in, bb + 7u) == 255u &&
                        sdf_clip_covers(cfg, in, &d, tx, ty);
            if (color != 0u) { p.tag = 2u; p.color = color; }
            else if (image) { p.tag = 13u; p.color = di; }
            else { p.tag = 9u; p.s0 = di; p.color = di; }
        }
        return p;
    }
    uint32_t bd = draw_backdrop(cfg, in->paths, &d, tx, ty);
    if (bd == NONE || (d.tag != 0u && d.tag != 5u && d.tag != 1u)) return p;
    uint32_t r0 = in->segment_ranges[2 * (size_t)bd], r1 = in->segment_ranges[2 * (size_t)bd + 1];
    uint32_t w = (uint32_t)in->backdrops[bd];
    if (r0 == r1 && w == 0u) return p;
    p.valid = 1;
    p.winding = w;
    p.fill_rule = d.fill_rule;
    p.s0 = r0;
    p.s1 = r1;
    if (d.tag == 1u) p.tag = 3u;
    else if (d.tag == 5u) p.tag = 11u;
    else if (d.solid_rect != 0u && d.lin[1] == 0.0f && d.lin[2] == 0.0f && nontransparent_solid(cfg, in, &d) && r0 == r1)
        p.tag = 2u;
    if (p.tag == 2u) p.color = solid_color(cfg, in, &d);
    else if (d.tag == 0u || d.tag == 5u) p.color = di;
    return p;
}

/* coarse_count_bins (coarse/count.hlsl:48-73) + store_tile_counts (coarse/draw_list.hlsli:19-24). */
typedef struct {
    const or_coarse_config *cfg;
    const or_coarse_inputs *in;
    uint32_t *work, *kinds;
} CoarseJob;
static void coarse_count_range(void *ctx, uint32_t t0, uint32_t t1);
static void coarse_emit_range(void *ctx, uint32_t t0, uint32_t t1);

void or_coarse_count_bins(const or_coarse_config *cfg, const or_coarse_inputs *in, uint32_t *work, uint32_t *kinds) {
    CoarseJob j = {cfg, in, work, kinds};
    par_for(cfg->tile_count, 64u, coarse_count_range, &j);
}
static void coarse_count_range(void *ctx, uint32_t t0, uint32_t t1) {
    const CoarseJob *j = (const CoarseJob *)ctx;
    const or_coarse_config *cfg = j->cfg;
    const or_coarse_inputs *in = j->in;
    uint32_t *work = j->work, *kinds = j->kinds;
    for (uint32_t tile = t0; tile < t1; ++tile) {
        uint32_t tx = tile % cfg->tiles_width, ty = tile / cfg->tiles_width;
        uint32_t w = stack_wrapper_count(cfg, in, tx, ty);
        uint32_t count = 0;
        if (w != NONE) {
            uint32_t start = in->tile_draws[2 * (size_t)tile], remaining = in->tile_draws[2 * (size_t)tile + 1];
            for (uint32_t i = 0; start != NONE && i < remaining; ++i) {
                uint32_t di = start + i < cfg->tile_draw_index_count ? in->tile_draw_indices[start + i] : 0u;
                count += draw_particle(cfg, in, di, tx, ty).valid ? 1u : 0u;
            }
        }
        if (count > 0u) count += w * 2u + 1u;
        work[(size_t)tile * 6u] = count;
        work[(size_t)tile * 6u + 3u] = 0u;
        kinds[tile] = count > 0u ? 0u : 1u;
    }
}

/* coarse_prefix_chunks / coarse_chunk_offsets / coarse_apply_chunk_offsets (coarse/prefix.hlsl:11-63). */
void or_coarse_prefix(uint32_t tile_count, uint32_t *work, uint32_t *chunk_records) {
    uint32_t chunks = (tile_count + 255u) / 256u;
    for (uint32_t c = 0; c < chunks; ++c) {
        uint32_t run[2] = {0u, 0u};
        for (uint32_t lane = 0; lane < 256u; ++lane) {
            uint32_t item = c * 256u + lane;
            if (item >= tile_count) break;
            uint32_t *rec = work + (size_t)item * 6u;
            for (int k = 0; k < 2; ++k) {
                uint32_t v = rec[3 * k];
                rec[3 * k + 1] = run[k];
                rec[3 * k + 2] = run[k] + v;
                run[k] += v;
            }
        }
        chunk_records[4 * c] = run[0];
        chunk_records[4 * c + 2] = run[1];
    }
    uint32_t carry[2] = {0u, 0u};
    for (uint32_t c = 0; c < chunks; ++c) {
        chunk_records[4 * c + 1] = carry[0];
        chunk_records[4 * c + 3] = carry[1];
        carry[0] += chunk_records[4 * c];
        carry[1] += chunk_records[4 * c + 2];
    }
    for (uint32_t item = 0; item < tile_count; ++item) {
        uint32_t c = item / 256u;
        uint32_t *rec = work + (size_t)item * 6u;
        rec[1] += chunk_records[4 * c + 1];
        rec[2] += chunk_records[4 * c + 1];
        rec[4] += chunk_records[4 * c + 3];
        rec[5] += chunk_records[4 * c + 3];
    }
}

static void store_particle(const or_coarse_config *cfg, uint32_t *work, uint32_t dst, uint32_t tag, uint32_t w,
                           uint32_t rule, uint32_t s0, uint32_t s1, uint32_t color) {
    if (dst >= cfg->ptcl_capacity) return;
    uint32_t *r = work + (size_t)cfg->tile_count * 6u + (size_t)dst * 6u;
    r[0] = tag; r[1] = w; r[2] = rule; r[3] = s0; r[4] = s1; r[5] = color;
}
static int noop_stack_clip(const or_coarse_config *cfg, const or_coarse_inputs *in, const uint32_t *layer,
                           const Draw *d, uint32_t tx, uint32_t ty) {
    if (layer[0] != 0u) return 0;
    if (has_sdf(d)) return sdf_clip_covers(cfg, in, d, tx, ty);
    uint32_t bd = draw_backdrop(cfg, in->paths, d, tx, ty);
    uint32_t r0 = in->segment_ranges[2 * (size_t)bd], r1 = in->segment_ranges[2 * (size_t)bd + 1];
    uint32_t w = (uint32_t)in->backdrops[bd];
    return r0 == r1 && (d->fill_rule == 1u ? (w & 1u) != 0u : w != 0u);
}
/* occluded_particle_prefix (coarse/occlusion.hlsli:12-34). */
static uint32_t occluded_prefix(const or_coarse_config *cfg, const uint32_t *work, uint32_t tile, uint32_t end,
                                uint32_t wrappers) {
    const uint32_t *rec = work + (size_t)tile * 6u;
    uint32_t start = rec[1];
    if (wrappers != 0u || end <= start || end - start <= 2u || end > cfg->ptcl_capacity || end > rec[2]) return 0u;
    const uint32_t *ptcl = work + (size_t)cfg->tile_count * 6u;
    if (ptcl[(size_t)(end - 1u) * 6u] != 0u) return 0u;
    uint32_t depth = 0;
    for (uint32_t cursor = end - 1u; cursor > start;) {
        --cursor;
        uint32_t tag = ptcl[(size_t)cursor * 6u];
        if (tag == 4u || tag == 6u || tag == 8u) depth++;
        else if (tag == 3u || tag == 12u || tag == 5u || tag == 7u) {
            if (depth == 0u) return 0u;
            depth--;
        } else if (tag == 2u && depth == 0u && (ptcl[(size_t)cursor * 6u + 5u] >> 24) == 255u) {
            return cursor - start;
        }
    }
    return 0u;
}

/* coarse_emit_bins (coarse/emit.hlsl:82-142) with emit_stack_begins/ends (coarse/emit_stack.hlsli:18-44). */
void or_coarse_emit_bins(const or_coarse_config *cfg, const or_coarse_inputs *in, uint32_t *work, uint32_t *kinds) {
    CoarseJob j = {cfg, in, work, kinds};
    par_for(cfg->tile_count, 64u, coarse_emit_range, &j);
}
static void coarse_emit_range(void *ctx, uint32_t t0, uint32_t t1) {
    const CoarseJob *j = (const CoarseJob *)ctx;
    const or_coarse_config *cfg = j->cfg;
    const or_coarse_inputs *in = j->in;
    uint32_t *work = j->work, *kinds = j->kinds;
    for (uint32_t tile = t0; tile < t1; ++tile) {
        uint32_t tx = tile % cfg->tiles_width, ty = tile / cfg->tiles_width;
        uint32_t *rec = work + (size_t)tile * 6u;
        uint32_t px = rec[1], py = rec[2], raster_start = rec[1];
        if (px >= py) { kinds[tile] = 1u; continue; }
        uint32_t wrappers = stack_wrapper_count(cfg, in, tx, ty);
        if (wrappers == NONE) {
            store_particle(cfg, work, px, 0u, 0, 0, 0, 0, 0);
            kinds[tile] = 1u;
            continue;
        }
        uint32_t flags = wrappers != 0u ? 4u : 0u;
        uint32_t dst = px;
        for (uint32_t i = cfg->layer_stack_start; i < cfg->layer_stack_end; ++i) {
            const uint32_t *layer = in->layers + (size_t)i * 3u;
            Draw d = load_draw(in->draws, cfg->draw_count, layer[1]);
            if (noop_stack_clip(cfg, in, layer, &d, tx, ty)) continue;
            if (has_sdf(&d)) {
                store_particle(cfg, work, dst++, 12u, 0, 0, 0, 0, layer[1]);
            } else {
                uint32_t bd = draw_backdrop(cfg, in->paths, &d, tx, ty);
                uint32_t tag = layer[0] == 1u ? 5u : (layer[0] == 2u ? 7u : 3u);
                store_particle(cfg, work, dst++, tag, (uint32_t)in->backdrops[bd], d.fill_rule,
                               in->segment_ranges[2 * (size_t)bd], in->segment_ranges[2 * (size_t)bd + 1], layer[2]);
            }
        }
        px += wrappers;
        uint32_t start = in->tile_draws[2 * (size_t)tile], remaining = in->tile_draws[2 * (size_t)tile + 1];
        for (uint32_t i = 0; start != NONE && i < remaining; ++i) {
            uint32_t di = start + i < cfg->tile_draw_index_count ? in->tile_draw_indices[start + i] : 0u;
            Particle p = draw_particle(cfg, in, di, tx, ty);
            if (!p.valid) continue;
            Draw d = load_draw(in->draws, cfg->draw_count, di);
            uint32_t cls = 4u;
            if (p.tag == 2u) cls = 1u;
            else if (p.tag == 13u) cls = 2u;
            else if (p.tag == 9u && nontransparent_solid(cfg, in, &d) && d.sdf_offset != NONE &&
                     d.shadow_offset == NONE && d.sdf_len != 0u) {
                uint32_t kind = pw(cfg, in, d.sdf_offset);
                if (kind == 1u || kind == 5u || kind == 14u) cls = 2u;
            }
            flags |= cls;
            if (wrappers == 0u && p.tag == 2u && (p.color >> 24) == 255u && px > raster_start) flags |= 8u;
            if (px < py) store_particle(cfg, work, px++, p.tag, p.winding, p.fill_rule, p.s0, p.s1, p.color);
        }
        uint32_t end_dst = px;
        for (uint32_t i = cfg->layer_stack_end; i > cfg->layer_stack_start;) {
            --i;
            const uint32_t *layer = in->layers + (size_t)i * 3u;
            Draw d = load_draw(in->draws, cfg->draw_count, layer[1]);
            if (noop_stack_clip(cfg, in, layer, &d, tx, ty)) continue;
            uint32_t tag = layer[0] == 1u ? 6u : (layer[0] == 2u ? 8u : 4u);
            store_particle(cfg, work, end_dst++, tag, 0, 0, 0, 0, 0);
        }
        store_particle(cfg, work, px + wrappers, 0u, 0, 0, 0, 0, 0);
        uint32_t skip = (flags & 8u) ? occluded_prefix(cfg, work, tile, px + wrappers + 1u, wrappers) : 0u;
        uint32_t kind;
        if (flags & 4u) kind = 0u;
        else if (flags & 2u) kind = (flags & 1u) ? 4u : 3u;
        else kind = (flags & 1u) ? 2u : 1u;
        kinds[tile] = kind | (skip << 3);
    }
}

/* ---------- tile draw bins with early occlusion -------------------------------------
 * TileDrawBins::reset_transient (src/shared/gpu_plan/tile_bins/transient.rs:7-146) with
 * Occlusion::update (tile_bins/occlusion.rs:48-160), opaque_tiles / rect_tiles /
 * batch_clips (occlusion/geometry.rs:17-135) and for_each_visible
 * (occlusion/transient.rs:9-66).  An immediate canvas records its draws in painter
 * order, so a batch's painter order is increasing draw index. */

typedef struct { uint32_t x0, y0, x1, y1; } TBox;
static uint32_t tbox_count(TBox b) { return (b.x1 > b.x0 ? b.x1 - b.x0 : 0u) * (b.y1 > b.y0 ? b.y1 - b.y0 : 0u); }
static TBox tbox_and(TBox a, TBox b) {
    TBox r = {umax(a.x0, b.x0), umax(a.y0, b.y0), umin(a.x1, b.x1), umin(a.y1, b.y1)};
    return r;
}
/* PixelBounds::tile_bbox (src/shared/bounds.rs:113-128). */
static TBox draw_tbox(const uint32_t *dw, uint32_t tw, uint32_t th) {
    TBox b = {tmin_of((int32_t)dw[10], tw), tmin_of((int32_t)dw[11], th), tmax_of((int32_t)dw[12], tw),
              tmax_of((int32_t)dw[13], th)};
    return b;
}
static double rect_radius(float r, double max_radius) {
    double v = (double)r;
    return v < 0.0 ? 0.0 : (v > max_radius ? max_radius : v);
}
static uint32_t whole_tiles(double t) { return t > 0.0 ? (uint32_t)t : 0u; }

/* rect_tiles (occlusion/geometry.rs:79-135). */
static TBox rect_tiles(const uint32_t *dw, const uint32_t *paint, uint32_t paint_sdf_words, uint32_t tw, uint32_t th) {
    TBox none = {0, 0, 0, 0};
    if (dw[0] != NONE || dw[1] != NONE || dw[4] != NONE) return none;
    if (bits_f(dw[19]) != 1.0f || bits_f(dw[20]) != 0.0f || bits_f(dw[21]) != 0.0f || bits_f(dw[22]) != 1.0f)
        return none;
    if (dw[2] == NONE || dw[3] != 17u || (uint64_t)dw[2] + 17u > paint_sdf_words) return none;
    const uint32_t *rec = paint + dw[2];
    if (rec[0] != 1u) return none; /* GPU_SDF_RECT */
    float v[10];
    int whole = 1;
    for (int i = 0; i < 8; ++i) v[i] = bits_f(rec[1 + i]);
    v[8] = bits_f(dw[23]);
    v[9] = bits_f(dw[24]);
    for (int i = 0; i < 10; ++i) {
        if (!isfinite(v[i]) || fabsf(v[i]) > 65536.0f) return none;
        if (v[i] != truncf(v[i])) whole = 0;
    }
    double x0 = fmin((double)v[0], (double)v[2]), x1 = fmax((double)v[0], (double)v[2]);
    double y0 = fmin((double)v[1], (double)v[3]), y1 = fmax((double)v[1], (double)v[3]);
    double max_radius = fmax(fmin(x1 - x0, y1 - y0) * 0.5, 0.0);
    double tl = rect_radius(v[4], max_radius), tr = rect_radius(v[5], max_radius);
    double bl = rect_radius(v[6], max_radius), br = rect_radius(v[7], max_radius);
    double guard = whole ? 0.0 : 1.0 / 32.0, ex = (double)v[8], ey = (double)v[9];
    TBox bb = draw_tbox(dw, tw, th), r;
    r.x0 = umax(whole_tiles(ceil((x0 + fmax(tl, bl) + ex + guard) / 16.0)), bb.x0);
    r.y0 = umax(whole_tiles(ceil((y0 + fmax(tl, tr) + ey + guard) / 16.0)), bb.y0);
    r.x1 = umin(whole_tiles(floor((x1 - fmax(tr, br) + ex - guard) / 16.0)), bb.x1);
    r.y1 = umin(whole_tiles(floor((y1 - fmax(bl, br) + ey - guard) / 16.0)), bb.y1);
    return r;
}

/* Visits the tiles of draw d's box outside `hidden` (tile index, draw). */
typedef struct {
    uint32_t *tile_draws, *indices, *counts;
    uint32_t tw, pass; /* pass 0 counts, pass 1 fills each list from its end */
} BinVisit;
static void bin_region(BinVisit *v, TBox r, uint32_t d) {
    for (uint32_t y = r.y0; y < r.y1; ++y)
        for (uint32_t x = r.x0; x < r.x1; ++x) {
            uint32_t t = y * v->tw + x;
            if (v->pass == 0) v->counts[t]++;
            else v->indices[v->tile_draws[2 * t] + --v->counts[t]] = d;
        }
}

int64_t or_tile_draw_bins(const or_scene *s, int cull, uint32_t *tile_draws, uint32_t *indices, uint64_t capacity) {
    const uint32_t tw = s->tiles_width, th = s->tiles_height, T = tw * th, D = s->draw_count, Q = s->batch_count;
    uint32_t *first = malloc(sizeof(uint32_t) * (Q + 1u));
    uint8_t *enabled = malloc(Q + 1u), *occluder = malloc(Q + 1u);
    TBox *clip = malloc(sizeof(TBox) * (Q + 1u)), *opaque = calloc(D + 1u, sizeof(TBox));
    uint32_t *order = malloc(sizeof(uint32_t) * (D + 1u)), *counts = calloc(T + 1u, sizeof(uint32_t));
    uint32_t *vis_batch = malloc(sizeof(uint32_t) * (D + 1u));
    const uint32_t sdf_words = s->paint_sdf_shadow_base;
    const uint32_t *brush = s->paint + s->paint_brush_base;
    const uint32_t brush_words = s->paint_words - s->paint_brush_base;
    uint32_t n = 0;
    int active = 0;
    for (uint32_t b = 0; b < Q; ++b) {
        first[b] = NONE;
        enabled[b] = 1;
        occluder[b] = 0;
        TBox e = {0, 0, 0, 0};
        clip[b] = e;
    }
    /* painter order: batches in order, each batch's draws by index */
    for (uint32_t b = 0; b < Q; ++b)
        for (uint32_t d = 0; d < D; ++d)
            if (s->draw_batch_ids[d] == b) {
                order[n++] = d;
                const uint32_t tag = s->draws[31 * d + 8];
                enabled[b] &= (uint8_t)(tag == 0u || tag == 5u); /* Brush, PathGlyph */
                if (first[b] == NONE) first[b] = d;
            }
    /* candidates (draw index order): opaque_tiles of every non-first member */
    for (uint32_t d = 0; d < D && cull; ++d) {
        const uint32_t *dw = s->draws + 31 * d, b = s->draw_batch_ids[d];
        if (b >= Q || !enabled[b] || first[b] == d) continue;
        if (dw[8] != 0u || dw[7] < 5u || (uint64_t)dw[6] + 5u > brush_words) continue;
        if (brush[dw[6]] != 1u || (brush[dw[6] + 4u] >> 24) != 255u) continue; /* solid, alpha 255 */
        opaque[d] = rect_tiles(dw, s->paint, sdf_words, tw, th);
        if (tbox_count(opaque[d])) occluder[b] = 1;
    }
    for (uint32_t b = 0; b < Q; ++b) {
        if (!occluder[b]) continue;
        TBox c = {0, 0, tw, th};
        for (uint32_t l = s->batches[2 * b]; l < s->batches[2 * b + 1]; ++l) {
            const uint32_t *layer = s->layers + 3 * l;
            if (layer[0] == 0u && layer[1] < D) /* LayerStackEntry::Clip */
                c = tbox_and(c, rect_tiles(s->draws + 31 * layer[1], s->paint, sdf_words, tw, th));
        }
        clip[b] = c;
    }
    for (uint32_t d = 0; d < D; ++d) {
        const uint32_t b = s->draw_batch_ids[d];
        vis_batch[d] = b < Q && enabled[b] ? b : NONE;
        if (vis_batch[d] != NONE) opaque[d] = tbox_and(opaque[d], clip[b]);
        else opaque[d].x1 = opaque[d].x0;
        if (tbox_count(opaque[d])) active = 1;
    }
    BinVisit v = {tile_draws, indices, counts, tw, 0};
    int64_t total = 0;
    for (uint32_t pass = 0; pass < 2; ++pass) {
        v.pass = pass;
        if (pass == 1) {
            for (uint32_t t = 0; t < T; ++t) {
                tile_draws[2 * t] = counts[t] ? (uint32_t)total : NONE;
                tile_draws[2 * t + 1] = counts[t];
                total += counts[t];
            }
            if ((uint64_t)total > capacity) break;
        }
        uint32_t batch = NONE;
        TBox cover = {0, 0, 0, 0};
        for (uint32_t k = n; k-- > 0;) {
            const uint32_t d = order[k], vb = active ? vis_batch[d] : NONE;
            if (batch != vb || batch == NONE) cover.x1 = cover.x0 = cover.y0 = cover.y1 = 0;
            batch = vb;
            TBox bb = draw_tbox(s->draws + 31 * d, tw, th), hid = tbox_and(bb, cover);
            if (tbox_count(hid) == 0) {
                bin_region(&v, bb, d);
            } else {
                TBox top = {bb.x0, bb.y0, bb.x1, hid.y0}, bottom = {bb.x0, hid.y1, bb.x1, bb.y1};
                TBox left = {bb.x0, hid.y0, hid.x0, hid.y1}, right = {hid.x1, hid.y0, bb.x1, hid.y1};
                bin_region(&v, top, d);
                bin_region(&v, bottom, d);
                bin_region(&v, left, d);
                bin_region(&v, right, d);
            }
            if (vb != NONE && tbox_count(opaque[d]) > tbox_count(cover)) cover = opaque[d];
        }
    }
    free(first); free(enabled); free(occluder); free(clip); free(opaque); free(order); free(counts); free(vis_batch);
    return (uint64_t)total > capacity ? -1 : total;
}

/* ---------- fine (fine/main.hlsl, interpreter.hlsli, specialized.hlsli, clip.hlsli) ---- */

typedef struct { uint32_t brush_offset, sdf_offset, shadow_offset; int32_t pb[4]; Aff inv; } FDraw;
static FDraw load_fdraw(const uint32_t *draws, uint32_t count, uint32_t index) {
    FDraw f;
    memset(&f, 0, sizeof(f));
    if (index >= count) return f;
    const uint32_t *r = draws + (size_t)index * 31u;
    f.brush_offset = r[6]; f.sdf_offset = r[2]; f.shadow_offset = r[4];
    for (int i = 0; i < 4; ++i) f.pb[i] = (int32_t)r[10 + i];
    f.inv.a = bits_f(r[25]); f.inv.b = bits_f(r[26]); f.inv.c = bits_f(r[27]);
    f.inv.d = bits_f(r[28]); f.inv.e = bits_f(r[29]); f.inv.f = bits_f(r[30]);
    return f;
}
static void aff_point(const Aff *t, float x, float y, float *ox, float *oy) {
    *ox = fmaf(t->a, x, fmaf(t->c, y, t->e));
    *oy = fmaf(t->b, x, fmaf(t->d, y, t->f));
}
static uint32_t draw_brush(const or_fine_config *cfg, const uint32_t *paint, const FDraw *d, float x, float y) {
    if (blob_word(paint, cfg->paint_words, cfg->paint_brush_base, d->brush_offset) == 1u)
        return blob_word(paint, cfg->paint_words, cfg->paint_brush_base, d->brush_offset + 4u);
    float lx, ly;
    aff_point(&d->inv, x, y, &lx, &ly);
    const Atlas atlas = {cfg->atlas, cfg->atlas_width, cfg->atlas_height, cfg->atlas_pages};
    return sample_brush_atlas(paint, cfg->paint_words, cfg->paint_brush_base, d->brush_offset, lx, ly, &atlas);
}
static float draw_sdf_coverage(const or_fine_config *cfg, const uint32_t *paint, const FDraw *d, float x, float y) {
    float lx, ly;
    aff_point(&d->inv, x, y, &lx, &ly);
    if (d->sdf_offset != NONE) return sdf_coverage_from_blob(paint, cfg->paint_words, d->sdf_offset, lx, ly, &d->inv);
    if (d->shadow_offset != NONE)
        return sdf_coverage_from_blob(paint, cfg->paint_words, cfg->paint_sdf_shadow_base + d->shadow_offset, lx, ly,
                                      &d->inv);
    return 0.0f;
}

/* One pixel of one batch, from the loaded target value. */
static Px fine_pixel(const or_fine_config *cfg, const uint32_t *draws, const uint32_t *paint, const uint32_t *segs,
                     const uint32_t *work, uint32_t kind_word, uint32_t tile, uint32_t lane, Px pixel,
                     uint8_t *spill /* clip_spill_depth bytes */) {
    uint32_t kind = kind_word & 7u;
    if (kind == 1u) return pixel;
    const uint32_t *rec = work + (size_t)tile * 6u;
    uint32_t start = rec[1] + (kind_word >> 3), end = rec[2];
    const uint32_t *ptcl = work + cfg->ptcl_base_words;
    uint32_t lx = lane % TILE, ly = lane / TILE;
    uint32_t gx = (tile % cfg->tiles_width) * TILE + lx, gy = (tile / cfg->tiles_width) * TILE + ly;
    float sx = (float)gx + 0.5f, sy = (float)gy + 0.5f;
    if (kind == 2u) {
        for (uint32_t i = start; i < end; ++i) {
            const uint32_t *r = ptcl + (size_t)i * 6u;
            if (r[0] == 0u) break;
            pixel = premul_opaque(r[5]) ? rgba8_to_unorm(r[5]) : src_over_premul_unorm(pixel, rgba8_to_unorm(r[5]));
        }
        return pixel;
    }
    if (kind == 3u || kind == 4u) {
        for (uint32_t i = start; i < end; ++i) {
            const uint32_t *r = ptcl + (size_t)i * 6u;
            if (r[0] == 0u) break;
            if (r[0] == 2u) {
                pixel = premul_opaque(r[5]) ? rgba8_to_unorm(r[5]) : src_over_premul_unorm(pixel, rgba8_to_unorm(r[5]));
            } else if (r[0] == 13u) { /* PTCL_IMAGE (fine/specialized.hlsli:84-91) */
                FDraw d = load_fdraw(draws, cfg->draw_count, r[5]);
                uint32_t color = draw_brush(cfg, paint, &d, sx, sy);
                pixel = premul_opaque(color) ? rgba8_to_unorm(color) : src_over_premul_unorm(pixel, rgba8_to_unorm(color));
            } else {
                FDraw d = load_fdraw(draws, cfg->draw_count, r[5]);
                float coverage = 0.0f;
                if ((int32_t)gx >= d.pb[0] && (int32_t)gy >= d.pb[1] && (int32_t)gx < d.pb[2] && (int32_t)gy < d.pb[3])
                    coverage = draw_sdf_coverage(cfg, paint, &d, sx, sy);
                uint32_t alpha = coverage_to_u8(coverage);
                if (alpha != 0u) {
                    uint32_t color = blob_word(paint, cfg->paint_words, cfg->paint_brush_base, d.brush_offset + 4u);
                    pixel = alpha == 255u && premul_opaque(color) ? rgba8_to_unorm(color)
                                                                  : src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(color, alpha));
                }
            }
        }
        return pixel;
    }
    /* tile_pixel (fine/interpreter.hlsli:21-219) */
    uint32_t clip_mask = 255u, depth = 0u, stack[4] = {255u, 255u, 255u, 255u};
    for (uint32_t i = start; i < end; ++i) {
        const uint32_t *r = ptcl + (size_t)i * 6u;
        uint32_t tag = r[0];
        if (tag == 0u) break;
        if (tag == 2u) {
            if (clip_mask != 0u) {
                if (clip_mask == 255u && premul_opaque(r[5])) pixel = rgba8_to_unorm(r[5]);
                else pixel = src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(r[5], clip_mask));
            }
        } else if (tag == 13u) { /* PTCL_IMAGE (fine/interpreter.hlsli:65-74) */
            if (clip_mask != 0u) {
                FDraw d = load_fdraw(draws, cfg->draw_count, r[5]);
                uint32_t color = draw_brush(cfg, paint, &d, sx, sy);
                if (clip_mask == 255u && premul_opaque(color)) pixel = rgba8_to_unorm(color);
                else pixel = src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(color, clip_mask));
            }
        } else if (tag == 9u) {
            if (clip_mask != 0u) {
                FDraw d = load_fdraw(draws, cfg->draw_count, r[5]);
                uint32_t alpha = combine_alpha(coverage_to_u8(draw_sdf_coverage(cfg, paint, &d, sx, sy)), clip_mask);
                if (alpha != 0u) {
                    uint32_t color = draw_brush(cfg, paint, &d, sx, sy);
                    if (alpha == 255u && premul_opaque(color)) pixel = rgba8_to_unorm(color);
                    else pixel = src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(color, alpha));
                }
            }
        } else if (tag == 4u) {
            if (depth > 0u) {
                depth -= 1u;
                if (depth < 4u) clip_mask = stack[depth];
                else if (depth - 4u < cfg->clip_spill_depth) clip_mask = spill[depth - 4u];
            } else {
                clip_mask = 255u;
            }
        } else if (tag == 12u || tag == 1u || tag == 3u || tag == 11u) {
            uint32_t parent = clip_mask;
            uint32_t alpha;
            if (tag == 12u) {
                alpha = 0u;
                if (parent != 0u) {
                    FDraw d = load_fdraw(draws, cfg->draw_count, r[5]);
                    alpha = coverage_to_u8(draw_sdf_coverage(cfg, paint, &d, sx, sy));
                }
            } else {
                alpha = coverage_to_alpha(fill_value_at(segs, (int32_t)r[1], r[3], r[4], lx, ly), r[2]);
            }
            if (tag == 1u || tag == 11u) {
                uint32_t masked = combine_alpha(alpha, clip_mask);
                if (masked != 0u) {
                    FDraw d = load_fdraw(draws, cfg->draw_count, r[5]);
                    uint32_t color = draw_brush(cfg, paint, &d, sx, sy);
                    if (tag == 11u) /* PTCL_PATH_GLYPH (fine/interpreter.hlsli:205-206) */
                        pixel = rgba8_to_unorm(src_over_mask_linear_auto_u8(unorm_to_rgba8(pixel), color, masked));
                    else if (masked == 255u && premul_opaque(color)) pixel = rgba8_to_unorm(color);
                    else pixel = src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(color, masked));
                }
            } else {
                /* push_clip (fine/clip.hlsli:7-38) */
                if (depth < 4u) stack[depth++] = parent;
                else if (depth - 4u < cfg->clip_spill_depth) { spill[depth - 4u] = (uint8_t)parent; depth++; }
                clip_mask = combine_alpha(parent, alpha);
            }
        }
    }
    return pixel;
}

/* fine_tile_main (fine/main.hlsl:23-40) with load_target: reads and writes the RGBA8 target. */
typedef struct {
    const or_fine_config *cfg;
    const uint32_t *draws, *paint, *segments, *work, *kinds;
    uint32_t *target;
} FineJob;
static void fine_range(void *ctx, uint32_t t0, uint32_t t1);

void or_fine(const or_fine_config *cfg, const uint32_t *draws, const uint32_t *paint, const uint32_t *segments,
             const uint32_t *work, const uint32_t *kinds, uint32_t *target) {
    FineJob j = {cfg, draws, paint, segments, work, kinds, target};
    par_for(cfg->tile_count, 16u, fine_range, &j);
}
static void fine_range(void *ctx, uint32_t t0, uint32_t t1) {
    const FineJob *j = (const FineJob *)ctx;
    const or_fine_config *cfg = j->cfg;
    const uint32_t *draws = j->draws, *paint = j->paint, *segments = j->segments, *work = j->work, *kinds = j->kinds;
    uint32_t *target = j->target;
    uint8_t *spill = (uint8_t *)malloc(cfg->clip_spill_depth + 1u);
    for (uint32_t tile = t0; tile < t1; ++tile) {
        for (uint32_t lane = 0; lane < 256u; ++lane) {
            uint32_t gx = (tile % cfg->tiles_width) * TILE + lane % TILE;
            uint32_t gy = (tile / cfg->tiles_width) * TILE + lane / TILE;
            if (gx >= cfg->width || gy >= cfg->height) continue;
            uint32_t *dst = target + (size_t)gy * cfg->width + gx;
            Px pixel = rgba8_to_unorm(*dst);
            pixel = fine_pixel(cfg, draws, paint, segments, work, kinds[tile], tile, lane, pixel, spill);
            *dst = unorm_to_rgba8(pixel);
        }
    }
    free(spill);
}

/* ---------- whole frame ------------------------------------------------------------ */

int or_render_frame(const or_scene *s, uint32_t *image, or_frame_outputs *out) {
    int status = 0;
    const uint32_t T = s->tiles_width * s->tiles_height;
    const uint32_t B = s->backdrop_len;
    or_scan_config sc;
    memset(&sc, 0, sizeof(sc));
    sc.backdrop_len = B;
    sc.path_count = s->path_count;
    sc.scan_chunk_count = s->scan_chunk_count;
    sc.line_count = s->line_count;
    sc.segment_capacity = s->segment_capacity;

    int32_t *backdrops = (int32_t *)calloc(B + 1u, 4);
    uint32_t *counts = (uint32_t *)calloc(B + 1u, 4);
    uint32_t *ranges = (uint32_t *)calloc(2u * (size_t)B + 2u, 4);
    uint32_t *cursors = (uint32_t *)calloc(B + 1u, 4);
    uint32_t *bumps = (uint32_t *)calloc(s->path_count + 1u, 4);
    uint32_t *totals = (uint32_t *)calloc(s->scan_chunk_count + 1u, 4);
    uint32_t *offsets = (uint32_t *)calloc(s->scan_chunk_count + 1u, 4);
    uint32_t *segments = (uint32_t *)calloc(5u * (size_t)s->segment_capacity + 5u, 4);

    double t0 = now_ms(), ms[6] = {0, 0, 0, 0, 0, 0}; /* count, prefix, emit, cumsum, coarse, fine */
    or_scan_count(&sc, s->lines, s->paths, s->path_count, backdrops, counts);
    double t1 = now_ms();
    or_scan_prefix_chunks(&sc, s->scan_chunks, counts, ranges, totals);
    or_scan_chunk_offsets(&sc, s->paths, s->scan_chunk_ranges, bumps, totals, offsets);
    or_scan_apply_chunk_offsets(&sc, s->scan_chunks, ranges, cursors, offsets);
    double t2 = now_ms();
    or_scan_emit(&sc, s->lines, s->paths, s->path_count, cursors, segments);
    double t3 = now_ms();
    for (uint32_t p = 0; p < s->path_count; ++p)
        if (bumps[p] > s->segment_capacity) status = 1;
    if (out && out->raw_backdrops) memcpy(out->raw_backdrops, backdrops, (size_t)B * 4u);
    double t4 = now_ms();
    or_cumsum(s->cumsum_row_count, s->cumsum_chunk_count, s->cumsum_chunk_offsets, s->cumsum_chunk_lens,
              s->cumsum_row_starts, s->cumsum_row_ends, backdrops);
    ms[0] = t1 - t0;
    ms[1] = t2 - t1;
    ms[2] = t3 - t2;
    ms[3] = now_ms() - t4;

    /* root clear (filter_clear_region), then coarse + fine per batch on the RGBA8 target */
    for (size_t i = 0; i < (size_t)s->width * s->height; ++i) image[i] = s->clear_color;
    uint32_t *work = NULL;
    size_t work_words = 0;
    uint32_t *kinds = (uint32_t *)calloc(T + 1u, 4);
    uint32_t chunk_count = (T + 255u) / 256u;
    uint32_t *chunk_records = (uint32_t *)calloc(4u * (size_t)chunk_count + 4u, 4);
    uint64_t base = 0;
    or_coarse_inputs in = {s->draws, s->paths, backdrops, ranges, s->layers, s->paint, s->draw_batch_ids,
                           s->tile_draws, s->tile_draw_indices};
    for (uint32_t b = 0; b < s->batch_count; ++b) {
        or_coarse_config cc;
        memset(&cc, 0, sizeof(cc));
        cc.tile_count = T;
        cc.tiles_width = s->tiles_width;
        cc.tiles_height = s->tiles_height;
        cc.batch_id = b;
        cc.layer_stack_start = s->batches[2 * b];
        cc.layer_stack_end = s->batches[2 * b + 1];
        cc.paint_brush_base = s->paint_brush_base;
        cc.draw_count = s->draw_count;
        cc.path_count = s->path_count;
        cc.paint_words = s->paint_words;
        cc.layer_count = s->layer_count;
        cc.tile_draw_index_count = s->tile_draw_index_count;
        /* count with a tile-record-only work buffer, then size the PTCL exactly */
        size_t need = (size_t)T * 6u;
        if (need > work_words) {
            free(work);
            work_words = need;
            work = (uint32_t *)calloc(work_words, 4);
        }
        cc.ptcl_capacity = 0;
        double tc = now_ms();
        or_coarse_count_bins(&cc, &in, work, kinds);
        or_coarse_prefix(T, work, chunk_records);
        uint32_t total = 0;
        for (uint32_t c = 0; c < chunk_count; ++c) total += chunk_records[4 * c];
        need = (size_t)T * 6u + (size_t)total * 6u;
        if (need > work_words) {
            uint32_t *grown = (uint32_t *)calloc(need, 4);
            memcpy(grown, work, (size_t)T * 6u * 4u);
            free(work);
            work = grown;
            work_words = need;
        }
        cc.ptcl_capacity = total;
        or_coarse_emit_bins(&cc, &in, work, kinds);
        ms[4] += now_ms() - tc;
        if (out) {
            for (uint32_t t = 0; t < T; ++t) {
                size_t idx = (size_t)b * T + t;
                if (out->tile_ptcl_counts) out->tile_ptcl_counts[idx] = work[(size_t)t * 6u];
                if (out->tile_ptcl_starts) out->tile_ptcl_starts[idx] = work[(size_t)t * 6u + 1u];
                if (out->tile_ptcl_ends) out->tile_ptcl_ends[idx] = work[(size_t)t * 6u + 2u];
                if (out->tile_kinds) out->tile_kinds[idx] = kinds[t];
            }
            if (out->batch_bases) out->batch_bases[b] = (uint32_t)base;
            if (out->ptcl) {
                if ((base + total) * 6u > out->ptcl_capacity_words) status = 2;
                else memcpy(out->ptcl + base * 6u, work + (size_t)T * 6u, (size_t)total * 24u);
            }
        }
        base += total;
        or_fine_config fc;
        memset(&fc, 0, sizeof(fc));
        fc.width = s->width;
        fc.height = s->height;
        fc.tile_count = T;
        fc.tiles_width = s->tiles_width;
        fc.tiles_height = s->tiles_height;
        fc.clip_spill_depth = s->max_clip_depth > 4u ? s->max_clip_depth - 4u : 0u;
        fc.paint_words = s->paint_words;
        fc.paint_sdf_shadow_base = s->paint_sdf_shadow_base;
        fc.paint_brush_base = s->paint_brush_base;
        fc.draw_count = s->draw_count;
        fc.ptcl_base_words = T * 6u;
        fc.atlas = s->atlas;
        fc.atlas_width = s->atlas_width;
        fc.atlas_height = s->atlas_height;
        fc.atlas_pages = s->atlas ? s->atlas_pages : 0u;
        double tf = now_ms();
        or_fine(&fc, s->draws, s->paint, segments, work, kinds, image);
        ms[5] += now_ms() - tf;
    }
    if (out && out->stage_ms) memcpy(out->stage_ms, ms, sizeof(ms));
    if (out && out->batch_bases) out->batch_bases[s->batch_count] = (uint32_t)base;
    if (out) {
        if (out->backdrops) memcpy(out->backdrops, backdrops, (size_t)B * 4u);
        if (out->segment_counts) memcpy(out->segment_counts, counts, (size_t)B * 4u);
        if (out->segment_ranges) memcpy(out->segment_ranges, ranges, (size_t)B * 8u);
        if (out->segment_cursors) memcpy(out->segment_cursors, cursors, (size_t)B * 4u);
        if (out->segment_bumps) memcpy(out->segment_bumps, bumps, (size_t)s->path_count * 4u);
        if (out->segments) memcpy(out->segments, segments, (size_t)s->segment_capacity * 20u);
    }
    free(work); free(kinds); free(chunk_records);
    free(backdrops); free(counts); free(ranges); free(cursors); free(bumps); free(totals); free(offsets); free(segments);
    return status;
}

/* ---------- validation entries ---------------------------------------------------- */

void or_fill_coverage_words(uint32_t count, const uint32_t *q, const uint32_t *segments, uint32_t *alpha_out,
                            float *coverage_out) {
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t *r = q + (size_t)i * 8u;
        float v = fill_value_at(segments, (int32_t)r[2], r[0], r[1], r[3], r[4]);
        alpha_out[i] = coverage_to_alpha(v, r[5]);
        if (coverage_out) coverage_out[i] = v;
    }
}

void or_fill_coverage_words_sequential(uint32_t count, const uint32_t *q, const uint32_t *segments,
                                       uint32_t *alpha_out, float *coverage_out) {
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t *r = q + (size_t)i * 8u;
        float v = fill_value_sequential(segments, (int32_t)r[2], r[0], r[1], r[3], r[4]);
        alpha_out[i] = coverage_to_alpha(v, r[5]);
        if (coverage_out) coverage_out[i] = v;
    }
}

void or_pixel_math_words(uint32_t count, const uint32_t *q, uint32_t *out) {
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t *r = q + (size_t)i * 8u;
        uint32_t *o = out + (size_t)i * 14u;
        uint32_t a = r[0], b = r[1], src = r[2], dst = r[3];
        float c = bits_f(r[4]), t = bits_f(r[5]);
        uint32_t rule = r[6];
        Px pixel = rgba8_to_unorm(src);
        o[0] = mul_div255(a, b);
        o[1] = combine_alpha(a, b);
        o[2] = scale_premul_u8(src, b);
        o[3] = src_over_premul_u8(dst, src);
        o[4] = src;
        o[5] = coverage_to_u8(c);
        o[6] = coverage_to_alpha(c, rule);
        o[7] = lerp_premul_u8(src, dst, t);
        o[8] = unorm_to_rgba8(src_over_premul_unorm(rgba8_to_unorm(dst), rgba8_to_unorm(src)));
        o[9] = unorm_to_rgba8(scale_premul_u8_to_unorm(src, b));
        o[10] = coverage_to_u8(straight_channel(a, b));
        o[11] = pack_premul_rgba8(pixel.c[0], pixel.c[1], pixel.c[2], pixel.c[3]);
        o[12] = coverage_to_u8(rem_euclid_f32(c, 3.0f));
        o[13] = f_bits(signum_f32(c));
    }
}

/* ---------- FP64 area oracles ------------------------------------------------------ */

/* Exact integral over y in [ya, yb] of (X + 1 - clamp(x(y), X, X + 1)), x linear in y. */
static double band_area(double x0, double y0, double x1, double y1, double X, double ya, double yb) {
    if (!(yb > ya)) return 0.0;
    double cuts[4] = {ya, yb, ya, ya};
    int n = 2;
    if (y1 != y0) {
        double k = (x1 - x0) / (y1 - y0);
        double edges[2] = {X, X + 1.0};
        for (int e = 0; e < 2; ++e) {
            if (k != 0.0) {
                double yc = y0 + (edges[e] - x0) / k;
                if (yc > ya && yc < yb) cuts[n++] = yc;
            }
        }
    }
    /* sort the (at most 4) cut points */
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && cuts[j] < cuts[j - 1]; --j) {
            double t = cuts[j]; cuts[j] = cuts[j - 1]; cuts[j - 1] = t;
        }
    double total = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
        double u = cuts[i], v = cuts[i + 1];
        if (!(v > u)) continue;
        double xu, xv;
        if (y1 != y0) {
            xu = x0 + (x1 - x0) * (u - y0) / (y1 - y0);
            xv = x0 + (x1 - x0) * (v - y0) / (y1 - y0);
        } else {
            xu = xv = x0;
        }
        /* within a piece the clamp is affine, so the trapezoid rule is exact */
        double fu = X + 1.0 - fmin(fmax(xu, X), X + 1.0);
        double fv = X + 1.0 - fmin(fmax(xv, X), X + 1.0);
        total += 0.5 * (fu + fv) * (v - u);
    }
    return total;
}

void or_fill_coverage_f64(uint32_t count, const uint32_t *q, const uint32_t *segments, double *out) {
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t *r = q + (size_t)i * 8u;
        double X = (double)r[3], Y = (double)r[4];
        double total = (double)(int32_t)r[2];
        for (uint32_t k = r[0]; k < r[1]; ++k) {
            const uint32_t *sgm = segments + (size_t)k * 5u;
            double p0x = bits_f(sgm[0]), p0y = bits_f(sgm[1]), p1x = bits_f(sgm[2]), p1y = bits_f(sgm[3]);
            double y_edge = bits_f(sgm[4]);
            double edge = (p1x - p0x < 0.0 ? -1.0 : 1.0) * fmin(fmax(Y - y_edge + 1.0, 0.0), 1.0);
            total += edge;
            double lo = fmax(fmin(p0y, p1y), Y), hi = fmin(fmax(p0y, p1y), Y + 1.0);
            double sign = p0y > p1y ? 1.0 : -1.0;
            total += sign * band_area(p0x, p0y, p1x, p1y, X, lo, hi);
        }
        out[i] = total;
    }
}

double or_polygon_pixel_area_f64(const uint32_t *lines, uint32_t n, double px, double py) {
    double total = 0.0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t *l = lines + (size_t)k * 6u;
        double x0 = bits_f(l[2]), y0 = bits_f(l[3]), x1 = bits_f(l[4]), y1 = bits_f(l[5]);
        if (y0 == y1) continue;
        double lo = fmax(fmin(y0, y1), py), hi = fmin(fmax(y0, y1), py + 1.0);
        double sign = y0 > y1 ? 1.0 : -1.0;
        total += sign * band_area(x0, y0, x1, y1, px, lo, hi);
    }
    return total;
}

/* ------------------------------------------------------------------------ */
/* The differentiable form in FP64: the checker of tk_frame_render_diff and   */
/* tk_frame_backward.  Not a restatement of a reference routine (the          */
/* reference has no backward pass): the forward is the fine composite         */
/* (fine/main.hlsl:23-40, shared/pixel.hlsli:7-66) with exact pixel areas in  */
/* place of the FP32 sums and no RGBA8 steps.                                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    int x0, y0, w, h; /* pixel box, clipped to the canvas */
} DiffBox;

static DiffBox diff_box(const double *lines, uint32_t start, uint32_t count, uint32_t width, uint32_t height) {
    double xl = INFINITY, xh = -INFINITY, yl = INFINITY, yh = -INFINITY;
    for (uint32_t i = start; i < start + count; ++i) {
        const double *l = lines + 4 * (size_t)i;
        xl = fmin(xl, fmin(l[0], l[2]));
        xh = fmax(xh, fmax(l[0], l[2]));
        yl = fmin(yl, fmin(l[1], l[3]));
        yh = fmax(yh, fmax(l[1], l[3]));
    }
    DiffBox b = {0, 0, 0, 0};
    if (!count || !(xh >= xl) || !(yh >= yl)) return b;
    double fx0 = fmax(floor(xl), 0.0), fy0 = fmax(floor(yl), 0.0);
    double fx1 = fmin(ceil(xh), (double)width), fy1 = fmin(ceil(yh), (double)height);
    if (fx1 <= fx0 || fy1 <= fy0) return b;
    b.x0 = (int)fx0;
    b.y0 = (int)fy0;
    b.w = (int)(fx1 - fx0);
    b.h = (int)(fy1 - fy0);
    return b;
}

/* Signed area accumulation: each line adds, in every row slab it crosses, its
 * exact signed area to the right of it inside each cell and the rest of its dy
 * to the next cell; a prefix sum along each row then gives the integral of the
 * winding number over every pixel of the box (closed paths: zero outside the
 * box of their lines).  dy is taken in line order (y at the first parameter
 * minus y at the second), the sign convention of segment_row_parts. */
static void diff_area(const double *lines, uint32_t start, uint32_t count, DiffBox b, double *area) {
    const int W = b.w + 1;
    double *acc = (double *)calloc((size_t)b.h * W, sizeof(double));
    for (uint32_t i = start; i < start + count; ++i) {
        const double *l = lines + 4 * (size_t)i;
        const double x0 = l[0], y0 = l[1], dx = l[2] - l[0], dy = l[3] - l[1];
        if (dy == 0.0) continue;
        int r0 = (int)floor(fmin(y0, l[3])), r1 = (int)ceil(fmax(y0, l[3])) - 1;
        if (r0 < b.y0) r0 = b.y0;
        if (r1 > b.y0 + b.h - 1) r1 = b.y0 + b.h - 1;
        for (int r = r0; r <= r1; ++r) {
            double ta = ((double)r - y0) / dy, tb = ((double)(r + 1) - y0) / dy;
            double sa = fmax(fmin(ta, tb), 0.0), sb = fmin(fmax(ta, tb), 1.0);
            if (!(sb > sa)) continue;
            double *row = acc + (size_t)(r - b.y0) * W;
            /* pieces of [sa, sb] between consecutive integer x */
            double xa = x0 + sa * dx, xb = x0 + sb * dx;
            int c0 = (int)floor(fmin(xa, xb)), c1 = dx == 0.0 ? c0 : (int)ceil(fmax(xa, xb)) - 1;
            for (int c = c0; c <= c1; ++c) {
                double ua = sa, ub = sb;
                if (dx != 0.0) {
                    double qa = ((double)c - x0) / dx, qb = ((double)(c + 1) - x0) / dx;
                    ua = fmax(fmin(qa, qb), sa);
                    ub = fmin(fmax(qa, qb), sb);
                }
                if (!(ub > ua)) continue;
                double py0 = y0 + ua * dy, py1 = y0 + ub * dy;
                double d = py0 - py1; /* this piece's dy */
                if (c < b.x0) {
                    row[0] += d;
                } else if (c < b.x0 + b.w) {
                    double mx = 0.5 * ((x0 + ua * dx) + (x0 + ub * dx));
                    double inside = d * ((double)(c + 1) - mx);
                    row[c - b.x0] += inside;
                    row[c - b.x0 + 1] += d - inside;
                }
            }
        }
    }
    for (int y = 0; y < b.h; ++y) {
        double s = 0.0;
        for (int x = 0; x < b.w; ++x) {
            s += acc[(size_t)y * W + x];
            area[(size_t)y * b.w + x] = s;
        }
    }
    free(acc);
}

static double diff_alpha64(double v, uint32_t rule, double *slope) {
    if (rule == 1u) {
        double w = v - 2.0 * nearbyint(0.5 * v);
        *slope = w < 0.0 ? -1.0 : 1.0;
        return fabs(w);
    }
    double m = fabs(v);
    *slope = m < 1.0 ? (v < 0.0 ? -1.0 : 1.0) : 0.0;
    return fmin(m, 1.0);
}

void or_diff_area_f64(const double *lines, uint32_t start, uint32_t count, uint32_t width, uint32_t height,
                      int32_t *box, double *area) {
    DiffBox b = diff_box(lines, start, count, width, height);
    box[0] = b.x0;
    box[1] = b.y0;
    box[2] = b.w;
    box[3] = b.h;
    if (area && b.w && b.h) diff_area(lines, start, count, b, area);
}

void or_diff_render_f64(uint32_t width, uint32_t height, const double *clear, uint32_t draw_count,
                        const uint32_t *line_start, const uint32_t *line_count, const uint32_t *rule,
                        const double *colors, const double *lines, double *image) {
    for (size_t p = 0; p < (size_t)width * height; ++p)
        for (int k = 0; k < 4; ++k) image[4 * p + k] = clear[k];
    for (uint32_t d = 0; d < draw_count; ++d) {
        DiffBox b = diff_box(lines, line_start[d], line_count[d], width, height);
        if (!b.w || !b.h) continue;
        double *area = (double *)malloc((size_t)b.w * b.h * sizeof(double));
        diff_area(lines, line_start[d], line_count[d], b, area);
        const double *c = colors + 4 * (size_t)d;
        for (int y = 0; y < b.h; ++y)
            for (int x = 0; x < b.w; ++x) {
                double slope, alpha = diff_alpha64(area[(size_t)y * b.w + x], rule[d], &slope);
                double f = 1.0 - alpha * c[3];
                double *px = image + 4 * ((size_t)(b.y0 + y) * width + b.x0 + x);
                for (int k = 0; k < 4; ++k) px[k] = px[k] * f + alpha * c[k];
            }
        free(area);
    }
}

/* dL/d(line endpoints) of lines [start, start + count) from dL/dcoverage over
 * their draw's box b (gmap, b.w * b.h doubles): the coverage of a pixel moves by
 * -n.(dp0 int(1-s) + dp1 int s) over the part of a line inside it. */
static void diff_line_pass(const double *lines, uint32_t start, uint32_t count, DiffBox b, const double *gmap,
                           double *line_grad) {
    for (uint32_t i = start; i < start + count; ++i) {
        const double *l = lines + 4 * (size_t)i;
        double *out = line_grad + 4 * (size_t)i;
        double g0 = 0.0, g1 = 0.0;
        const double x0 = l[0], y0 = l[1], dx = l[2] - l[0], dy = l[3] - l[1];
        if (b.w && b.h) {
            int r0 = (int)floor(fmin(y0, l[3])), r1 = dy == 0.0 ? r0 : (int)ceil(fmax(y0, l[3])) - 1;
            if (r0 < b.y0) r0 = b.y0;
            if (r1 > b.y0 + b.h - 1) r1 = b.y0 + b.h - 1;
            for (int r = r0; r <= r1; ++r) {
                double sa = 0.0, sb = 1.0;
                if (dy != 0.0) {
                    double ta = ((double)r - y0) / dy, tb = ((double)(r + 1) - y0) / dy;
                    sa = fmax(fmin(ta, tb), 0.0);
                    sb = fmin(fmax(ta, tb), 1.0);
                }
                if (!(sb > sa)) continue;
                double xa = x0 + sa * dx, xb = x0 + sb * dx;
                int c0 = (int)floor(fmin(xa, xb)), c1 = dx == 0.0 ? c0 : (int)ceil(fmax(xa, xb)) - 1;
                if (c0 < b.x0) c0 = b.x0;
                if (c1 > b.x0 + b.w - 1) c1 = b.x0 + b.w - 1;
                for (int c = c0; c <= c1; ++c) {
                    double ua = sa, ub = sb;
                    if (dx != 0.0) {
                        double qa = ((double)c - x0) / dx, qb = ((double)(c + 1) - x0) / dx;
                        ua = fmax(fmin(qa, qb), sa);
                        ub = fmin(fmax(qa, qb), sb);
                    }
                    if (!(ub > ua)) continue;
                    double gv = gmap[(size_t)(r - b.y0) * b.w + (c - b.x0)];
                    double i1 = 0.5 * (ub * ub - ua * ua);
                    g0 += gv * ((ub - ua) - i1);
                    g1 += gv * i1;
                }
            }
        }
        double nx = y0 - l[3], ny = dx;
        out[0] = -nx * g0;
        out[1] = -ny * g0;
        out[2] = -nx * g1;
        out[3] = -ny * g1;
    }
}

/* dL/d(line endpoints) and dL/d(colour) for a given dL/d(image).  Per pixel,
 * with T_k the product of (1 - alpha_j c_j.a) over the draws above k and R the
 * composite below k: dL/dalpha_k = T_k (g.c_k - c_k.a g.R).  The line pass
 * integrates dL/dcoverage along each line: over the part of the line with
 * parameter in [ua, ub] inside a pixel, the pixel's coverage moves by
 * -n.(dp0 int(1-s) + dp1 int s), n = (p0.y - p1.y, p1.x - p0.x). */
void or_diff_backward_f64(uint32_t width, uint32_t height, const double *clear, uint32_t draw_count,
                          const uint32_t *line_start, const uint32_t *line_count, const uint32_t *rule,
                          const double *colors, const double *lines, const double *image_grad, double *line_grad,
                          double *color_grad) {
    const size_t N = (size_t)width * height;
    DiffBox *boxes = (DiffBox *)calloc(draw_count ? draw_count : 1, sizeof(DiffBox));
    double **maps = (double **)calloc(draw_count ? draw_count : 1, sizeof(double *));
    double **trans = (double **)calloc(draw_count ? draw_count : 1, sizeof(double *));
    for (uint32_t d = 0; d < draw_count; ++d) {
        boxes[d] = diff_box(lines, line_start[d], line_count[d], width, height);
        size_t n = (size_t)boxes[d].w * boxes[d].h;
        if (!n) continue;
        maps[d] = (double *)malloc(n * sizeof(double));
        trans[d] = (double *)malloc(n * sizeof(double));
        diff_area(lines, line_start[d], line_count[d], boxes[d], maps[d]);
    }
    /* transmittance above each draw, top to bottom */
    double *T = (double *)malloc((N ? N : 1) * sizeof(double));
    for (size_t p = 0; p < N; ++p) T[p] = 1.0;
    for (uint32_t d = draw_count; d-- > 0;) {
        DiffBox b = boxes[d];
        for (int y = 0; y < b.h; ++y)
            for (int x = 0; x < b.w; ++x) {
                size_t p = (size_t)(b.y0 + y) * width + b.x0 + x, q = (size_t)y * b.w + x;
                double slope, alpha = diff_alpha64(maps[d][q], rule[d], &slope);
                trans[d][q] = T[p];
                T[p] *= 1.0 - alpha * colors[4 * (size_t)d + 3];
            }
    }
    /* bottom to top: R, the colour gradients, and dL/dcoverage in place of the coverage */
    double *R = (double *)malloc((N ? N : 1) * 4 * sizeof(double));
    for (size_t p = 0; p < N; ++p)
        for (int k = 0; k < 4; ++k) R[4 * p + k] = clear[k];
    for (uint32_t d = 0; d < draw_count; ++d) {
        DiffBox b = boxes[d];
        const double *c = colors + 4 * (size_t)d;
        double *cg = color_grad + 4 * (size_t)d;
        cg[0] = cg[1] = cg[2] = cg[3] = 0.0;
        for (int y = 0; y < b.h; ++y)
            for (int x = 0; x < b.w; ++x) {
                size_t p = (size_t)(b.y0 + y) * width + b.x0 + x, q = (size_t)y * b.w + x;
                double slope, alpha = diff_alpha64(maps[d][q], rule[d], &slope);
                const double *g = image_grad + 4 * p;
                double *r = R + 4 * p;
                double t = trans[d][q];
                double g_r = g[0] * r[0] + g[1] * r[1] + g[2] * r[2] + g[3] * r[3];
                double g_c = g[0] * c[0] + g[1] * c[1] + g[2] * c[2] + g[3] * c[3];
                maps[d][q] = t * (g_c - c[3] * g_r) * slope;
                cg[0] += t * alpha * g[0];
                cg[1] += t * alpha * g[1];
                cg[2] += t * alpha * g[2];
                cg[3] += t * alpha * (g[3] - g_r);
                double f = 1.0 - alpha * c[3];
                for (int k = 0; k < 4; ++k) r[k] = r[k] * f + alpha * c[k];
            }
    }
    for (uint32_t d = 0; d < draw_count; ++d) diff_line_pass(lines, line_start[d], line_count[d], boxes[d], maps[d], line_grad);
    for (uint32_t d = 0; d < draw_count; ++d) {
        free(maps[d]);
        free(trans[d]);
    }
    free(maps);
    free(trans);
    free(boxes);
    free(T);
    free(R);
}

/* The clip mask one SDF clip draw gives each pixel centre of a width x height
 * canvas: its FP32 coverage as fine evaluates it (sdf_coverage_from_blob at the
 * draw's inverse transform), before the u8 step. */
void or_sdf_draw_coverage(const uint32_t *draws, uint32_t draw_count, uint32_t draw_index, const uint32_t *paint,
                          uint32_t paint_words, uint32_t paint_sdf_shadow_base, uint32_t width, uint32_t height,
                          float *out) {
    or_fine_config cfg;
    memset(&cfg, 0, sizeof(cfg));
    cfg.paint_words = paint_words;
    cfg.paint_sdf_shadow_base = paint_sdf_shadow_base;
    const FDraw d = load_fdraw(draws, draw_count, draw_index);
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x)
            out[(size_t)y * width + x] = draw_sdf_coverage(&cfg, paint, &d, (float)x + 0.5f, (float)y + 0.5f);
}

void or_diff_line_grad_f64(const double *lines, uint32_t start, uint32_t count, uint32_t width, uint32_t height,
                           const double *coverage_grad, double *line_grad) {
    DiffBox b = diff_box(lines, start, count, width, height);
    diff_line_pass(lines, start, count, b, coverage_grad, line_grad);
}
