// gfx950 kernels of the compute-raster pipeline.
//
// Built with: hipcc --genco --offload-arch=gfx950 -O3 -ffp-contract=off
//             -Xclang -target-feature -Xclang -packed-fp32-ops
// and loaded by the host through hipModuleLoadData.  Each kernel takes one
// argument block (tk_kernel_args.h) by value.
//
// Stage map (reference entry -> kernel):
//   scan_clear                       -> tk_k_scan_clear
//   scan_count                       -> tk_k_scan_count
//   scan_prefix_chunks + scan_chunk_offsets + scan_apply_chunk_offsets
//                                    -> tk_k_scan_prefix (single pass, decoupled lookback)
//   scan_emit                        -> tk_k_scan_emit (slot order canonicalised when fine stages a tile)
//   cumsum_*                         -> tk_k_cumsum
//   coarse_count_bins   (x batches)  -> tk_k_coarse_count  (all batches, one pass over each tile list)
//   coarse_prefix_*     (x batches)  -> tk_k_coarse_scan   (one decoupled-lookback pass over batch x tile)
//   coarse_emit_bins    (x batches)  -> tk_k_coarse_emit
//   fine_tile_main      (x batches)  -> tk_k_fine          (all batches per 16x16 tile workgroup)

#include <hip/hip_runtime.h>

#include "raster_device.h"
#include "tk_kernel_args.h"

using namespace tkd;

namespace {

// ---------------------------------------------------------------------------
// Block-wide scan and decoupled lookback
// ---------------------------------------------------------------------------

constexpr uint32_t kFlagAggregate = 1u;
constexpr uint32_t kFlagPrefix = 2u;

// Exclusive wrapping-u32 scan over a 256-thread block (four 64-lane waves).
__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t *s_wave, uint32_t &total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63u) s_wave[wave] = x;
    __syncthreads();
    uint32_t before = 0u, sum = 0u;
#pragma unroll
    for (uint32_t w = 0; w < 4u; ++w) {
        uint32_t s = s_wave[w];
        if (w < wave) before += s;
        sum += s;
    }
    total = sum;
    __syncthreads();
    return before + x - v;
}

// Exclusive wrapping-u32 scan across one 64-lane wave.
__device__ __forceinline__ uint32_t wave_exclusive_scan(uint32_t v, uint32_t lane, uint32_t &total) {
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
    return v;
}

// Workgroups are dealt to the eight XCDs round-robin in launch order.  This
// renumbers a launch of `grid` workgroups (a multiple of 8 * chunk) so that each
// XCD receives runs of `chunk` consecutive block indices, every eighth run: the
// cache lines that neighbouring blocks share (tile table words, a path's
// consecutive lines and their counters) are then fetched into one XCD's L2, and
// the runs stay interleaved across the XCDs so that spatially uneven work is
// still spread evenly.  chunk is a power of two, chosen at launch; 0 keeps the
// launch order.
__device__ __forceinline__ uint32_t xcd_chunked_block(uint32_t linear, uint32_t chunk) {
    if (chunk == 0u) return linear;
    const uint32_t x = linear & 7u, k = linear >> 3, shift = (uint32_t)__builtin_ctz(chunk);
    return ((((k >> shift) << 3) + x) << shift) + (k & (chunk - 1u));
}
__device__ __forceinline__ uint32_t xcd_grid(uint64_t blocks, uint32_t chunk) {
    const uint64_t unit = chunk ? 8u * (uint64_t)chunk : 1u;
    return (uint32_t)((blocks + unit - 1u) & ~(unit - 1u));
}

// Work queues of the persistent launches: eight counters, one per XCD.  The
// k-th item of XCD x is xcd_chunked_block(8k + x, run), the runs that the
// one-workgroup-per-item launch deals to that XCD, so each XCD walks its own
// runs in order and keeps their cache lines in its L2.  A caller whose XCD's
// items are used up takes items from the other XCDs' counters in turn (`qi`
// counts the counters it has left behind; it starts at 0 on its home XCD).
// Returns kNone once all eight are used up.  Every call advances a counter, so
// every caller reaches kNone after finitely many calls.
__device__ uint32_t queue_next(uint32_t *queue, uint32_t home, uint32_t &qi, uint32_t run, uint32_t items) {
    while (qi < 8u) {
        const uint32_t x = (home + qi) & 7u;
        const uint64_t linear = (uint64_t)atomicAdd(&queue[x], 1u) * 8u + x;
        if (linear <= 0xFFFFFFFFull) {
            const uint32_t item = xcd_chunked_block((uint32_t)linear, run);
            if (item < items) return item;
        }
        ++qi;
    }
    return ~0u;
}

__device__ __forceinline__ unsigned long long pack_status(uint32_t value, uint32_t flag, uint32_t epoch) {
    return (unsigned long long)value | ((unsigned long long)flag << 32) | ((unsigned long long)(epoch & 0x3FFFFFFFu) << 34);
}

__device__ __forceinline__ void publish(unsigned long long *status, uint32_t index, uint32_t value, uint32_t flag,
                                        uint32_t epoch) {
    __hip_atomic_store(&status[index], pack_status(value, flag, epoch), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of block k by decoupled lookback, run by one full wave: lane i
// reads the status word of block hi - i, so each step inspects 64 predecessors.
// The lanes up to the nearest inclusive prefix are summed once all of them have
// published; otherwise the published run nearest to k is consumed and the rest
// is read again.  Blocks before 0 count as an inclusive prefix of 0.
__device__ uint32_t wave_lookback(unsigned long long *status, uint32_t k, uint32_t epoch, uint32_t lane) {
    uint32_t acc = 0u;
    int64_t hi = (int64_t)k - 1;
    const uint32_t want = epoch & 0x3FFFFFFFu;
    while (hi >= 0) {
        const int64_t j = hi - (int64_t)lane;
        uint32_t flag = kFlagPrefix, value = 0u;
        if (j >= 0) {
            const unsigned long long st = __hip_atomic_load(&status[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            flag = ((uint32_t)(st >> 34) & 0x3FFFFFFFu) == want ? (uint32_t)(st >> 32) & 3u : 0u;
            value = (uint32_t)st;
        }
        const unsigned long long prefix_lanes = __ballot(flag == kFlagPrefix);
        const unsigned long long missing_lanes = __ballot(flag == 0u);
        const uint32_t stop = prefix_lanes ? (uint32_t)__builtin_ctzll(prefix_lanes) : 64u;
        const uint32_t gap = missing_lanes ? (uint32_t)__builtin_ctzll(missing_lanes) : 64u;
        if (stop < gap) {
            acc += wave_sum_u32(lane <= stop ? value : 0u);
            break;
        }
        acc += wave_sum_u32(lane < gap ? value : 0u);
        hi -= (int64_t)gap;
        if (gap < 64u) __builtin_amdgcn_s_sleep(1);
    }
    return acc;
}

// Dynamic block order: ticket k is only handed out after tickets < k, so a
// block only ever waits on blocks that are already resident.
__device__ __forceinline__ uint32_t draw_ticket(uint32_t *ticket, uint32_t block_count, uint32_t *s_ticket) {
    if (threadIdx.x == 0) {
        uint32_t t = atomicAdd(ticket, 1u);
        if (t == block_count - 1u) atomicExch(ticket, 0u);
        *s_ticket = t;
    }
    __syncthreads();
    return *s_ticket;
}

// ---------------------------------------------------------------------------
// Coarse helpers (shared/draw.hlsli, coarse/{classify,sdf_clip,paint,emit_draw,emit_stack}.hlsli)
// ---------------------------------------------------------------------------

struct DrawData {
    uint32_t path_id, glyph_run_id, sdf_offset, sdf_len, shadow_offset, tag, fill_rule, brush_offset, solid_rect;
    int32_t pb[4];
    float la, lb, lc, ld, te, tf;
};

__device__ __forceinline__ DrawData load_draw(const tk_draw_record *draws, uint32_t count, uint32_t index) {
    DrawData d{};
    if (index >= count) return d;  // robust read of an absent record: zeros
    const tk_draw_record &r = draws[index];
    d.path_id = r.path_id;
    d.glyph_run_id = r.glyph_run_id;
    d.sdf_offset = r.sdf_offset;
    d.sdf_len = r.sdf_len;
    d.shadow_offset = r.sdf_shadow_offset;
    d.brush_offset = r.brush_offset;
    d.solid_rect = r.solid_rect;
    d.tag = r.tag;
    d.fill_rule = r.fill_rule;
    d.pb[0] = r.pixel_bounds.x0;
    d.pb[1] = r.pixel_bounds.y0;
    d.pb[2] = r.pixel_bounds.x1;
    d.pb[3] = r.pixel_bounds.y1;
    d.la = r.transform.a;
    d.lb = r.transform.b;
    d.lc = r.transform.c;
    d.ld = r.transform.d;
    d.te = r.transform.e;
    d.tf = r.transform.f;
    return d;
}

__device__ __forceinline__ bool draw_has_sdf(const DrawData &d) { return d.sdf_offset != kNone || d.shadow_offset != kNone; }
__device__ __forceinline__ uint32_t tile_min_of(int32_t pixel, uint32_t limit) {
    return pixel > 0 ? min((uint32_t)pixel / 16u, limit) : 0u;
}
__device__ __forceinline__ uint32_t tile_max_of(int32_t pixel, uint32_t limit) {
    return pixel > 0 ? min(((uint32_t)pixel + 15u) / 16u, limit) : 0u;
}
__device__ __forceinline__ bool draw_hits_tile(const DrawData &d, uint32_t tx, uint32_t ty, uint32_t tw, uint32_t th) {
    uint32_t lx = tile_min_of(d.pb[0], tw), ly = tile_min_of(d.pb[1], th);
    uint32_t ux = tile_max_of(d.pb[2], tw), uy = tile_max_of(d.pb[3], th);
    return tx >= lx && ty >= ly && tx < ux && ty < uy;
}
__device__ __forceinline__ uint32_t draw_backdrop(const tk_path_record *paths, uint32_t path_count, const DrawData &d,
                                                  uint32_t tx, uint32_t ty, uint32_t tw, uint32_t th) {
    if (d.path_id == kNone || d.tag > TK_DRAW_PATH_GLYPH || d.path_id >= path_count || !draw_hits_tile(d, tx, ty, tw, th))
        return kNone;
    const tk_path_record &p = paths[d.path_id];
    if (p.tile_x1 <= p.tile_x0 || tx < p.tile_x0 || ty < p.tile_y0 || tx >= p.tile_x1 || ty >= p.tile_y1) return kNone;
    return p.data_offset + (ty - p.tile_y0) * (p.tile_x1 - p.tile_x0) + tx - p.tile_x0;
}

__device__ __forceinline__ bool rounded_corner_covers(float px, float py, float cx, float cy, float ox, float oy,
                                                      float radius, float inset) {
    bool in_square = radius > 0.0f && fabsf(px - cx) < radius && fabsf(py - cy) < radius;
    float inner = radius - inset;
    float dx = px - ox, dy = py - oy;
    return !in_square || (inner > 0.0f && dx * dx + dy * dy <= inner * inner);
}

__device__ bool rect_covers_tile(const float rect[4], const float radii_in[4], const DrawData &d, uint32_t tx,
                                 uint32_t ty, float inset) {
    float lx = fminf(rect[0], rect[2]), ly = fminf(rect[1], rect[3]);
    float ux = fmaxf(rect[0], rect[2]), uy = fmaxf(rect[1], rect[3]);
    float tlx = (float)(tx * 16u) + 0.5f - d.te, tly = (float)(ty * 16u) + 0.5f - d.tf;
    float tux = tlx + 15.0f, tuy = tly + 15.0f;
    float sx = ux - lx, sy = uy - ly;
    float need = 15.0f + 2.0f * inset;
    if (sx < need || sy < need || tlx < lx + inset || tly < ly + inset || tux > ux - inset || tuy > uy - inset)
        return false;
    float limit = fminf(sx, sy) * 0.5f;
    for (int i = 0; i < 4; ++i)
        if (!(fabsf(rect[i]) <= 65536.0f) || !(fabsf(radii_in[i]) <= 65536.0f)) return false;
    if (!(fabsf(d.te) <= 65536.0f) || !(fabsf(d.tf) <= 65536.0f)) return false;
    float r[4];
    for (int i = 0; i < 4; ++i) r[i] = fminf(fmaxf(radii_in[i], 0.0f), limit);
    return rounded_corner_covers(tlx, tly, lx, ly, lx + r[0], ly + r[0], r[0], inset) &&
           rounded_corner_covers(tux, tly, ux, ly, ux - r[1], ly + r[1], r[1], inset) &&
           rounded_corner_covers(tlx, tuy, lx, uy, lx + r[2], uy - r[2], r[2], inset) &&
           rounded_corner_covers(tux, tuy, ux, uy, ux - r[3], uy - r[3], r[3], inset);
}

__device__ __forceinline__ bool identity_linear(const DrawData &d) {
    return d.la == 1.0f && d.lb == 0.0f && d.lc == 0.0f && d.ld == 1.0f;
}

__device__ bool sdf_clip_covers(const uint32_t *paint, uint32_t pc, const DrawData &d, uint32_t tx, uint32_t ty) {
    if (d.sdf_offset == kNone || d.shadow_offset != kNone || d.sdf_len < 9u) return false;
    uint32_t base = d.sdf_offset;
    if (paint_word(paint, pc, base) != TK_SDF_KIND_RECT || !identity_linear(d)) return false;
    float rect[4], radii[4];
    for (int i = 0; i < 4; ++i) {
        rect[i] = asfloat(paint_word(paint, pc, base + 1u + (uint32_t)i));
        radii[i] = asfloat(paint_word(paint, pc, base + 5u + (uint32_t)i));
    }
    return rect_covers_tile(rect, radii, d, tx, ty, 0.5f);
}

__device__ bool sdf_stroke_empty(const uint32_t *paint, uint32_t pc, const DrawData &d, uint32_t tx, uint32_t ty) {
    if (d.sdf_offset == kNone || d.shadow_offset != kNone || d.sdf_len < 13u || !identity_linear(d)) return false;
    uint32_t base = d.sdf_offset;
    if (paint_word(paint, pc, base) != 3u) return false;  // SDF_KIND_RECT_STROKE
    float rect[4], radii[4], stroke[4];
    for (int i = 0; i < 4; ++i) {
        rect[i] = asfloat(paint_word(paint, pc, base + 1u + (uint32_t)i));
        radii[i] = asfloat(paint_word(paint, pc, base + 5u + (uint32_t)i));
        stroke[i] = asfloat(paint_word(paint, pc, base + 9u + (uint32_t)i));
    }
    for (int i = 0; i < 4; ++i)
        if (!(fabsf(rect[i]) <= 65536.0f) || !(fabsf(radii[i]) <= 65536.0f) || !(fabsf(stroke[i]) <= 65536.0f))
            return false;
    float hw[4];
    for (int i = 0; i < 4; ++i) hw[i] = fmaxf(stroke[i], 0.0f);
    float lx = fminf(rect[0], rect[2]), ly = fminf(rect[1], rect[3]);
    float ux = fmaxf(rect[0], rect[2]), uy = fmaxf(rect[1], rect[3]);
    float inner[4] = {lx + hw[3], ly + hw[0], ux - hw[1], uy - hw[2]};
    if (inner[0] >= inner[2] || inner[1] >= inner[3]) return false;
    float corner[4] = {fmaxf(hw[0], hw[3]), fmaxf(hw[0], hw[1]), fmaxf(hw[2], hw[3]), fmaxf(hw[2], hw[1])};
    float r[4];
    for (int i = 0; i < 4; ++i) r[i] = fmaxf(radii[i] - corner[i], 0.0f);
    return rect_covers_tile(inner, r, d, tx, ty, 1.0f);
}

__device__ __forceinline__ uint32_t brush_base_of(uint32_t paint_brush_base, const DrawData &d) {
    return paint_brush_base + d.brush_offset;
}
__device__ __forceinline__ bool nontransparent_solid(const uint32_t *paint, uint32_t pc, uint32_t pbb, const DrawData &d) {
    if (d.brush_offset == kNone) return false;
    uint32_t base = brush_base_of(pbb, d);
    return paint_word(paint, pc, base) == TK_BRUSH_SOLID && paint_word(paint, pc, base + 4u) != 0u;
}
__device__ __forceinline__ uint32_t solid_color(const uint32_t *paint, uint32_t pc, uint32_t pbb, const DrawData &d) {
    return paint_word(paint, pc, brush_base_of(pbb, d) + 4u);
}
__device__ bool solid_supported_sdf(const uint32_t *paint, uint32_t pc, uint32_t pbb, const DrawData &d) {
    if (!nontransparent_solid(paint, pc, pbb, d) || d.sdf_offset == kNone || d.shadow_offset != kNone || d.sdf_len == 0u)
        return false;
    uint32_t kind = paint_word(paint, pc, d.sdf_offset);
    return kind == TK_SDF_KIND_RECT || kind == 5u || kind == 14u;
}
__device__ __forceinline__ uint32_t full_tile_solid(const uint32_t *paint, uint32_t pc, uint32_t pbb, const DrawData &d,
                                                    uint32_t tx, uint32_t ty) {
    return nontransparent_solid(paint, pc, pbb, d) && sdf_clip_covers(paint, pc, d, tx, ty) ? solid_color(paint, pc, pbb, d)
                                                                                            : 0u;
}
__device__ __forceinline__ bool full_tile_image(const uint32_t *paint, uint32_t pc, uint32_t pbb, const DrawData &d,
                                                uint32_t tx, uint32_t ty) {
    if (d.brush_offset == kNone) return false;
    uint32_t base = brush_base_of(pbb, d);
    return paint_word(paint, pc, base) == 7u && paint_word(paint, pc, base + 7u) == 255u &&
           sdf_clip_covers(paint, pc, d, tx, ty);
}

constexpr uint32_t kClassColor = 1u, kClassSdf = 2u, kClassOther = 4u;
constexpr uint32_t kPtclImage = 13u, kPtclPathGlyph = 11u;
constexpr uint32_t kPtclBeginOpacity = 5u, kPtclEndOpacity = 6u, kPtclBeginBlend = 7u, kPtclEndBlend = 8u;

__device__ __forceinline__ uint32_t classify_tile_flags(uint32_t flags) {
    if (flags & kClassOther) return TK_TILE_KIND_INTERPRETER;
    if (flags & kClassSdf) return (flags & kClassColor) ? TK_TILE_KIND_MIXED : TK_TILE_KIND_SDF;
    return (flags & kClassColor) ? TK_TILE_KIND_COLOR : TK_TILE_KIND_EMPTY;
}

struct CoarseCtx {
    const tk_coarse_config *cfg;
    const tk_coarse_inputs *in;
    uint32_t tx, ty;
    const uint8_t *class_row;      // this tile's layer-pair classes, or null
    const uint32_t *class_pair_of; // per layer record: its pair
    uint32_t tw, th;               // the tile grid of this tile's canvas
};

// Where a table column lives: its canvas's tile grid, its batches [q0, q1), and
// its table entries base + (batch - q0) * cols + local.  One scene: the whole
// table, batch-major over `columns` columns.
struct SceneView {
    uint32_t tw, th, width, height, clear, q0, q1, cols, local;
    uint64_t base, image;
};
template <bool kScenes>
__device__ __forceinline__ SceneView scene_view(const tk_scene_slots &sl, uint32_t column, uint32_t columns,
                                                uint32_t batch_count, uint32_t tw, uint32_t th, uint32_t width,
                                                uint32_t height, uint32_t clear) {
    SceneView v{tw, th, width, height, clear, 0u, batch_count, columns, column, 0ull, 0ull};
    if constexpr (kScenes) {
        const tk_scene_slot &s = sl.slots[sl.scene_of_column[column]];
        v.tw = s.tiles_width;
        v.th = s.tiles_height;
        v.width = s.width;
        v.height = s.height;
        v.clear = s.clear_color;
        v.q0 = s.batch_start;
        v.q1 = s.batch_end;
        v.cols = s.tiles_width * s.tiles_height;
        v.local = column - s.column_start;
        v.base = s.table_base;
        v.image = s.image_offset;
    }
    return v;
}
__device__ __forceinline__ size_t table_index(const SceneView &v, uint32_t batch) {
    return (size_t)(v.base + (uint64_t)(batch - v.q0) * v.cols + v.local);
}

struct Particle {
    bool valid;
    uint32_t tag, winding, fill_rule, seg_start, seg_end, color;
};

// draw_particle (coarse/emit_draw.hlsli:10-52); draw_particle_count is its validity.
__device__ Particle draw_particle(const CoarseCtx &c, uint32_t batch, uint32_t draw_index) {
    const tk_coarse_config &cfg = *c.cfg;
    const tk_coarse_inputs &in = *c.in;
    Particle p{false, TK_PTCL_FILL, 0u, 0u, 0u, 0u, 0u};
    if (draw_index >= cfg.draw_count || in.draw_batch_ids[draw_index] != batch) return p;
    DrawData d = load_draw(in.draws, cfg.draw_count, draw_index);
    if (draw_has_sdf(d)) {
        if (d.tag == TK_DRAW_BRUSH && !sdf_stroke_empty(in.paint, cfg.paint_words, d, c.tx, c.ty)) {
            p.valid = true;
            uint32_t color = full_tile_solid(in.paint, cfg.paint_words, cfg.paint_brush_base, d, c.tx, c.ty);
            if (color != 0u) {
                p.tag = TK_PTCL_COLOR;
                p.color = color;
            } else if (full_tile_image(in.paint, cfg.paint_words, cfg.paint_brush_base, d, c.tx, c.ty)) {
                p.tag = kPtclImage;
                p.color = draw_index;
            } else {
                p.tag = TK_PTCL_SDF;
                p.seg_start = draw_index;
                p.color = draw_index;
            }
        }
        return p;
    }
    uint32_t bd = draw_backdrop(in.paths, cfg.path_count, d, c.tx, c.ty, c.tw, c.th);
    if (bd == kNone || (d.tag != TK_DRAW_BRUSH && d.tag != TK_DRAW_PATH_GLYPH && d.tag != TK_DRAW_CLIP)) return p;
    tk_segment_range range = in.segment_ranges[bd];
    uint32_t winding = (uint32_t)in.backdrops[bd];
    if (range.start == range.end && winding == 0u) return p;
    p.valid = true;
    p.winding = winding;
    p.fill_rule = d.fill_rule;
    p.seg_start = range.start;
    p.seg_end = range.end;
    if (d.tag == TK_DRAW_CLIP) p.tag = TK_PTCL_BEGIN_CLIP;
    else if (d.tag == TK_DRAW_PATH_GLYPH) p.tag = kPtclPathGlyph;
    else if (d.solid_rect != 0u && d.lb == 0.0f && d.lc == 0.0f &&
             nontransparent_solid(in.paint, cfg.paint_words, cfg.paint_brush_base, d) && range.start == range.end)
        p.tag = TK_PTCL_COLOR;
    if (p.tag == TK_PTCL_COLOR) p.color = solid_color(in.paint, cfg.paint_words, cfg.paint_brush_base, d);
    else if (d.tag == TK_DRAW_BRUSH || d.tag == TK_DRAW_PATH_GLYPH) p.color = draw_index;
    return p;
}

__device__ __forceinline__ uint32_t particle_class_flags(const CoarseCtx &c, uint32_t draw_index, uint32_t tag) {
    if (tag == TK_PTCL_COLOR) return kClassColor;
    if (tag == kPtclImage) return kClassSdf;
    if (tag == TK_PTCL_SDF) {
        DrawData d = load_draw(c.in->draws, c.cfg->draw_count, draw_index);
        if (solid_supported_sdf(c.in->paint, c.cfg->paint_words, c.cfg->paint_brush_base, d)) return kClassSdf;
    }
    return kClassOther;
}

// noop_stack_clip (coarse/emit_stack.hlsli:9-17).
__device__ bool noop_stack_clip(const CoarseCtx &c, const tk_layer_record &layer, const DrawData &d) {
    if (layer.tag != TK_LAYER_CLIP) return false;
    if (draw_has_sdf(d)) return sdf_clip_covers(c.in->paint, c.cfg->paint_words, d, c.tx, c.ty);
    uint32_t bd = draw_backdrop(c.in->paths, c.cfg->path_count, d, c.tx, c.ty, c.tw, c.th);
    if (bd == kNone) return false;  // excluded by stack_wrapper_count before any emission
    tk_segment_range range = c.in->segment_ranges[bd];
    uint32_t winding = (uint32_t)c.in->backdrops[bd];
    return range.start == range.end && (d.fill_rule == TK_FILL_EVEN_ODD ? (winding & 1u) != 0u : winding != 0u);
}

// A layer's verdict on this tile, as stack_wrapper_count (coarse/classify.hlsli:
// 11-33) and emit_stack_begins (coarse/emit_stack.hlsli:18-31) decide it:
// kLayerInvalid rejects the tile for the whole batch, kLayerCounted counts a
// wrapper, kLayerEmits writes BEGIN/END records, kLayerSdf marks an SDF clip.
constexpr uint32_t kLayerInvalid = 1u, kLayerCounted = 2u, kLayerEmits = 4u, kLayerSdf = 8u;
__device__ uint32_t layer_class_bits(const CoarseCtx &c, const tk_layer_record &layer) {
    const tk_coarse_config &cfg = *c.cfg;
    const tk_coarse_inputs &in = *c.in;
    const DrawData d = load_draw(in.draws, cfg.draw_count, layer.draw);
    uint32_t bits = 0u;
    if (draw_has_sdf(d)) {
        const bool covers = sdf_clip_covers(in.paint, cfg.paint_words, d, c.tx, c.ty);
        if (!covers) bits |= draw_hits_tile(d, c.tx, c.ty, c.tw, c.th) ? kLayerCounted : kLayerInvalid;
        if (!(layer.tag == TK_LAYER_CLIP && covers)) bits |= kLayerEmits;
        return bits | kLayerSdf;
    }
    const uint32_t bd = draw_backdrop(in.paths, cfg.path_count, d, c.tx, c.ty, c.tw, c.th);
    if (bd == kNone) return kLayerInvalid;
    const tk_segment_range range = in.segment_ranges[bd];
    const uint32_t winding = (uint32_t)in.backdrops[bd];
    const bool empty = range.start == range.end;
    const bool full = d.fill_rule == TK_FILL_EVEN_ODD ? (winding & 1u) != 0u : winding != 0u;
    if (layer.tag == TK_LAYER_CLIP && empty && full) return 0u;  // a no-op clip
    return kLayerEmits | ((empty && winding == 0u) ? kLayerInvalid : kLayerCounted);
}

// The verdict of layer record i: from the per-tile class table when the frame
// built one, else evaluated here.
__device__ __forceinline__ uint32_t layer_bits_at(const CoarseCtx &c, uint32_t i, const tk_layer_record &layer) {
    if (c.class_row) return c.class_row[c.class_pair_of[i]];
    return layer_class_bits(c, layer);
}

// The coarse view of table column `column`, whose tile is `tile` of its canvas.
__device__ __forceinline__ CoarseCtx coarse_ctx(const tk_args_coarse &a, const SceneView &v, uint32_t tile,
                                                uint32_t column) {
    CoarseCtx c{&a.cfg, &a.in, tile % v.tw, tile / v.tw, nullptr, nullptr, v.tw, v.th};
    if (a.classes.bits) {
        c.class_row = a.classes.bits + (size_t)column * a.classes.pair_count;
        c.class_pair_of = a.classes.pair_of;
    }
    return c;
}

struct Emitter {
    tk_ptcl_record *ptcl;
    uint64_t capacity;
    uint32_t *status;
    __device__ void store(uint32_t dst, uint32_t tag, uint32_t winding, uint32_t fill_rule, uint32_t s0, uint32_t s1,
                          uint32_t color) {
        if ((uint64_t)dst >= capacity) {
            atomicOr(&status[0], 1u);
            return;
        }
        tk_ptcl_record r;
        r.tag = tag;
        r.backdrop = (int32_t)winding;
        r.fill_rule = fill_rule;
        r.segment_start = s0;
        r.segment_end = s1;
        r.color = color;
        ptcl[dst] = r;
    }
};

__device__ __forceinline__ bool is_begin_tag(uint32_t tag) {
    return tag == TK_PTCL_BEGIN_CLIP || tag == TK_PTCL_BEGIN_SDF_CLIP || tag == kPtclBeginOpacity || tag == kPtclBeginBlend;
}

// ---------------------------------------------------------------------------
// Fine helpers (fine/{interpreter,specialized,clip,draw,records}.hlsli)
// ---------------------------------------------------------------------------

struct FineDraw {
    uint32_t brush_offset, sdf_offset, shadow_offset;
    int32_t pb[4];
    Affine inv;
};

__device__ __forceinline__ FineDraw load_fine_draw(const tk_draw_record *draws, uint32_t count, uint32_t index) {
    FineDraw f{};
    if (index >= count) return f;
    const tk_draw_record &r = draws[index];
    f.brush_offset = r.brush_offset;
    f.sdf_offset = r.sdf_offset;
    f.shadow_offset = r.sdf_shadow_offset;
    f.pb[0] = r.pixel_bounds.x0;
    f.pb[1] = r.pixel_bounds.y0;
    f.pb[2] = r.pixel_bounds.x1;
    f.pb[3] = r.pixel_bounds.y1;
    f.inv = Affine{r.inverse_transform.a, r.inverse_transform.b, r.inverse_transform.c,
                   r.inverse_transform.d, r.inverse_transform.e, r.inverse_transform.f};
    return f;
}

__device__ __forceinline__ uint32_t sample_draw_brush(const tk_fine_config &cfg, const uint32_t *paint,
                                                      const FineDraw &d, float x, float y, const AtlasView &atlas) {
    if (blob_word(paint, cfg.paint_words, cfg.paint_brush_base, d.brush_offset) == TK_BRUSH_SOLID)
        return blob_word(paint, cfg.paint_words, cfg.paint_brush_base, d.brush_offset + 4u);
    float lx, ly;
    affine_point(d.inv, x, y, lx, ly);
    return sample_brush(paint, cfg.paint_words, cfg.paint_brush_base, d.brush_offset, lx, ly, atlas);
}
__device__ __forceinline__ AtlasView atlas_of(const tk_fine_inputs &in) {
    return AtlasView{in.atlas, in.atlas_width, in.atlas_height, in.atlas_pages};
}

__device__ __forceinline__ float sdf_coverage_from_draw(const tk_fine_config &cfg, const uint32_t *paint,
                                                        const FineDraw &d, float x, float y) {
    float lx, ly;
    affine_point(d.inv, x, y, lx, ly);
    if (d.sdf_offset != kNone) return sdf_coverage_from_blob(paint, cfg.paint_words, d.sdf_offset, lx, ly, d.inv);
    if (d.shadow_offset != kNone)
        return sdf_coverage_from_blob(paint, cfg.paint_words, cfg.paint_sdf_shadow_base + d.shadow_offset, lx, ly, d.inv);
    return 0.0f;
}

__device__ __forceinline__ bool premul_opaque(uint32_t c) { return (c >> 24) == 255u; }

// The text composite out of line: only path-glyph records reach it, and inlined
// into fine's record loop its temporaries cost the loop a spilled register.
__device__ __attribute__((noinline)) uint32_t glyph_composite(uint32_t destination, uint32_t source, uint32_t coverage) {
    return src_over_mask_linear_auto_u8(destination, source, coverage);
}

// Exact coverage sums of one FILL / BEGIN_CLIP record over a 16x16 tile, built
// segment by segment instead of pixel by pixel, with 64-bit integer atomics in
// LDS (exact, so the order of arrival does not matter):
//   base[row]     += the row's edge term,
//   run[row][fs]  += dy, where fs = full_start (prefix-summed along the row later),
//   part[row][x]  += area(x) * dy for the straddled pixels x in [partial_start, fs).
struct CoverageAcc {
    long long base[16];
    long long run[256];
    long long part[256];
};

// All terms of segments [0, n) of one record: lane l takes (segment l / 16,
// row l % 16), evaluates segment_row_parts for that row once, and adds its terms.
__device__ __forceinline__ void accumulate_record(const tk_line_segment *segs, uint32_t n, CoverageAcc &acc,
                                                  uint32_t lane) {
    const uint32_t row = lane & 15u;
    // A 32-bit index (n is 32-bit): no 64-bit counter compare and step per pass.
    for (uint32_t i = lane >> 4; i < n; i += 16u) {
        const tk_line_segment seg = segs[i];
        const RowParts p = segment_row_parts(seg.p0x, seg.p0y, seg.p1x, seg.p1y, seg.y_edge, row);
        if (p.edge != 0.0f) atomicAdd((unsigned long long *)&acc.base[row], (unsigned long long)coverage_fixed(p.edge));
        if (p.dy != 0.0f) {
            const int fs = clamp_cell16(ceilf(p.xmax));
            const int ps = clamp_cell16(floorf(p.xmin));
            if (fs < 16) atomicAdd((unsigned long long *)&acc.run[row * 16u + fs], (unsigned long long)coverage_fixed(p.dy));
            float xf = (float)ps;  // exact: 0..16
            for (int x = ps; x < fs; ++x, xf += 1.0f)
                atomicAdd((unsigned long long *)&acc.part[row * 16u + x],
                          (unsigned long long)coverage_fixed(segment_area_at_f(p.xmin, p.xmax, xf) * p.dy));
        }
    }
}

// Inclusive prefix sum over each 16-lane DPP row (one pixel row of the tile).
__device__ __forceinline__ long long row16_inclusive_sum(long long v) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)((unsigned long long)v >> 32);
#define TK_ROW_SHR_STEP(ctrl)                                                                   \
    {                                                                                           \
        uint32_t slo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, ctrl, 0xf, 0xf, true); \
        uint32_t shi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, ctrl, 0xf, 0xf, true); \
        uint32_t nlo = lo + slo;                                                                \
        hi = hi + shi + (nlo < lo ? 1u : 0u);                                                   \
        lo = nlo;                                                                               \
    }
    TK_ROW_SHR_STEP(0x111)  // row_shr:1
    TK_ROW_SHR_STEP(0x112)  // row_shr:2
    TK_ROW_SHR_STEP(0x114)  // row_shr:4
    TK_ROW_SHR_STEP(0x118)  // row_shr:8
#undef TK_ROW_SHR_STEP
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// Reads this lane's pixel (x = lane & 15, y = lane >> 4) out of `acc` and clears
// the entries it owns.  All 256 lanes call it together, after the barrier that
// follows accumulate_record; the 16 lanes of a pixel row share one wave, so the
// row's base word is read by all of them before lane x = 0 clears it.
__device__ __forceinline__ float take_coverage(CoverageAcc &acc, int32_t backdrop, uint32_t lane) {
    const uint32_t row = lane >> 4;
    const long long base = acc.base[row] + backdrop_fixed(backdrop);
    const long long running = row16_inclusive_sum(acc.run[lane]);
    const long long partial = acc.part[lane];
    acc.run[lane] = 0;
    acc.part[lane] = 0;
    if ((lane & 15u) == 0u) acc.base[row] = 0;
    return fixed_to_float(base + running + partial);
}

}  // namespace

// ===========================================================================
// Kernels
// ===========================================================================

extern "C" __global__ __launch_bounds__(256) void tk_k_fill(tk_args_fill a) {
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (; i < a.count; i += stride) a.dst[i] = a.value;
}

// Plain device copy, 16 bytes per lane per step (traffic calibration).
extern "C" __global__ __launch_bounds__(256) void tk_k_copy(tk_args_copy a) {
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (; i < a.count; i += stride) a.dst[i] = a.src[i];
}

// scan_clear (scan/clear.hlsl:16-32), non-incremental.
extern "C" __global__ __launch_bounds__(256) void tk_k_scan_clear(tk_args_scan_clear a) {
    uint32_t index = blockIdx.x * 256u + threadIdx.x;
    if (index >= a.cfg.clear_len) return;
    if (index < a.cfg.backdrop_len) {
        if (a.backdrops) a.backdrops[index] = 0;
        if (a.ranges) a.ranges[index] = tk_segment_range{0u, 0u};
        if (a.counts) a.counts[index] = 0u;
        if (a.cursors) a.cursors[index] = 0u;
    }
    if (index < a.cfg.path_count && a.bumps) a.bumps[index] = 0u;
    if (index < a.cfg.scan_chunk_count) {
        if (a.chunk_totals) a.chunk_totals[index] = 0u;
        if (a.chunk_offsets) a.chunk_offsets[index] = 0u;
    }
}

// Run-length aggregation over the active lanes: a run is a maximal stretch of
// consecutive active lanes with the same index (consecutive lines of a path
// usually share their tile).  Returns the run's first lane and its length
// (meaningful in that first lane); constant cost however many distinct indices
// the wave holds.  Called where every lane that reaches it takes part (any subset
// of the wave may be active); indices repeated in lanes that are not adjacent
// form separate runs, each with its own atomic.
struct LaneRun {
    uint32_t head, length;
};
__device__ __forceinline__ LaneRun lane_run(uint32_t index) {
    const uint32_t lane = __lane_id();
    const unsigned long long active = __ballot(1);
    const uint32_t prev = (uint32_t)__shfl_up((int)index, 1, 64);
    const bool head = lane == 0u || ((active >> (lane - 1u)) & 1ull) == 0ull || prev != index;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
    LaneRun r;
    r.head = 63u - (uint32_t)__clzll(heads & upto);
    const unsigned long long stops = (heads | ~active) & ~upto;  // lanes after this one that end the run
    r.length = (stops ? (uint32_t)__builtin_ctzll(stops) : 64u) - lane;
    return r;
}

// Adds +1 or -1 per lane to words[index], one atomic per run of equal indices.
__device__ __forceinline__ void wave_add_pm1_runs(uint32_t *words, uint32_t index, bool plus) {
    const uint32_t lane = __lane_id();
    const LaneRun r = lane_run(index);
    const unsigned long long ups = __ballot(plus);
    if (r.head == lane) {
        const unsigned long long span = (r.length >= 64u ? ~0ull : ((1ull << r.length) - 1ull)) << lane;
        const uint32_t n_up = (uint32_t)__popcll(ups & span);
        atomicAdd(&words[index], n_up - (r.length - n_up));
    }
}

// Incremental scan: a path takes part when its tile box meets a damaged tile,
// by one summed-area-table query over the damage bitmap.
extern "C" __global__ __launch_bounds__(256) void tk_k_active_paths(tk_args_active_paths a) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= a.path_count) return;
    const tk_path_record &r = a.paths[p];
    const uint32_t x0 = min(r.tile_x0, a.tiles_width), x1 = min(r.tile_x1, a.tiles_width);
    const uint32_t y0 = min(r.tile_y0, a.tiles_height), y1 = min(r.tile_y1, a.tiles_height);
    uint32_t active = 0u;
    if (x0 < x1 && y0 < y1) {
        const uint32_t w = a.tiles_width + 1u;
        const uint32_t *t = a.damage_sat;
        active = (t[y1 * w + x1] - t[y0 * w + x1] - t[y1 * w + x0] + t[y0 * w + x0]) != 0u ? 1u : 0u;
    }
    a.path_active[p] = active;
}

// scan_count (scan/count.hlsl:13-40).
extern "C" __global__ __launch_bounds__(1024) void tk_k_scan_count(tk_args_scan_lines a) {
    const uint32_t block = xcd_chunked_block(blockIdx.x, a.xcd_run);
    uint32_t id = block * blockDim.x + threadIdx.x;  // lines per workgroup: a launch parameter
    if (id >= a.cfg.line_count) return;
    if (a.path_active) {  // incremental: only paths whose tile box meets the damage
        const uint32_t path = a.lines[id].path_id;
        if (path < a.path_record_count && a.path_active[path] == 0u) return;
    }
    ScanTraversal s;
    if (!scan_geometry(a.lines, a.paths, a.path_record_count, id, s)) return;
    uint32_t *backdrops = reinterpret_cast<uint32_t *>(a.backdrops);
    const bool up = !s.down;  // delta = down ? -1 : +1
    uint32_t stride = s.bz - s.bx;
    for (int32_t y = s.ymin; y < s.ymax; y++) {
        uint32_t local = (uint32_t)(y - (int32_t)s.by) * stride;
        wave_add_pm1_runs(backdrops, s.data_offset + local, up);
    }
    float last_z = floorf(s.a * ((float)s.imin - 1.0f) + s.b);
    for (uint32_t i = s.imin; i < s.imax; i++) {
        float z = floorf(s.a * (float)i + s.b);
        int32_t tx, ty;
        scan_tile(s, i, z, tx, ty);
        if (scan_tile_inside(s, tx, ty)) {
            bool top = last_z == z;
            if (i == 0u) top = fabsf(s.y0 - s.py0 * (1.0f / 16.0f)) <= kScanEpsilon;
            if (top && tx + 1 < (int32_t)s.bz) {
                int32_t x = max(tx + 1, (int32_t)s.bx);
                uint32_t local = (uint32_t)(ty - (int32_t)s.by) * stride + (uint32_t)(x - (int32_t)s.bx);
                wave_add_pm1_runs(backdrops, s.data_offset + local, up);
            }
            wave_add_pm1_runs(a.counts_or_cursors, scan_tile_offset(s, tx, ty), true);
        }
        last_z = z;
    }
}


// Shared by both scan passes: the chunk's local scan (four tiles per lane) and
// whether the offsets pass of the chunk's path covers it.
struct ScanChunkView {
    tk_scan_chunk chunk;
    uint32_t len;            // tiles this chunk's lanes touch: min(chunk.len, 256)
    uint32_t v[4];           // counts of tiles lane*4 .. lane*4+3
    uint32_t excl, total;    // lane-exclusive prefix and chunk total
    uint32_t path;           // chunk.path_id
    tk_scan_chunk_range r;   // the path's chunk range (empty when the path id is invalid)
    bool in_range;           // r.start <= k < r.end
    uint32_t seg_start;      // the path's segment_start
};

__device__ __forceinline__ void load_scan_chunk(const tk_args_scan_prefix &a, uint32_t k, uint32_t lane, ScanChunkView &c) {
    c.chunk = a.chunks[k];
    c.len = min(c.chunk.len, 256u);
    uint32_t mine = 0u;
#pragma unroll
    for (uint32_t i = 0; i < 4u; ++i) {
        const uint32_t e = lane * 4u + i;
        c.v[i] = e < c.len ? a.counts[c.chunk.backdrop_offset + e] : 0u;
        mine += c.v[i];
    }
    c.excl = wave_exclusive_scan(mine, lane, c.total);
    c.path = c.chunk.path_id;
    c.r = tk_scan_chunk_range{0u, 0u};
    c.seg_start = 0u;
    if (c.path < a.cfg.path_count) {
        c.r = a.chunk_ranges[c.path];
        c.seg_start = a.paths[c.path].segment_start;
    }
    c.in_range = k >= c.r.start && k < c.r.end;
}

// scan_apply_chunk_offsets for one chunk whose offset is `base`.
__device__ __forceinline__ void apply_scan_chunk(const tk_args_scan_prefix &a, const ScanChunkView &c, uint32_t lane,
                                                 uint32_t base) {
    uint32_t running = base + c.excl;
#pragma unroll
    for (uint32_t i = 0; i < 4u; ++i) {
        const uint32_t e = lane * 4u + i;
        if (e < c.len) {
            a.ranges[c.chunk.backdrop_offset + e] = tk_segment_range{running, running + c.v[i]};
            a.cursors[c.chunk.backdrop_offset + e] = running;
        }
        running += c.v[i];
    }
}

// The chunk total as scan_prefix_chunks stores it: only a chunk of 1..256 tiles
// writes its word (prefix_chunks.hlsl:37); any other keeps the buffer's value.
__device__ __forceinline__ bool scan_chunk_writes_total(const ScanChunkView &c) {
    return c.chunk.len >= 1u && c.chunk.len <= 256u;
}

// Pass A of scan_prefix_chunks + scan_chunk_offsets + scan_apply_chunk_offsets
// (scan/prefix_chunks.hlsl:16-38, chunk_offsets.hlsl:14-26,
// apply_chunk_offsets.hlsl:15-27).  One 64-lane wave per chunk.  Each path's
// offsets form an independent chain over its own chunks, and almost every path
// is a single chunk, so this pass finishes every chunk whose offset needs no
// other chunk (a path's first chunk, or a chunk outside its path's range) and
// stores every chunk total; tk_k_scan_chunk_apply finishes the rest.  No wave
// ever waits on another.
extern "C" __global__ __launch_bounds__(256) void tk_k_scan_prefix(tk_args_scan_prefix a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (k >= a.wave_count) return;

    // Paths without chunks still receive a zero bump (scan_chunk_offsets runs per path).
    const uint32_t p = k * 64u + lane;
    if (p < a.cfg.path_count) {
        const tk_scan_chunk_range r = a.chunk_ranges[p];
        if (r.start >= r.end) a.bumps[p] = 0u;
    }
    if (k >= a.cfg.scan_chunk_count) return;

    ScanChunkView c;
    load_scan_chunk(a, k, lane, c);
    if (c.in_range && k != c.r.start) {
        if (lane == 0u && scan_chunk_writes_total(c)) a.chunk_totals[k] = c.total;
        return;  // offset depends on earlier chunks of the path: pass B
    }
    uint32_t base = 0u;
    if (lane == 0u) {
        uint32_t stored_total = c.total;
        if (scan_chunk_writes_total(c)) a.chunk_totals[k] = c.total;
        else stored_total = a.chunk_totals[k];
        if (c.in_range) {
            base = c.seg_start;
            a.chunk_offsets[k] = base;
            if (k + 1u == c.r.end) a.bumps[c.path] = stored_total;
        } else {
            base = a.chunk_offsets[k];  // no path's offsets pass reaches it: keeps the buffer's offset
        }
    }
    base = (uint32_t)__shfl((int)base, 0, 64);
    apply_scan_chunk(a, c, lane, base);
}


// Pass B: the chunks after the first of their path.  The offset is the path's
// segment_start plus the stored totals of its earlier chunks (chunk_offsets.hlsl
// :21-24), summed 64 at a time by one wave; the last chunk also stores the bump.
extern "C" __global__ __launch_bounds__(256) void tk_k_scan_chunk_apply(tk_args_scan_prefix a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t k = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (k >= a.cfg.scan_chunk_count) return;
    const tk_scan_chunk head = a.chunks[k];
    if (head.path_id >= a.cfg.path_count) return;
    const tk_scan_chunk_range r = a.chunk_ranges[head.path_id];
    if (!(k > r.start && k < r.end)) return;

    ScanChunkView c;
    load_scan_chunk(a, k, lane, c);
    uint32_t before = 0u;
    for (uint32_t j0 = r.start; j0 < k; j0 += 64u) {
        const uint32_t j = j0 + lane;
        before += j < k ? a.chunk_totals[j] : 0u;
    }
    const uint32_t base = c.seg_start + wave_sum_u32(before);
    if (lane == 0u) {
        a.chunk_offsets[k] = base;
        if (k + 1u == r.end) a.bumps[c.path] = base + a.chunk_totals[k] - c.seg_start;
    }
    apply_scan_chunk(a, c, lane, base);
}

// scan_emit (scan/emit.hlsl:14-28): each segment takes a slot of its (path, tile)
// range from that range's cursor.  The reference claims one slot per segment with
// an atomic; here a workgroup first gathers its segments: each goes into a
// staging area in LDS with its (path, tile) key, whose entry in an LDS key table
// counts the workgroup's segments of that key and gives the segment its rank.
// Then one cursor atomic per key claims the key's whole run of slots, and the
// staged segments are written to run start + rank.  The slots a range receives,
// and so its multiset of segments, are the reference's; a workgroup's segments
// of one tile land contiguous (written together, so L2 merges their lines), and
// the cursor atomics drop to one per (workgroup, key).  A segment that finds the
// staging area or its key's probe window full claims its slot on its own.
constexpr uint32_t kEmitNoKey = 0xFFFFFFFFu;
constexpr uint32_t kEmitProbes = 16u;
extern "C" __global__ __launch_bounds__(1024) void tk_k_scan_emit(tk_args_scan_lines a) {
    extern __shared__ uint32_t emit_lds[];
    const uint32_t H = a.hash_slots, S = a.stage_slots;
    uint32_t *hkey = emit_lds, *hcount = emit_lds + H, *hbase = emit_lds + 2u * H, *stage_n = emit_lds + 3u * H;
    // 8 words per staged segment, two 16-byte LDS accesses: p0x p0y p1x p1y | y_edge key-slot rank -
    uint4 *stage = reinterpret_cast<uint4 *>(emit_lds + 3u * H + 4u);
    for (uint32_t h = threadIdx.x; h < H; h += blockDim.x) {
        hkey[h] = kEmitNoKey;
        hcount[h] = 0u;
    }
    if (threadIdx.x == 0u) *stage_n = 0u;
    __syncthreads();

    const uint32_t block = xcd_chunked_block(blockIdx.x, a.xcd_run);
    const uint32_t id = block * blockDim.x + threadIdx.x;  // lines per workgroup: a launch parameter
    bool live = id < a.cfg.line_count;
    if (live && a.path_active) {  // incremental: only paths whose tile box meets the damage
        const uint32_t path = a.lines[id].path_id;
        if (path < a.path_record_count && a.path_active[path] == 0u) live = false;
    }
    ScanTraversal s;
    if (live) live = scan_geometry(a.lines, a.paths, a.path_record_count, id, s);
    uint32_t i = live ? s.imin : 0u;
    const uint32_t iend = live ? s.imax : 0u;
    while (__ballot(i < iend)) {  // the wave steps together; a lane past its tiles idles
        bool have = false;
        uint32_t key = 0u;
        tk_line_segment seg{};
        if (i < iend) {
            const float z = floorf(s.a * (float)i + s.b);
            int32_t tx, ty;
            scan_tile(s, i, z, tx, ty);
            if (scan_tile_inside(s, tx, ty)) {
                have = true;
                key = scan_tile_offset(s, tx, ty);
                seg = clipped_segment(s, i, z, tx, ty);
            }
            ++i;
        }
        // this step's staging positions: one LDS atomic per wave
        const unsigned long long m = __ballot(have);
        uint32_t first = 0u;
        if (m) {
            const uint32_t leader = (uint32_t)__ffsll(m) - 1u;
            if (__lane_id() == leader) first = atomicAdd(stage_n, (uint32_t)__popcll(m));
            first = (uint32_t)__shfl((int)first, (int)leader, 64);
        }
        bool staged = false;
        if (have) {
            const uint32_t pos = first + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (pos < S) {
                uint32_t h = (key * 2654435761u) & (H - 1u), slot = kEmitNoKey;
                for (uint32_t probe = 0; probe < kEmitProbes; ++probe, h = (h + 1u) & (H - 1u)) {
                    const uint32_t old = atomicCAS(&hkey[h], kEmitNoKey, key);
                    if (old == kEmitNoKey || old == key) {
                        slot = h;
                        break;
                    }
                }
                uint32_t rank = 0u;
                if (slot != kEmitNoKey) {
                    stage[2u * pos] = make_uint4(__float_as_uint(seg.p0x), __float_as_uint(seg.p0y),
                                                 __float_as_uint(seg.p1x), __float_as_uint(seg.p1y));
                    rank = atomicAdd(&hcount[slot], 1u);
                    staged = true;
                }
                // slot kEmitNoKey: a hole, skipped below
                stage[2u * pos + 1u] = make_uint4(__float_as_uint(seg.y_edge), slot, rank, 0u);
            }
        }
        if (have && !staged) {
            const uint32_t dst = atomicAdd(&a.counts_or_cursors[key], 1u);
            if (dst < a.cfg.segment_capacity) a.segments[dst] = seg;
        }
    }
    __syncthreads();
    // one claim per key: the key's run of slots
    for (uint32_t h = threadIdx.x; h < H; h += blockDim.x) {
        const uint32_t n = hcount[h];
        if (n) hbase[h] = atomicAdd(&a.counts_or_cursors[hkey[h]], n);
    }
    __syncthreads();
    const uint32_t staged_total = min(*stage_n, S);
    for (uint32_t e = threadIdx.x; e < staged_total; e += blockDim.x) {
        const uint4 tail = stage[2u * e + 1u];
        if (tail.y == kEmitNoKey) continue;
        const uint32_t dst = hbase[tail.y] + tail.z;
        if (dst < a.cfg.segment_capacity) {
            const uint4 head = stage[2u * e];
            a.segments[dst] = tk_line_segment{__uint_as_float(head.x), __uint_as_float(head.y), __uint_as_float(head.z),
                                              __uint_as_float(head.w), __uint_as_float(tail.x)};
        }
    }
}

// cumsum_prefix_chunks + cumsum_chunk_offsets + cumsum_apply_chunk_offsets
// (cumsum.hlsl:16-71): one thread carries one backdrop row across its chunks.
extern "C" __global__ __launch_bounds__(256) void tk_k_cumsum(tk_args_cumsum a) {
    uint32_t row = blockIdx.x * 256u + threadIdx.x;
    if (row >= a.cfg.row_count) return;
    uint32_t carry = 0u;
    uint32_t end = a.row_ends[row];
    for (uint32_t c = a.row_starts[row]; c < end && c < a.cfg.chunk_count; ++c) {
        uint32_t off = a.chunk_offsets_in[c];
        uint32_t len = min(a.chunk_lens[c], (uint32_t)TK_CUMSUM_CHUNK_SIZE);
        if (a.chunk_carries) a.chunk_carries[c] = carry;
        uint32_t sum = 0u;
        for (uint32_t i = 0; i < len; ++i) {
            sum += (uint32_t)a.backdrops[off + i];
            a.backdrops[off + i] = (int32_t)(carry + sum);
        }
        if (a.chunk_totals) a.chunk_totals[c] = sum;
        carry += sum;
    }
}

// ---------------------------------------------------------------------------
// Coarse: one 64-lane wave per tile.  Lanes take 64 consecutive entries of the
// tile's painter-ordered draw list at a time; a DrawBatch is a contiguous run of
// entries, so per-batch particle counts and emission offsets come from ballots.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t lanes_below(uint32_t lane) { return lane ? (~0ull >> (64u - lane)) : 0ull; }

// stack_wrapper_count with the layers of one batch evaluated across the wave:
// the stack is invalid when any layer rejects the tile, otherwise the count of
// layers that are not no-ops.
__device__ uint32_t stack_wrapper_count_wave(const CoarseCtx &c, uint32_t batch, uint32_t lane) {
    const tk_coarse_config &cfg = *c.cfg;
    const tk_coarse_inputs &in = *c.in;
    const tk_batch_record br = in.batches[batch];
    const uint32_t end = min(br.layer_end, cfg.layer_count);
    uint32_t count = 0u;
    for (uint32_t base = br.layer_start; base < end; base += 64u) {
        const uint32_t i = base + lane;
        bool invalid = false, counted = false;
        if (i < end) {
            const tk_layer_record layer = in.layers[i];
            const DrawData d = load_draw(in.draws, cfg.draw_count, layer.draw);
            if (draw_has_sdf(d)) {
                if (!sdf_clip_covers(in.paint, cfg.paint_words, d, c.tx, c.ty)) {
                    if (!draw_hits_tile(d, c.tx, c.ty, c.tw, c.th)) invalid = true;
                    else counted = true;
                }
            } else {
                const uint32_t bd = draw_backdrop(in.paths, cfg.path_count, d, c.tx, c.ty, c.tw, c.th);
                if (bd == kNone) {
                    invalid = true;
                } else {
                    const tk_segment_range range = in.segment_ranges[bd];
                    const int32_t winding = in.backdrops[bd];
                    const bool full = d.fill_rule == TK_FILL_EVEN_ODD ? (((uint32_t)winding & 1u) != 0u) : winding != 0;
                    if (!(layer.tag == TK_LAYER_CLIP && range.start == range.end && full)) {
                        if (range.start == range.end && winding == 0) invalid = true;
                        else counted = true;
                    }
                }
            }
        }
        if (__ballot(invalid)) return kNone;
        count += (uint32_t)__popcll(__ballot(counted));
    }
    return count;
}

// Writes the stack's BEGIN records (forward layer order) from dst; returns the next slot.
__device__ uint32_t emit_stack_begins_wave(const CoarseCtx &c, uint32_t batch, Emitter &e, uint32_t dst, uint32_t lane) {
    const tk_batch_record br = c.in->batches[batch];
    const uint32_t end = min(br.layer_end, c.cfg->layer_count);
    for (uint32_t base = br.layer_start; base < end; base += 64u) {
        const uint32_t i = base + lane;
        bool emit = false;
        tk_layer_record layer{};
        DrawData d{};
        if (i < end) {
            layer = c.in->layers[i];
            d = load_draw(c.in->draws, c.cfg->draw_count, layer.draw);
            emit = !noop_stack_clip(c, layer, d);
        }
        const uint64_t mask = __ballot(emit);
        if (emit) {
            const uint32_t slot = dst + (uint32_t)__popcll(mask & lanes_below(lane));
            if (draw_has_sdf(d)) {
                e.store(slot, TK_PTCL_BEGIN_SDF_CLIP, 0u, 0u, 0u, 0u, layer.draw);
            } else {
                const uint32_t bd = draw_backdrop(c.in->paths, c.cfg->path_count, d, c.tx, c.ty, c.tw, c.th);
                const uint32_t tag =
                    layer.tag == 1u ? kPtclBeginOpacity : (layer.tag == 2u ? kPtclBeginBlend : TK_PTCL_BEGIN_CLIP);
                tk_segment_range range{0u, 0u};
                uint32_t winding = 0u;
                if (bd != kNone) {
                    range = c.in->segment_ranges[bd];
                    winding = (uint32_t)c.in->backdrops[bd];
                }
                e.store(slot, tag, winding, d.fill_rule, range.start, range.end, layer.payload);
            }
        }
        dst += (uint32_t)__popcll(mask);
    }
    return dst;
}

// One pass over a stack of at most 64 layers (lane k takes layer layer_start + k):
// the validity and wrapper count of stack_wrapper_count (coarse/classify.hlsli:
// 11-33) and the BEGIN records of emit_stack_begins (coarse/emit_stack.hlsli:18-31)
// from the same loads.  Returns the slot after the BEGIN records; wrappers is
// kNone (and nothing is written) when a layer rejects the tile.  emitted and
// end_tag keep what emit_stack_ends needs, so the END records are written
// without loading the layers again (stack_ends_from_registers).
__device__ uint32_t stack_begins_single_pass(const CoarseCtx &c, uint32_t batch, Emitter &e, uint32_t dst,
                                             uint32_t lane, uint32_t &wrappers, uint64_t &emitted,
                                             uint32_t &end_tag) {
    const tk_coarse_config &cfg = *c.cfg;
    const tk_coarse_inputs &in = *c.in;
    const tk_batch_record br = in.batches[batch];
    const uint32_t end = min(br.layer_end, cfg.layer_count);
    const uint32_t i = br.layer_start + lane;
    bool invalid = false, counted = false, emit = false;
    uint32_t tag = TK_PTCL_BEGIN_CLIP, winding = 0u, fill_rule = 0u, payload = 0u;
    tk_segment_range range{0u, 0u};
    end_tag = TK_PTCL_END_CLIP;
    if (i < end) {
        const tk_layer_record layer = in.layers[i];
        end_tag = layer.tag == 1u ? kPtclEndOpacity : (layer.tag == 2u ? kPtclEndBlend : TK_PTCL_END_CLIP);
        const uint32_t bits = layer_bits_at(c, i, layer);
        invalid = (bits & kLayerInvalid) != 0u;
        counted = (bits & kLayerCounted) != 0u;
        emit = (bits & kLayerEmits) != 0u;
        if (bits & kLayerSdf) {
            tag = TK_PTCL_BEGIN_SDF_CLIP;
            payload = layer.draw;
        } else {
            tag = layer.tag == 1u ? kPtclBeginOpacity : (layer.tag == 2u ? kPtclBeginBlend : TK_PTCL_BEGIN_CLIP);
            payload = layer.payload;
            if (emit && !invalid) {  // a path clip's BEGIN record carries its tile segments and winding
                const DrawData d = load_draw(in.draws, cfg.draw_count, layer.draw);
                fill_rule = d.fill_rule;
                const uint32_t bd =
                    draw_backdrop(in.paths, cfg.path_count, d, c.tx, c.ty, c.tw, c.th);
                if (bd != kNone) {
                    range = in.segment_ranges[bd];
                    winding = (uint32_t)in.backdrops[bd];
                }
            }
        }
    }
    emitted = 0ull;
    if (__ballot(invalid)) {
        wrappers = kNone;
        return dst;
    }
    wrappers = (uint32_t)__popcll(__ballot(counted));
    emitted = __ballot(emit);
    if (emit) e.store(dst + (uint32_t)__popcll(emitted & lanes_below(lane)), tag, winding, fill_rule, range.start,
                      range.end, payload);
    return dst + (uint32_t)__popcll(emitted);
}

// emit_stack_ends for a stack handled by stack_begins_single_pass: the emitted
// layers' END records in reverse layer order from dst; returns the next slot.
__device__ __forceinline__ uint32_t stack_ends_from_registers(Emitter &e, uint32_t dst, uint32_t lane, uint64_t emitted,
                                                              uint32_t end_tag) {
    const uint32_t n = (uint32_t)__popcll(emitted);
    if ((emitted >> lane) & 1ull)
        e.store(dst + n - 1u - (uint32_t)__popcll(emitted & lanes_below(lane)), end_tag, 0u, 0u, 0u, 0u, 0u);
    return dst + n;
}

// Writes the stack's END records (reverse layer order) from dst; returns the next slot.
__device__ uint32_t emit_stack_ends_wave(const CoarseCtx &c, uint32_t batch, Emitter &e, uint32_t dst, uint32_t lane) {
    const tk_batch_record br = c.in->batches[batch];
    const uint32_t end = min(br.layer_end, c.cfg->layer_count);
    for (uint32_t top = end; top > br.layer_start;) {
        const uint32_t span = min(64u, top - br.layer_start);
        // lane k takes layer top-1-k, so lane order is emission order
        bool emit = false;
        uint32_t tag = TK_PTCL_END_CLIP;
        if (lane < span) {
            const tk_layer_record layer = c.in->layers[top - 1u - lane];
            const DrawData d = load_draw(c.in->draws, c.cfg->draw_count, layer.draw);
            emit = !noop_stack_clip(c, layer, d);
            tag = layer.tag == 1u ? kPtclEndOpacity : (layer.tag == 2u ? kPtclEndBlend : TK_PTCL_END_CLIP);
        }
        const uint64_t mask = __ballot(emit);
        if (emit) e.store(dst + (uint32_t)__popcll(mask & lanes_below(lane)), tag, 0u, 0u, 0u, 0u, 0u);
        dst += (uint32_t)__popcll(mask);
        top -= span;
    }
    return dst;
}

// One chunk of up to 64 list entries: lane's draw and batch (kNone when absent).
__device__ __forceinline__ void load_list_entry(const tk_args_coarse &a, const tk_tile_draw_record &list, uint32_t base,
                                                uint32_t lane, const SceneView &v, uint32_t &draw, uint32_t &batch) {
    draw = kNone;
    batch = kNone;
    const uint32_t i = base + lane;
    if (i < list.len) {
        const uint32_t slot = list.start + i;
        if (slot < a.cfg.tile_draw_index_count) {
            draw = a.in.tile_draw_indices[slot];
            if (draw < a.cfg.draw_count) batch = a.in.draw_batch_ids[draw];
            if (batch < v.q0 || batch >= v.q1) batch = kNone;  // outside this canvas's batches
        }
    }
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}

// The smallest batch id at or above `from` among a tile list's entries (kNone if
// none): the next batch of a list whose entries are not in batch order.
__device__ uint32_t next_listed_batch(const tk_args_coarse &a, const tk_tile_draw_record &list, uint32_t len,
                                      uint32_t lane, const SceneView &v, uint32_t from) {
    uint32_t m = kNone;
    for (uint32_t base = 0; base < len; base += 64u) {
        uint32_t d, b;
        load_list_entry(a, list, base, lane, v, d, b);
        if (b != kNone && b >= from) m = min(m, b);
    }
    return wave_min_u32(m);
}

// Wrapper validity and count (stack_wrapper_count, coarse/classify.hlsli:11-33)
// of up to 64 batches at once: lane j describes batch `batch` with `particles`
// particles; the (batch, layer) pairs of the batches that have particles are laid
// end to end and dealt to the lanes, so every layer's loads are in flight
// together.  Returns, in lane j, the wrapper count of batch j or kNone.
__device__ uint32_t stack_wrappers_of_batches(const CoarseCtx &c, uint32_t pending, uint32_t batch,
                                              uint32_t particles, uint32_t lane, uint32_t *invalid_flags,
                                              uint32_t *counted) {
    const tk_coarse_config &cfg = *c.cfg;
    const tk_coarse_inputs &in = *c.in;
    uint32_t first_layer = 0u, depth = 0u;
    if (lane < pending && particles > 0u) {
        const tk_batch_record br = in.batches[batch];
        const uint32_t end = min(br.layer_end, cfg.layer_count);
        first_layer = br.layer_start;
        depth = end > first_layer ? end - first_layer : 0u;
    }
    invalid_flags[lane] = 0u;
    counted[lane] = 0u;
    uint32_t total;
    const uint32_t off = wave_exclusive_scan(depth, lane, total);
    const uint32_t incl = off + depth;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t t0 = 0u; t0 < total; t0 += 64u) {
        const uint32_t t = t0 + lane;
        uint32_t j = 0u;  // owner: the number of batches whose layers all precede pair t
#pragma unroll
        for (uint32_t step = 32u; step >= 1u; step >>= 1) {
            if ((uint32_t)__shfl((int)incl, (int)(j + step - 1u), 64) <= t) j += step;
        }
        j = min(j, 63u);
        const uint32_t i = (uint32_t)__shfl((int)first_layer, (int)j, 64) + t - (uint32_t)__shfl((int)off, (int)j, 64);
        if (t < total) {
            const uint32_t bits = c.class_row ? c.class_row[c.class_pair_of[i]] : layer_class_bits(c, in.layers[i]);
            if (bits & kLayerInvalid) atomicOr(&invalid_flags[j], 1u);
            if (bits & kLayerCounted) atomicAdd(&counted[j], 1u);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return invalid_flags[lane] ? kNone : counted[lane];
}

// The class table: layer_class_bits of every distinct (draw, tag) pair of the layer
// records on every table column's tile, bits[column * pair_count + pair].  A
// workgroup takes one pair over 256 consecutive columns, so the layer, its draw
// record and its SDF words are the same in every lane (scalar loads, loaded once
// per wave) and only the tile position varies.
extern "C" __global__ __launch_bounds__(256) void tk_k_layer_classes(tk_args_layer_classes a) {
    const uint32_t groups = (a.list.columns + 255u) / 256u;  // workgroups per pair
    const uint32_t pair = blockIdx.x / groups;
    const uint32_t column = (blockIdx.x - pair * groups) * 256u + threadIdx.x;
    if (pair >= a.pair_count || column >= a.list.columns) return;
    const uint32_t tile = a.list.tiles ? a.list.tiles[column] : column;
    const CoarseCtx c{&a.cfg,   &a.in,   tile % a.cfg.tiles_width, tile / a.cfg.tiles_width, nullptr, nullptr,
                      a.cfg.tiles_width, a.cfg.tiles_height};
    a.bits[(uint64_t)column * a.pair_count + pair] = (uint8_t)layer_class_bits(c, a.in.layers[a.pair_layer[pair]]);
}

// coarse_count_bins for every batch (coarse/count.hlsl:48-73): one wave per tile.
// The tile's draw list is read 64 entries at a time and its particles counted per
// batch; the batches met are kept one per lane and their stacks resolved together
// (stack_wrappers_of_batches) once 64 are pending or the list ends.
// kScenes: a batch of scenes (per-tile canvases from a.scenes); else one canvas.
template <bool kScenes>
__device__ __forceinline__ void coarse_count_tile(const tk_args_coarse &a, uint32_t column) {
    __shared__ uint32_t invalid_flags[4][64], counted[4][64];
    const tk_coarse_config &cfg = a.cfg;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    if (column >= a.list.columns) return;
    const SceneView v = scene_view<kScenes>(a.scenes, column, a.list.columns, cfg.batch_count, cfg.tiles_width,
                                            cfg.tiles_height, 0u, 0u, 0u);
    const uint32_t tile = a.list.tiles ? a.list.tiles[column] : v.local;
    const CoarseCtx c = coarse_ctx(a, v, tile, column);
    uint32_t *counts = a.out.tile_ptcl_counts;
    const tk_tile_draw_record list = a.in.tile_draws[kScenes ? column : tile];
    const uint32_t len = list.start == kNone ? 0u : list.len;
    const uint32_t Q = v.q1;

    uint32_t next = v.q0;                   // first batch whose count is not written yet
    uint32_t pending = 0u;                  // batches held, one per lane
    uint32_t my_batch = 0u, my_particles = 0u;
    auto resolve = [&]() {
        if (pending == 0u) return;
        const uint32_t w = stack_wrappers_of_batches(c, pending, my_batch, my_particles, lane, invalid_flags[wave],
                                                     counted[wave]);
        uint32_t prev = (uint32_t)__shfl_up((int)my_batch, 1u, 64);
        if (lane == 0u) prev = next - 1u;  // wraps to ~0u when next == 0: zeroing then starts at batch 0
        if (lane < pending) {
            for (uint32_t bb = prev + 1u; bb < my_batch; ++bb) counts[table_index(v, bb)] = 0u;
            const uint32_t count = (my_particles > 0u && w != kNone) ? my_particles + 2u * w + 1u : 0u;
            counts[table_index(v, my_batch)] = count;
        }
        next = (uint32_t)__shfl((int)my_batch, (int)(pending - 1u), 64) + 1u;
        pending = 0u;
    };
    uint32_t cur = kNone, particles = 0u;
    auto flush = [&]() {
        if (cur == kNone) return;
        if (lane == pending) {
            my_batch = cur;
            my_particles = particles;
        }
        if (++pending == 64u) resolve();
    };
    // A list out of batch order (the recorder never produces one; the stage entry
    // accepts it, as the reference's per-batch dispatches filter any list by
    // batch) is counted again one batch at a time, in batch order, each over the
    // whole list filtered to that batch (`only`); every count is then rewritten.
    bool disordered = false;
    uint32_t only = kNone;
    for (;;) {
        for (uint32_t base = 0; base < len; base += 64u) {
            uint32_t d, b;
            load_list_entry(a, list, base, lane, v, d, b);
            if (only != kNone && b != only) b = kNone;
            const bool valid = b != kNone && draw_particle(c, b, d).valid;
            uint64_t pending_runs = __ballot(b != kNone);
            const uint64_t valid_mask = __ballot(valid);
            while (pending_runs) {
                const uint32_t first = (uint32_t)__ffsll((unsigned long long)pending_runs) - 1u;
                const uint32_t bb = (uint32_t)__shfl((int)b, (int)first, 64);
                const uint64_t run = __ballot(b == bb) & pending_runs;
                pending_runs &= ~run;
                if (bb != cur) {
                    if (cur != kNone && bb < cur) {  // a batch met again after a later one
                        disordered = true;
                        continue;
                    }
                    flush();
                    cur = bb;
                    particles = 0u;
                }
                particles += (uint32_t)__popcll(run & valid_mask);
            }
        }
        if (!disordered) break;
        uint32_t from = only + 1u;
        if (only == kNone) {  // restart: drop what the first walk held
            if (lane == 0u) atomicOr(&a.out.status[3], 1u);
            pending = 0u;
            next = v.q0;
            cur = kNone;
            from = v.q0;
        }
        only = next_listed_batch(a, list, len, lane, v, from);
        if (only == kNone) break;
    }
    flush();
    resolve();
    for (uint32_t bb = next + lane; bb < Q; bb += 64u) counts[table_index(v, bb)] = 0u;
}

// A coarse launch: one tile per wave, up to four per workgroup in XCD runs, or
// (kPersistent: a grid of resident workgroups) each wave taking tiles from the
// per-XCD work queues (the eight counters at a.queue + first) until they are
// used up.
template <bool kPersistent, class TileFn>
__device__ __forceinline__ void coarse_tiles(const tk_args_coarse &a, uint32_t first, TileFn tile) {
    if (!kPersistent) {  // blockDim.x / 64 tiles per workgroup (1, 2 or 4: a launch choice)
        tile(xcd_chunked_block(blockIdx.x, a.xcd_run) * (blockDim.x >> 6) + (threadIdx.x >> 6));
        return;
    }
    uint32_t *counters = a.queue + first;
    uint32_t qi = 0u;
    for (;;) {
        uint32_t column = 0u;
        if ((threadIdx.x & 63u) == 0u) column = queue_next(counters, blockIdx.x & 7u, qi, a.xcd_run * (blockDim.x >> 6), a.list.columns);
        column = (uint32_t)__shfl((int)column, 0, 64);
        if (column == ~0u) return;
        tile(column);
    }
}

extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void tk_k_coarse_count(tk_args_coarse a) {
    coarse_tiles<false>(a, 0u, [&](uint32_t column) { coarse_count_tile<false>(a, column); });
}
extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_count_scenes(tk_args_coarse a) {
    coarse_tiles<false>(a, 0u, [&](uint32_t column) { coarse_count_tile<true>(a, column); });
}
extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void tk_k_coarse_count_persistent(
    tk_args_coarse a) {
    coarse_tiles<true>(a, 0u, [&](uint32_t column) { coarse_count_tile<false>(a, column); });
}
extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_count_scenes_persistent(tk_args_coarse a) {
    coarse_tiles<true>(a, 0u, [&](uint32_t column) { coarse_count_tile<true>(a, column); });
}

// coarse_prefix_chunks + coarse_chunk_offsets + coarse_apply_chunk_offsets for
// every batch: one wrapping exclusive scan over the batch-major count table.
// A batch's reference (per-batch) offsets are start - batch_ptcl_bases[batch].
// kItems consecutive counts per lane, 256 * kItems per workgroup (a.scan_blocks
// workgroups): the scan block size is a launch choice among 1024-8192 counts.
template <uint32_t kItems>
__device__ __forceinline__ void coarse_scan(const tk_args_coarse &a) {
    __shared__ uint32_t s_ticket, s_base;
    __shared__ uint32_t s_wave[4];
    const uint32_t k = draw_ticket(a.ticket, a.scan_blocks, &s_ticket);
    const uint64_t n = a.scenes.slots ? a.scenes.table_entries : (uint64_t)a.cfg.batch_count * a.list.columns;
    const uint64_t first = ((uint64_t)k * 256u + threadIdx.x) * kItems;
    uint32_t v[kItems];
    uint32_t sum = 0u;
    if (first + kItems <= n && ((uintptr_t)a.out.tile_ptcl_counts & 15u) == 0u) {
        const uint4 *src = reinterpret_cast<const uint4 *>(a.out.tile_ptcl_counts + first);
#pragma unroll
        for (uint32_t q = 0; q < kItems / 4u; ++q) {
            const uint4 w = src[q];
            v[4 * q] = w.x;
            v[4 * q + 1] = w.y;
            v[4 * q + 2] = w.z;
            v[4 * q + 3] = w.w;
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < kItems; ++i) v[i] = first + i < n ? a.out.tile_ptcl_counts[first + i] : 0u;
    }
#pragma unroll
    for (uint32_t i = 0; i < kItems; ++i) sum += v[i];
    uint32_t total;
    const uint32_t excl = block_exclusive_scan_256(sum, s_wave, total);
    if (threadIdx.x < 64u) {
        uint32_t base = 0u;
        if (k > 0u) {
            if (threadIdx.x == 0u) publish(a.scan_status, k, total, kFlagAggregate, a.epoch);
            base = wave_lookback(a.scan_status, k, a.epoch, threadIdx.x);
        }
        if (threadIdx.x == 0u) {
            publish(a.scan_status, k, base + total, kFlagPrefix, a.epoch);
            s_base = base;
            if (k + 1u == a.scan_blocks) {
                a.out.batch_ptcl_bases[a.cfg.batch_count] = base + total;
                a.out.status[1] = base + total;
            }
        }
    }
    __syncthreads();
    if (first >= n) return;
    uint32_t running = s_base + excl;
    uint64_t batch = first / a.list.columns;
    uint32_t offset = (uint32_t)(first - batch * a.list.columns);
#pragma unroll
    for (uint32_t i = 0; i < kItems; ++i) {
        const uint64_t idx = first + i;
        if (idx < n) {
            a.out.tile_ptcl_ranges[idx] = tk_segment_range{running, running + v[i]};
            if (offset == 0u && !a.scenes.slots) a.out.batch_ptcl_bases[batch] = running;
        }
        running += v[i];
        if (++offset == a.list.columns) {
            offset = 0u;
            ++batch;
        }
    }
}

extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_scan(tk_args_coarse a) {
    coarse_scan<TK_COARSE_SCAN_ITEMS / 256u>(a);  // 4096 counts per workgroup
}
extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_scan_1k(tk_args_coarse a) { coarse_scan<4u>(a); }
extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_scan_2k(tk_args_coarse a) { coarse_scan<8u>(a); }
extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_scan_8k(tk_args_coarse a) { coarse_scan<32u>(a); }

// coarse_emit_bins for every batch (coarse/emit.hlsl:82-142): writes
// [stack begins][particles][stack ends][END] for each (batch, tile) with a
// nonzero count and the tile kind word, including the opaque-prefix skip that
// occluded_particle_prefix (coarse/occlusion.hlsli:12-34) derives: the last
// opaque colour particle, unless a begin record follows it.
template <bool kScenes>
__device__ __forceinline__ void coarse_emit_tile(const tk_args_coarse &a, uint32_t column) {
    const tk_coarse_config &cfg = a.cfg;
    const uint32_t lane = threadIdx.x & 63u;
    if (column >= a.list.columns) return;
    const SceneView v = scene_view<kScenes>(a.scenes, column, a.list.columns, cfg.batch_count, cfg.tiles_width,
                                            cfg.tiles_height, 0u, 0u, 0u);
    const uint32_t tile = a.list.tiles ? a.list.tiles[column] : v.local;
    const CoarseCtx c = coarse_ctx(a, v, tile, column);
    const uint32_t Q = v.q1;
    Emitter e{a.out.ptcl, cfg.ptcl_capacity, a.out.status};
    uint32_t *kinds = a.out.tile_kinds;
    const tk_tile_draw_record list = a.in.tile_draws[kScenes ? column : tile];
    const uint32_t len = list.start == kNone ? 0u : list.len;

    // Uniform per-batch state.
    uint32_t cur = kNone, next = v.q0;
    bool active = false;
    uint32_t start = 0u, end = 0u, wrappers = 0u, dst = 0u, flags = 0u, last_opaque = 0u;
    bool begin_after = false;
    bool single_pass = false;  // the batch's stack is at most 64 layers: ENDs come from registers
    uint64_t stack_emitted = 0ull;
    uint32_t stack_end_tag = TK_PTCL_END_CLIP;
    auto empty_until = [&](uint32_t stop) {
        for (uint32_t bb = next + lane; bb < stop; bb += 64u) kinds[table_index(v, bb)] = TK_TILE_KIND_EMPTY;
    };
    auto finish = [&]() {
        if (cur == kNone) return;
        uint32_t kind = TK_TILE_KIND_EMPTY;
        if (active) {
            dst = single_pass ? stack_ends_from_registers(e, dst, lane, stack_emitted, stack_end_tag)
                              : emit_stack_ends_wave(c, cur, e, dst, lane);
            if (lane == 0u) e.store(dst, TK_PTCL_END, 0u, 0u, 0u, 0u, 0u);
            uint32_t skip = 0u;
            if ((flags & 8u) && wrappers == 0u && end - start > 2u && !begin_after) skip = last_opaque;
            kind = classify_tile_flags(flags) | (skip << TK_TILE_PREFIX_SHIFT);
        }
        if (lane == 0u) kinds[table_index(v, cur)] = kind;
        next = cur + 1u;
    };
    auto begin = [&](uint32_t bb) {
        cur = bb;
        const tk_segment_range r = a.out.tile_ptcl_ranges[table_index(v, bb)];
        start = r.start;
        end = r.end;
        active = end > start;
        flags = 0u;
        last_opaque = 0u;
        begin_after = false;
        wrappers = 0u;
        dst = start;
        if (active) {
            const tk_batch_record br = a.in.batches[bb];
            const uint32_t layer_end = min(br.layer_end, cfg.layer_count);
            single_pass = layer_end <= br.layer_start || layer_end - br.layer_start <= 64u;
            if (single_pass) {
                dst = stack_begins_single_pass(c, bb, e, start, lane, wrappers, stack_emitted, stack_end_tag);
            } else {
                wrappers = stack_wrapper_count_wave(c, bb, lane);
            }
            if (wrappers == kNone) {  // cannot happen with a consistent count; keep the stream terminated
                if (lane == 0u) e.store(start, TK_PTCL_END, 0u, 0u, 0u, 0u, 0u);
                active = false;
                wrappers = 0u;
                return;
            }
            flags = wrappers != 0u ? kClassOther : 0u;
            if (!single_pass) dst = emit_stack_begins_wave(c, bb, e, start, lane);
        }
    };

    // A list out of batch order (see coarse_count_tile) is emitted again one batch
    // at a time in batch order, from the whole list filtered to that batch
    // (`only`); every stream and kind word of the tile is then rewritten.
    bool disordered = false;
    uint32_t only = kNone;
    for (;;) {
        for (uint32_t base = 0; base < len; base += 64u) {
            uint32_t d, b;
            load_list_entry(a, list, base, lane, v, d, b);
            if (only != kNone && b != only) b = kNone;
            // Every lane's particle and class at once (a lane's batch is its own b), so
            // the particle code runs once per 64 entries however many batches they span.
            Particle p{false, TK_PTCL_FILL, 0u, 0u, 0u, 0u, 0u};
            if (b != kNone) p = draw_particle(c, b, d);
            uint32_t cls = 0u;
            if (p.valid) cls = particle_class_flags(c, d, p.tag);
            const bool p_opaque = p.valid && p.tag == TK_PTCL_COLOR && (p.color >> 24) == 255u;
            const bool p_begins = p.valid && is_begin_tag(p.tag);
            uint64_t pending = __ballot(b != kNone);
            while (pending) {
                const uint32_t first = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
                const uint32_t bb = (uint32_t)__shfl((int)b, (int)first, 64);
                const uint64_t run = __ballot(b == bb) & pending;
                pending &= ~run;
                if (bb != cur) {
                    if (cur != kNone && bb < cur) {  // a batch met again after a later one
                        disordered = true;
                        continue;
                    }
                    finish();
                    empty_until(bb);
                    begin(bb);
                }
                if (!active) continue;
                const bool mine = (run >> lane) & 1ull;
                const bool emits = mine && p.valid;
                const uint64_t valid = __ballot(emits);
                const uint32_t pos = dst + (uint32_t)__popcll(valid & lanes_below(lane));
                const uint32_t my_cls = emits ? cls : 0u;
                const bool opaque = emits && p_opaque, begins = emits && p_begins;
                if (emits && pos < end) e.store(pos, p.tag, p.winding, p.fill_rule, p.seg_start, p.seg_end, p.color);
                if (__ballot(my_cls & kClassColor)) flags |= kClassColor;
                if (__ballot(my_cls & kClassSdf)) flags |= kClassSdf;
                if (__ballot(my_cls & kClassOther)) flags |= kClassOther;
                if (__ballot(opaque && wrappers == 0u && pos > start)) flags |= 8u;
                const uint64_t opaque_mask = __ballot(opaque), begin_mask = __ballot(begins);
                if (opaque_mask) {
                    const uint32_t h = 63u - (uint32_t)__clzll((long long)opaque_mask);
                    last_opaque = dst + (uint32_t)__popcll(valid & lanes_below(h)) - start;
                    begin_after = (begin_mask >> h) > 1ull;
                } else if (begin_mask) {
                    begin_after = true;
                }
                dst += (uint32_t)__popcll(valid);
            }
        }
        if (!disordered) break;
        uint32_t from = only + 1u;
        if (only == kNone) {  // restart: the first walk's streams are all rewritten
            cur = kNone;
            next = v.q0;
            from = v.q0;
        }
        only = next_listed_batch(a, list, len, lane, v, from);
        if (only == kNone) break;
    }
    finish();
    empty_until(Q);
}

extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_emit(tk_args_coarse a) {
    coarse_tiles<false>(a, 8u, [&](uint32_t column) { coarse_emit_tile<false>(a, column); });
}
extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_emit_scenes(tk_args_coarse a) {
    coarse_tiles<false>(a, 8u, [&](uint32_t column) { coarse_emit_tile<true>(a, column); });
}
extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_emit_persistent(tk_args_coarse a) {
    coarse_tiles<true>(a, 8u, [&](uint32_t column) { coarse_emit_tile<false>(a, column); });
}
extern "C" __global__ __launch_bounds__(256) void tk_k_coarse_emit_scenes_persistent(tk_args_coarse a) {
    coarse_tiles<true>(a, 8u, [&](uint32_t column) { coarse_emit_tile<true>(a, column); });
}

// The colour sample_draw_brush returns for a draw whose brush does not depend on
// the pixel (solid, and the kinds sample_brush maps to the header colour or to
// 0), so that a fill reads one word pair instead of the draw record and then
// its brush header (shared/brush/sample.hlsli:15-42).
extern "C" __global__ __launch_bounds__(256) void tk_k_draw_brush(tk_args_draw_brush a) {
    const uint32_t d = blockIdx.x * 256u + threadIdx.x;
    if (d >= a.draw_count) return;
    const uint32_t offset = a.draws[d].brush_offset;
    const uint32_t kind = blob_word(a.paint, a.paint_words, a.paint_brush_base, offset);
    const uint32_t color = blob_word(a.paint, a.paint_words, a.paint_brush_base, offset + 4u);
    uint32_t flag = 1u, value = color;
    if (kind == TK_BRUSH_LINEAR || kind == TK_BRUSH_RADIAL || kind == TK_BRUSH_SWEEP || kind == TK_BRUSH_FOUR_CORNER ||
        kind == TK_BRUSH_IMAGE)
        flag = 0u, value = 0u;
    else if (kind == 6u)  // BRUSH_PATTERN
        value = 0u;
    a.table[2u * d] = flag;
    a.table[2u * d + 1u] = value;
}

// fine_tile_main for every batch: one 256-lane workgroup per tile, one lane per
// pixel.  Between batches the pixel goes through the RGBA8 store/load round trip
// the reference performs on its target.  kScenes: a batch of scenes.
template <bool kScenes, bool kPersistent>
__device__ __forceinline__ void fine_tile(const tk_args_fine &a) {
    // Clip levels beyond the four register slots: clip_spill_depth * 256 bytes of
    // dynamic LDS, sized by the launch.
    extern __shared__ uint8_t clip_spill[];
    __shared__ CoverageAcc cov_acc[2];  // double-buffered: one barrier per fill
    __shared__ uint32_t batch_kind[256];
    __shared__ tk_segment_range batch_range[256];
    __shared__ uint32_t s_item;
    const tk_fine_config &cfg = a.cfg;
    const uint32_t columns = a.list.columns;  // table columns: listed tiles, or every tile
    const uint32_t lane = threadIdx.x;

    // Cleared once per workgroup: each take_coverage clears the entries it read.
    cov_acc[0].run[lane] = 0;
    cov_acc[0].part[lane] = 0;
    cov_acc[1].run[lane] = 0;
    cov_acc[1].part[lane] = 0;
    if (lane < 16u) {
        cov_acc[0].base[lane] = 0;
        cov_acc[1].base[lane] = 0;
    }
    uint32_t acc_sel = 0u;

    auto tile = [&](uint32_t column) {
        const SceneView v = scene_view<kScenes>(a.scenes, column, columns, cfg.batch_count, cfg.tiles_width,
                                                cfg.tiles_height, cfg.width, cfg.height, cfg.clear_color);
        const uint32_t tile_ix = a.list.tiles ? a.list.tiles[column] : v.local;  // the tile of its canvas
        if (tile_ix >= cfg.tile_count) return;
        const uint32_t lx = lane & 15u, ly = lane >> 4;
        const uint32_t gx = (tile_ix % v.tw) * 16u + lx;
        const uint32_t gy = (tile_ix / v.tw) * 16u + ly;
        const float sx = (float)gx + 0.5f, sy = (float)gy + 0.5f;
        const uint32_t *paint = a.in.paint;
        const tk_ptcl_record *ptcl = a.in.ptcl;
        const tk_line_segment *segs = a.in.segments;

        Px pixel = rgba8_to_unorm(v.clear);
        for (uint32_t b0 = v.q0; b0 < v.q1; b0 += 256u) {
          // This tile's (kind, range) word pairs of up to 256 batches, one load per lane.
          __syncthreads();
          if (b0 + lane < v.q1) {
              const size_t idx = table_index(v, b0 + lane);
              batch_kind[lane] = a.in.tile_kinds[idx];
              batch_range[lane] = a.in.tile_ptcl_ranges[idx];
          }
          __syncthreads();
          const uint32_t chunk = min(256u, v.q1 - b0);
          for (uint32_t bj = 0; bj < chunk; ++bj) {
            const uint32_t kind_word = batch_kind[bj];
            const uint32_t kind = kind_word & TK_TILE_KIND_MASK;
            if (kind == TK_TILE_KIND_EMPTY) continue;
            const tk_segment_range r = batch_range[bj];
            const uint32_t pstart = r.start + (kind_word >> TK_TILE_PREFIX_SHIFT);
            const uint32_t pend = min(r.end, cfg.ptcl_capacity);
            if (kind == TK_TILE_KIND_INTERPRETER) {
                uint32_t clip_mask = 255u, clip_depth = 0u;
                uint32_t cs0 = 255u, cs1 = 255u, cs2 = 255u, cs3 = 255u;
                auto push_clip = [&](uint32_t mask) {
                    uint32_t dd = clip_depth;
                    if (dd == 0u) { cs0 = mask; clip_depth = 1u; }
                    else if (dd == 1u) { cs1 = mask; clip_depth = 2u; }
                    else if (dd == 2u) { cs2 = mask; clip_depth = 3u; }
                    else if (dd == 3u) { cs3 = mask; clip_depth = 4u; }
                    else {
                        uint32_t spill = dd - 4u;
                        if (spill < cfg.clip_spill_depth) {
                            clip_spill[spill * 256u + lane] = (uint8_t)mask;
                            clip_depth = dd + 1u;
                        } else if (a.in.status) {
                            atomicOr(a.in.status, 2u);  // deeper than the scene's declared clip depth
                        }
                    }
                };
                for (uint32_t pi = pstart; pi < pend; ++pi) {
                    const tk_ptcl_record rec = ptcl[pi];
                    const uint32_t tag = rec.tag;
                    if (tag == TK_PTCL_END) break;
                    if (tag == TK_PTCL_COLOR) {
                        if (clip_mask != 0u) {
                            if (clip_mask == 255u && premul_opaque(rec.color)) pixel = rgba8_to_unorm(rec.color);
                            else pixel = src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(rec.color, clip_mask));
                        }
                    } else if (tag == kPtclImage) {
                        // interpreter.hlsli:65-74: an image covering the whole tile
                        if (clip_mask != 0u) {
                            const uint32_t color = sample_draw_brush(
                                cfg, paint, load_fine_draw(a.in.draws, cfg.draw_count, rec.color), sx, sy, atlas_of(a.in));
                            if (clip_mask == 255u && premul_opaque(color)) pixel = rgba8_to_unorm(color);
                            else pixel = src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(color, clip_mask));
                        }
                    } else if (tag == TK_PTCL_SDF) {
                        if (clip_mask != 0u) {
                            FineDraw d = load_fine_draw(a.in.draws, cfg.draw_count, rec.color);
                            float coverage = sdf_coverage_from_draw(cfg, paint, d, sx, sy);
                            uint32_t alpha = combine_alpha(coverage_to_u8(coverage), clip_mask);
                            if (alpha != 0u) {
                                uint32_t color = sample_draw_brush(cfg, paint, d, sx, sy, atlas_of(a.in));
                                if (alpha == 255u && premul_opaque(color)) pixel = rgba8_to_unorm(color);
                                else pixel = src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(color, alpha));
                            }
                        }
                    } else if (tag == TK_PTCL_END_CLIP) {
                        if (clip_depth > 0u) {
                            clip_depth -= 1u;
                            if (clip_depth == 0u) clip_mask = cs0;
                            else if (clip_depth == 1u) clip_mask = cs1;
                            else if (clip_depth == 2u) clip_mask = cs2;
                            else if (clip_depth == 3u) clip_mask = cs3;
                            else {
                                uint32_t spill = clip_depth - 4u;
                                if (spill < cfg.clip_spill_depth) clip_mask = clip_spill[spill * 256u + lane];
                            }
                        } else {
                            clip_mask = 255u;
                        }
                    } else if (tag == TK_PTCL_BEGIN_SDF_CLIP) {
                        uint32_t parent = clip_mask, alpha = 0u;
                        if (parent != 0u) {
                            FineDraw d = load_fine_draw(a.in.draws, cfg.draw_count, rec.color);
                            alpha = coverage_to_u8(sdf_coverage_from_draw(cfg, paint, d, sx, sy));
                        }
                        push_clip(parent);
                        clip_mask = combine_alpha(parent, alpha);
                    } else if (tag == TK_PTCL_FILL || tag == TK_PTCL_BEGIN_CLIP || tag == kPtclPathGlyph) {
                        const uint32_t s0 = rec.segment_start, s1 = min(rec.segment_end, cfg.segment_capacity);
                        const uint32_t n = s1 > s0 ? s1 - s0 : 0u;
                        // Issued before the coverage work so that its latency overlaps it.
                        uint32_t brush_flag = 0u, brush_color = 0u;
                        if (tag != TK_PTCL_BEGIN_CLIP && rec.color < cfg.draw_count) {
                            brush_flag = a.draw_brush[2u * rec.color];
                            brush_color = a.draw_brush[2u * rec.color + 1u];
                        }
                        CoverageAcc &acc = cov_acc[acc_sel];
                        acc_sel ^= 1u;
                        accumulate_record(segs + s0, n, acc, lane);
                        __syncthreads();
                        const float v = take_coverage(acc, rec.backdrop, lane);
                        uint32_t alpha = coverage_to_alpha(v, rec.fill_rule);
                        if (tag == TK_PTCL_BEGIN_CLIP) {
                            push_clip(clip_mask);
                            clip_mask = combine_alpha(clip_mask, alpha);
                        } else {
                            uint32_t masked = combine_alpha(alpha, clip_mask);
                            if (masked != 0u) {
                                uint32_t color = brush_color;
                                if (brush_flag == 0u)
                                    color = sample_draw_brush(cfg, paint, load_fine_draw(a.in.draws, cfg.draw_count, rec.color),
                                                              sx, sy, atlas_of(a.in));
                                if (tag == kPtclPathGlyph)  // text compositing (interpreter.hlsli:205-206)
                                    pixel = rgba8_to_unorm(glyph_composite(unorm_to_rgba8(pixel), color, masked));
                                else if (masked == 255u && premul_opaque(color)) pixel = rgba8_to_unorm(color);
                                else pixel = src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(color, masked));
                            }
                        }
                    }
                    // Glyph-image, opacity and blend records are not produced on this path.
                }
            } else if (kind == TK_TILE_KIND_COLOR) {
                for (uint32_t pi = pstart; pi < pend; ++pi) {
                    const tk_ptcl_record rec = ptcl[pi];
                    if (rec.tag == TK_PTCL_END) break;
                    if (premul_opaque(rec.color)) pixel = rgba8_to_unorm(rec.color);
                    else pixel = src_over_premul_unorm(pixel, rgba8_to_unorm(rec.color));
                }
            } else {  // TILE_KIND_SDF / TILE_KIND_MIXED
                for (uint32_t pi = pstart; pi < pend; ++pi) {
                    const tk_ptcl_record rec = ptcl[pi];
                    if (rec.tag == TK_PTCL_END) break;
                    if (rec.tag == TK_PTCL_COLOR) {
                        if (premul_opaque(rec.color)) pixel = rgba8_to_unorm(rec.color);
                        else pixel = src_over_premul_unorm(pixel, rgba8_to_unorm(rec.color));
                    } else if (rec.tag == kPtclImage) {
                        const uint32_t color = sample_draw_brush(
                            cfg, paint, load_fine_draw(a.in.draws, cfg.draw_count, rec.color), sx, sy, atlas_of(a.in));
                        if (premul_opaque(color)) pixel = rgba8_to_unorm(color);
                        else pixel = src_over_premul_unorm(pixel, rgba8_to_unorm(color));
                    } else {
                        FineDraw d = load_fine_draw(a.in.draws, cfg.draw_count, rec.color);
                        float coverage = 0.0f;
                        if ((int32_t)gx >= d.pb[0] && (int32_t)gy >= d.pb[1] && (int32_t)gx < d.pb[2] && (int32_t)gy < d.pb[3])
                            coverage = sdf_coverage_from_draw(cfg, paint, d, sx, sy);
                        uint32_t alpha = coverage_to_u8(coverage);
                        if (alpha != 0u) {
                            uint32_t color = blob_word(paint, cfg.paint_words, cfg.paint_brush_base, d.brush_offset + 4u);
                            if (alpha == 255u && premul_opaque(color)) pixel = rgba8_to_unorm(color);
                            else pixel = src_over_premul_unorm(pixel, scale_premul_u8_to_unorm(color, alpha));
                        }
                    }
                }
            }
            pixel = rgba8_to_unorm(unorm_to_rgba8(pixel));
          }
        }
        if (gx < v.width && gy < v.height) a.rgba[v.image + (size_t)gy * v.width + gx] = unorm_to_rgba8(pixel);
    };

    if (!kPersistent) {  // one tile per workgroup
        const uint32_t linear = blockIdx.x + blockIdx.y * gridDim.x;
        if (linear >= xcd_grid(columns, a.xcd_run)) return;
        const uint32_t column = xcd_chunked_block(linear, a.xcd_run);
        if (column < columns) tile(column);
        return;
    }
    // A persistent launch: tiles from the per-XCD work queues until they are used up.
    uint32_t qi = 0u;
    for (;;) {
        if (lane == 0u) s_item = queue_next(a.queue, blockIdx.x & 7u, qi, a.xcd_run, columns);
        __syncthreads();
        const uint32_t column = s_item;
        __syncthreads();
        if (column == ~0u) return;
        tile(column);
    }
}

// Seven waves per SIMD: at most 72 VGPRs (a few dwords spill, in the per-batch-run
// setup outside the record loop).
extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void tk_k_fine(tk_args_fine a) {
    fine_tile<false, false>(a);
}
extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void tk_k_fine_scenes(
    tk_args_fine a) {
    fine_tile<true, false>(a);
}
// Persistent forms: a grid of resident workgroups draining the per-XCD queues.
extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void tk_k_fine_persistent(
    tk_args_fine a) {
    fine_tile<false, true>(a);
}
extern "C" __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void tk_k_fine_scenes_persistent(
    tk_args_fine a) {
    fine_tile<true, true>(a);
}

// fill_coverage_words (validation/geometry.hlsl:22-29) plus the raw FP32 area.
extern "C" __global__ __launch_bounds__(256) void tk_k_fill_coverage(tk_args_fill_coverage a) {
    uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= a.count) return;
    const uint32_t *q = a.requests + (size_t)id * 8u;
    float v = exact_fill_value(a.segments, (int32_t)q[2], q[0], q[1], q[3], q[4]);
    a.alpha_out[id] = coverage_to_alpha(v, q[5]);
    if (a.coverage_out) a.coverage_out[id] = v;
}

// pixel_math_words (validation/pixel.hlsl:9-34).
extern "C" __global__ __launch_bounds__(256) void tk_k_pixel_math(tk_args_pixel_math a) {
    uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= a.count) return;
    const uint32_t *q = a.requests + (size_t)id * 8u;
    uint32_t *o = a.out + (size_t)id * 14u;
    uint32_t av = q[0], bv = q[1], src = q[2], dst = q[3];
    float c = asfloat(q[4]), t = asfloat(q[5]);
    uint32_t rule = q[6];
    Px pixel = rgba8_to_unorm(src);
    o[0] = mul_div255(av, bv);
    o[1] = combine_alpha(av, bv);
    o[2] = scale_premul_u8(src, bv);
    o[3] = src_over_premul_u8(dst, src);
    o[4] = src;
    o[5] = coverage_to_u8(c);
    o[6] = coverage_to_alpha(c, rule);
    o[7] = lerp_premul_u8(src, dst, t);
    o[8] = unorm_to_rgba8(src_over_premul_unorm(rgba8_to_unorm(dst), rgba8_to_unorm(src)));
    o[9] = unorm_to_rgba8(scale_premul_u8_to_unorm(src, bv));
    o[10] = coverage_to_u8(straight_channel(av, bv));
    o[11] = pack_premul_rgba8(pixel.r, pixel.g, pixel.b, pixel.a);
    o[12] = coverage_to_u8(rem_euclid_f32(c, 3.0f));
    o[13] = asuint(signum_f32(c));
}

// brush_words (validation/brush.hlsl:14-18).
extern "C" __global__ __launch_bounds__(256) void tk_k_brush_words(tk_args_brush_words a) {
    const uint32_t id = blockIdx.x * 256u + threadIdx.x;
    if (id >= a.count) return;
    const uint32_t *q = a.requests + (size_t)id * 4u;
    a.out[id] = sample_brush(a.paint, a.paint_words, a.paint_brush_base, q[0], asfloat(q[1]), asfloat(q[2]));
}

// ===========================================================================
// The differentiable form.
//
// Forward: the fine walk over FILL records with an FP32 premultiplied
// source-over composite, per-draw float colours and no RGBA8 quantization;
// coverage is computed exactly as in tk_k_fine.  Each (record, pixel)
// coverage is kept in `adj` for the backward pass.
//
// Backward, per pixel, with g = dL/dI and f_k = 1 - alpha_k * c_k.a:
//   I = sum_k alpha_k c_k T_k,  T_k = prod_{j>k} f_j,
//   dL/dalpha_k = T_k (g.c_k - c_k.a g.R_{k-1}),   R_{k-1} = composite below k,
//   dL/dc_k     = T_k alpha_k g  (minus T_k alpha_k g.R_{k-1} on the alpha channel).
// The forward keeps each pixel's product of the nonzero f_j over its whole
// stream (mantissa and exponent apart, so it never underflows) and the count of
// zero ones; the backward rebuilds R_{k-1} bottom-up and takes T_k as the ratio
// of that total to its running product (zero while a zero factor lies above k).  The
// coverage derivative (dL/dcoverage = dL/dalpha * dalpha/dcoverage) replaces
// the coverage in `adj`, and the line kernel turns it into endpoint gradients
// with the boundary form of the area derivative: a pixel's coverage is the
// integral of the winding number over it, so moving a line changes it by the
// integral, over the part of the line inside the pixel, of the motion's
// component along the line's normal.
// ===========================================================================

namespace {

// The FP32 alpha fill_alpha_at derives from a coverage (before the u8 step),
// with its derivative: NonZero min(|v|, 1), EvenOdd |v - 2 rint(v / 2)|.
__device__ __forceinline__ float diff_alpha(float v, uint32_t rule, float &slope) {
    if (rule == 1u) {
        const float w = v - 2.0f * rintf(0.5f * v);
        slope = w < 0.0f ? -1.0f : 1.0f;
        return fabsf(w);
    }
    const float m = fabsf(v);
    slope = m < 1.0f ? (v < 0.0f ? -1.0f : 1.0f) : 0.0f;
    return fminf(m, 1.0f);
}

template <int kCtrl, int kRowMask, int kBankMask>
__device__ __forceinline__ float dpp_add(float v, float src) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(src), kCtrl, kRowMask, kBankMask, true));
}

template <int kCtrl>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xf, 0xf, false));
}

// Sums four per-lane values over the wave at the cost of about one: two quad
// exchanges (quad_perm [1,0,3,2], then [2,3,0,1]) leave each lane with one of the
// four channels summed over its quad -- channel ((l & 1) << 1) | ((l >> 1) & 1) --
// row shifts by 4 and 8 sum the quads of a row into lanes 12-15, and two xor
// shuffles add the rows.  Lanes 12-15 of every row end with the wave totals.
__device__ __forceinline__ uint32_t sum4_channel(uint32_t lane) { return ((lane & 1u) << 1) | ((lane >> 1) & 1u); }
__device__ __forceinline__ float wave_sum4(float v0, float v1, float v2, float v3, uint32_t lane) {
    const bool odd = (lane & 1u) != 0u, hi = (lane & 2u) != 0u;
    const float x = (odd ? v2 : v0) + dpp_mov<0xB1>(odd ? v0 : v2);
    const float y = (odd ? v3 : v1) + dpp_mov<0xB1>(odd ? v1 : v3);
    float z = (hi ? y : x) + dpp_mov<0x4E>(hi ? x : y);
    z = dpp_add<0x114, 0xf, 0xe>(z, z);  // row_shr:4, banks 1-3
    z = dpp_add<0x118, 0xf, 0xc>(z, z);  // row_shr:8, banks 2-3
    z += __shfl_xor(z, 16, 64);
    z += __shfl_xor(z, 32, 64);
    return z;
}

// Product of the nonzero factors as mantissa * 2^exponent, and the zero count.
struct Transmit {
    float m;
    int e, zeros;
    __device__ __forceinline__ void mul(float f) {
        if (f == 0.0f) {
            ++zeros;
        } else {
            int x;
            m = frexpf(m * f, &x);
            e += x;
        }
    }
};

// A draw's paint in the differentiable form, 32 floats: {kind, x0, y0, x1, y1
// (the gradient line, pixels), stop count n (1-4), 4 stop offsets, 6 unused,
// 4 stops of premultiplied RGBA}.  Kinds: 0 solid (the first stop), 1 linear
// gradient, 2 radial gradient, 3 sweep gradient, 4 four-corner gradient.
// Gradients are evaluated exactly (no 8-bit ramp), in premultiplied space.  The
// first three take a position t and interpolate the stops linearly in t,
// constant beyond the end stops: t is, for a linear gradient, the projection
// of the pixel centre on the line, clamped to [0, 1]; for a radial one, the
// centre's distance from (x0, y0) over the line's length, clamped to [0, 1];
// for a sweep, the angle from the line's direction to the centre about
// (x0, y0), turning from +x towards +y, over a full turn, in [0, 1).  A
// four-corner gradient (n = 4) is bilinear over the box (x0, y0)-(x1, y1) with
// stops at its corners (x0, y0), (x1, y0), (x0, y1), (x1, y1), the position
// clamped to the box.  w holds each stop's weight in the colour.
constexpr uint32_t kDiffPaintFloats = 32u;
struct DiffPaint {
    float4 c;
    float w0, w1, w2, w3;
};
__device__ __forceinline__ DiffPaint diff_paint(const float *paint, uint32_t draw, float x, float y) {
    const float4 *r = reinterpret_cast<const float4 *>(paint + (size_t)draw * kDiffPaintFloats);
    const float4 h0 = r[0];
    DiffPaint p{r[4], 1.0f, 0.0f, 0.0f, 0.0f};
    if (h0.x == 0.0f) return p;
    const float4 h1 = r[1], h2 = r[2];
    const float dx = h0.w - h0.y, dy = h1.x - h0.z, den = dx * dx + dy * dy;
    const float px = x - h0.y, py = y - h0.z;
    float w0 = 0.0f, w1 = 0.0f, w2 = 0.0f, w3 = 0.0f;
    if (h0.x == 4.0f) {
        const float u = dx != 0.0f ? fminf(fmaxf(px / dx, 0.0f), 1.0f) : 0.0f;
        const float v = dy != 0.0f ? fminf(fmaxf(py / dy, 0.0f), 1.0f) : 0.0f;
        w0 = (1.0f - u) * (1.0f - v);
        w1 = u * (1.0f - v);
        w2 = (1.0f - u) * v;
        w3 = u * v;
    } else {
        float t = 0.0f;
        if (h0.x == 3.0f) {
            const float start = den > 0.0f ? sweep_atan2(dy, dx) : 0.0f;
            const float angle = px != 0.0f || py != 0.0f ? sweep_atan2(py, px) : 0.0f;
            t = (angle - start) / kGradientTau;
            t -= floorf(t);
        } else if (den > 0.0f) {
            t = h0.x == 2.0f ? sqrtf((px * px + py * py) / den) : (px * dx + py * dy) / den;
        }
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        const float n = h1.y, o0 = h1.z, o1 = h1.w, o2 = h2.x, o3 = h2.y;
        if (n <= 1.0f || t <= o0) {
            w0 = 1.0f;
        } else if (t < o1) {
            const float u = (t - o0) / (o1 - o0);
            w0 = 1.0f - u;
            w1 = u;
        } else if (n <= 2.0f) {
            w1 = 1.0f;
        } else if (t < o2) {
            const float u = (t - o1) / (o2 - o1);
            w1 = 1.0f - u;
            w2 = u;
        } else if (n <= 3.0f) {
            w2 = 1.0f;
        } else if (t < o3) {
            const float u = (t - o2) / (o3 - o2);
            w2 = 1.0f - u;
            w3 = u;
        } else {
            w3 = 1.0f;
        }
    }
    const float4 s0 = r[4], s1 = r[5], s2 = r[6], s3 = r[7];
    p.c = make_float4(w0 * s0.x + w1 * s1.x + w2 * s2.x + w3 * s3.x, w0 * s0.y + w1 * s1.y + w2 * s2.y + w3 * s3.y,
                      w0 * s0.z + w1 * s1.z + w2 * s2.z + w3 * s3.z, w0 * s0.w + w1 * s1.w + w2 * s2.w + w3 * s3.w);
    p.w0 = w0;
    p.w1 = w1;
    p.w2 = w2;
    p.w3 = w3;
    return p;
}

struct DiffTile {
    SceneView v;
    uint32_t column, tile, lane, gx, gy;
    bool inside;
};

template <bool kScenes>
__device__ __forceinline__ bool diff_tile(const tk_args_fine_diff &a, DiffTile &t) {
    const uint32_t linear = blockIdx.x + blockIdx.y * gridDim.x;
    const uint32_t columns = a.list.columns;
    if (linear >= xcd_grid(columns, a.xcd_run)) return false;
    const uint32_t column = xcd_chunked_block(linear, a.xcd_run);
    if (column >= columns) return false;
    t.column = column;
    const tk_fine_config &cfg = a.cfg;
    t.v = scene_view<kScenes>(a.scenes, column, columns, cfg.batch_count, cfg.tiles_width, cfg.tiles_height, cfg.width,
                              cfg.height, cfg.clear_color);
    t.tile = a.list.tiles ? a.list.tiles[column] : t.v.local;
    if (t.tile >= cfg.tile_count) return false;
    t.lane = threadIdx.x;
    t.gx = (t.tile % t.v.tw) * 16u + (t.lane & 15u);
    t.gy = (t.tile / t.v.tw) * 16u + (t.lane >> 4);
    t.inside = t.gx < t.v.width && t.gy < t.v.height;
    return true;
}

// The clip mask of the SDF clip layers open at a pixel: the product of their
// FP32 coverages (the mask fine keeps as combine_alpha of u8 coverages), with up
// to four levels; deeper nesting is flagged.
struct ClipMask {
    float m, s0, s1, s2, s3;
    uint32_t depth;
    // Selects on every slot (depth is uniform), so the stack stays in registers.
    __device__ __forceinline__ bool push(float coverage) {
        if (depth >= 4u) return false;
        s0 = depth == 0u ? m : s0;
        s1 = depth == 1u ? m : s1;
        s2 = depth == 2u ? m : s2;
        s3 = depth == 3u ? m : s3;
        ++depth;
        m *= coverage;
        return true;
    }
    __device__ __forceinline__ void pop() {
        if (depth == 0u) return;
        --depth;
        m = depth == 0u ? s0 : m;
        m = depth == 1u ? s1 : m;
        m = depth == 2u ? s2 : m;
        m = depth == 3u ? s3 : m;
    }
};

// Takes a clip record (BEGIN_SDF_CLIP, END_CLIP, and in the backward (kKept)
// BEGIN_CLIP) into the mask; false for a record the differentiable form does
// not take.  The forward evaluates an SDF clip's coverage here (a path clip's,
// which needs the tile's coverage accumulator, in for_each_fill) and keeps it in
// the record's slot of `adj`; the backward reads it from there.
template <bool kKept>
__device__ __forceinline__ bool clip_record(const tk_args_fine_diff &a, const DiffTile &t, uint32_t pi,
                                            const tk_ptcl_record &rec, ClipMask &clip) {
    if (rec.tag == TK_PTCL_END_CLIP) {
        clip.pop();
        return true;
    }
    if (rec.tag != TK_PTCL_BEGIN_SDF_CLIP && !(kKept && rec.tag == TK_PTCL_BEGIN_CLIP)) return false;
    const size_t slot = (size_t)pi * 256u + t.lane;
    float coverage = 0.0f;
    if constexpr (kKept) {
        if (pi < a.adj_records) coverage = a.adj[slot];
    } else {
        const FineDraw d = load_fine_draw(a.in.draws, a.cfg.draw_count, rec.color);
        coverage = sdf_coverage_from_draw(a.cfg, a.in.paint, d, (float)t.gx + 0.5f, (float)t.gy + 0.5f);
        if (pi < a.adj_records) a.adj[slot] = coverage;
    }
    return clip.push(coverage);
}

// Calls body(pi, rec, mask) for every FILL record of the tile's streams, bottom
// to top, with the pixel's clip mask; flags any record kind it does not take.
// kClips: the scene has clip layers (their records are taken into the mask; a
// path clip's coverage, its alpha under its fill rule, comes from
// path_clip(pi, rec), called by the whole workgroup).
template <bool kClips, class Body, class PathClip>
__device__ __forceinline__ void for_each_fill(const tk_args_fine_diff &a, const DiffTile &t, Body body,
                                              PathClip path_clip) {
    const tk_fine_config &cfg = a.cfg;
    for (uint32_t b = t.v.q0; b < t.v.q1; ++b) {
        const size_t idx = table_index(t.v, b);
        const uint32_t kind_word = a.in.tile_kinds[idx];
        if ((kind_word & TK_TILE_KIND_MASK) == TK_TILE_KIND_EMPTY) continue;
        const tk_segment_range r = a.in.tile_ptcl_ranges[idx];
        const uint32_t pend = min(r.end, cfg.ptcl_capacity);
        ClipMask clip{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 0u};
        for (uint32_t pi = r.start + (kind_word >> TK_TILE_PREFIX_SHIFT); pi < pend; ++pi) {
            const tk_ptcl_record rec = a.in.ptcl[pi];
            if (rec.tag == TK_PTCL_END) break;
            if (rec.tag == TK_PTCL_FILL && rec.color < cfg.draw_count) {
                body(pi, rec, clip.m);
            } else if (kClips && rec.tag == TK_PTCL_BEGIN_CLIP) {
                if (!clip.push(path_clip(pi, rec)) && t.lane == 0u && a.status) atomicOr(a.status, 1u);
            } else if (!kClips || !clip_record<false>(a, t, pi, rec, clip)) {
                if (t.lane == 0u && a.status) atomicOr(a.status, 1u);
            }
        }
    }
}

// kKeep: keep each record's coverage (a path clip's mask) in its slot of `adj`
// for the backward; otherwise the backward recomputes it from the segments.
template <bool kScenes, bool kClips, bool kKeep>
__device__ __forceinline__ void fine_diff_forward_tile(const tk_args_fine_diff &a) {
    __shared__ CoverageAcc cov_acc[2];
    DiffTile t;
    if (!diff_tile<kScenes>(a, t)) return;
    const uint32_t lane = t.lane;
    cov_acc[0].run[lane] = 0;
    cov_acc[0].part[lane] = 0;
    cov_acc[1].run[lane] = 0;
    cov_acc[1].part[lane] = 0;
    if (lane < 16u) {
        cov_acc[0].base[lane] = 0;
        cov_acc[1].base[lane] = 0;
    }
    uint32_t sel = 0u;
    __syncthreads();
    const float sx = (float)t.gx + 0.5f, sy = (float)t.gy + 0.5f;
    Px px = rgba8_to_unorm(t.v.clear);
    bool overflow = false;
    Transmit total{1.0f, 0, 0};
    const uint32_t tx = t.tile % t.v.tw, ty = t.tile / t.v.tw;
    // A path clip's mask: its coverage turned into alpha by its fill rule, kept
    // in the record's slot of `adj` for the backward.
    auto path_clip = [&](uint32_t pi, const tk_ptcl_record &rec) {
        const uint32_t s0 = rec.segment_start, s1 = min(rec.segment_end, a.cfg.segment_capacity);
        CoverageAcc &acc = cov_acc[sel];
        sel ^= 1u;
        accumulate_record(a.in.segments + s0, s1 > s0 ? s1 - s0 : 0u, acc, lane);
        __syncthreads();
        float slope;
        const float m = diff_alpha(take_coverage(acc, rec.backdrop, lane), rec.fill_rule, slope);
        if (kKeep) {
            if (pi < a.adj_records) a.adj[(size_t)pi * 256u + lane] = m;
            else overflow = true;
        }
        return m;
    };
    for_each_fill<kClips>(a, t, [&](uint32_t pi, const tk_ptcl_record &rec, float mask) {
        const uint32_t s0 = rec.segment_start, s1 = min(rec.segment_end, a.cfg.segment_capacity);
        const float4 c = diff_paint(a.colors, rec.color, sx, sy).c;
        // Issued before the coverage work so that its latency overlaps it.
        const uint4 box = reinterpret_cast<const uint4 *>(a.draw_tiles)[rec.color];
        CoverageAcc &acc = cov_acc[sel];
        sel ^= 1u;
        accumulate_record(a.in.segments + s0, s1 > s0 ? s1 - s0 : 0u, acc, lane);
        __syncthreads();
        const float cov = take_coverage(acc, rec.backdrop, lane);
        float slope;
        const float alpha = diff_alpha(cov, rec.fill_rule, slope) * mask;
        const float f = 1.0f - alpha * c.w;
        total.mul(f);
        px.r = px.r * f + alpha * c.x;
        px.g = px.g * f + alpha * c.y;
        px.b = px.b * f + alpha * c.z;
        px.a = px.a * f + alpha * c.w;
        if (kKeep) {
            if (pi < a.adj_records) a.adj[(size_t)pi * 256u + lane] = cov;
            else overflow = true;
        }
        if (lane == 0u && box.w != 0u) {  // where the line pass finds this path's record in this tile
            const uint32_t x0 = box.y & 0xFFFFu, y0 = box.y >> 16, x1 = box.z & 0xFFFFu, y1 = box.z >> 16;
            if (tx >= x0 && ty >= y0 && tx < x1 && ty < y1) {
                const uint32_t bd = box.x + (ty - y0) * (x1 - x0) + tx - x0;
                if (bd < a.rec_of_count) a.rec_of[bd] = pi;
            }
        }
    }, path_clip);
    if (overflow && lane == 0u && a.status) atomicOr(a.status, 2u);
    reinterpret_cast<uint4 *>(a.trans)[(size_t)t.column * 256u + lane] =
        make_uint4(__float_as_uint(total.m), (uint32_t)total.e, (uint32_t)total.zeros, 0u);
    if (t.inside)
        reinterpret_cast<float4 *>(a.image)[t.v.image + (size_t)t.gy * t.v.width + t.gx] =
            make_float4(px.r, px.g, px.b, px.a);
}

// kKeep: the coverages come from the forward's slots of `adj`; otherwise they
// are recomputed here from the tile's segments, exactly as the forward computed
// them (the exact sums do not depend on the segment order).
template <bool kScenes, bool kClips, bool kKeep>
__device__ __forceinline__ void fine_diff_backward_tile(const tk_args_fine_diff &a) {
    __shared__ CoverageAcc cov_acc[kKeep ? 1 : 2];
    DiffTile t;
    if (!diff_tile<kScenes>(a, t)) return;
    const uint32_t lane = t.lane;
    const float sx = (float)t.gx + 0.5f, sy = (float)t.gy + 0.5f;
    float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (t.inside) g = reinterpret_cast<const float4 *>(a.image_grad)[t.v.image + (size_t)t.gy * t.v.width + t.gx];
    auto coverage = [&](uint32_t pi) { return pi < a.adj_records ? a.adj[(size_t)pi * 256u + lane] : 0.0f; };
    const uint4 tw = reinterpret_cast<const uint4 *>(a.trans)[(size_t)t.column * 256u + lane];
    const Transmit total{__uint_as_float(tw.x), (int)tw.y, (int)tw.z};
    Transmit below{1.0f, 0, 0};
    Px res = rgba8_to_unorm(t.v.clear);
    auto record = [&](uint32_t pi, const tk_ptcl_record &rec, float cov, float mask) {
        const DiffPaint paint = diff_paint(a.colors, rec.color, sx, sy);
        const float4 c = paint.c;
        float slope;
        const float alpha = diff_alpha(cov, rec.fill_rule, slope) * mask;
        slope *= mask;
        const float f = 1.0f - alpha * c.w;
        below.mul(f);
        const float T = total.zeros > below.zeros ? 0.0f : ldexpf(total.m / below.m, total.e - below.e);
        const float g_res = g.x * res.r + g.y * res.g + g.z * res.b + g.w * res.a;
        const float g_c = g.x * c.x + g.y * c.y + g.z * c.z + g.w * c.w;
        const float d_alpha = T * (g_c - c.w * g_res);
        if (pi < a.adj_records) a.adj[(size_t)pi * 256u + lane] = d_alpha * slope;
        // dL/d(stop j) = w_j T alpha g (the alpha channel less T alpha w_j g.R); one
        // four-channel wave sum per stop the paint has (the stop count is uniform).
        const float w = T * alpha;
        const float gr = w * g.x, gg = w * g.y, gb = w * g.z, ga = w * (g.w - g_res);
        float *out = a.color_grad + (size_t)rec.color * kDiffPaintFloats + 16u + sum4_channel(lane);
        const float stops = a.colors[(size_t)rec.color * kDiffPaintFloats] == 0.0f
                                ? 1.0f
                                : a.colors[(size_t)rec.color * kDiffPaintFloats + 5u];
        const float sum0 = wave_sum4(paint.w0 * gr, paint.w0 * gg, paint.w0 * gb, paint.w0 * ga, lane);
        if ((lane & 63u) >= 60u) atomicAdd(out, sum0);
        if (stops > 1.0f) {
            const float sum1 = wave_sum4(paint.w1 * gr, paint.w1 * gg, paint.w1 * gb, paint.w1 * ga, lane);
            if ((lane & 63u) >= 60u) atomicAdd(out + 4, sum1);
        }
        if (stops > 2.0f) {
            const float sum2 = wave_sum4(paint.w2 * gr, paint.w2 * gg, paint.w2 * gb, paint.w2 * ga, lane);
            if ((lane & 63u) >= 60u) atomicAdd(out + 8, sum2);
        }
        if (stops > 3.0f) {
            const float sum3 = wave_sum4(paint.w3 * gr, paint.w3 * gg, paint.w3 * gb, paint.w3 * ga, lane);
            if ((lane & 63u) >= 60u) atomicAdd(out + 12, sum3);
        }
        res.r = res.r * f + alpha * c.x;
        res.g = res.g * f + alpha * c.y;
        res.b = res.b * f + alpha * c.z;
        res.a = res.a * f + alpha * c.w;
    };
    if constexpr (!kKeep) {
        cov_acc[0].run[lane] = 0;
        cov_acc[0].part[lane] = 0;
        cov_acc[1].run[lane] = 0;
        cov_acc[1].part[lane] = 0;
        if (lane < 16u) {
            cov_acc[0].base[lane] = 0;
            cov_acc[1].base[lane] = 0;
        }
        uint32_t sel = 0u;
        __syncthreads();
        auto cover = [&](const tk_ptcl_record &rec) {
            const uint32_t s0 = rec.segment_start, s1 = min(rec.segment_end, a.cfg.segment_capacity);
            CoverageAcc &acc = cov_acc[sel];
            sel ^= 1u;
            accumulate_record(a.in.segments + s0, s1 > s0 ? s1 - s0 : 0u, acc, lane);
            __syncthreads();
            return take_coverage(acc, rec.backdrop, lane);
        };
        auto path_clip = [&](uint32_t, const tk_ptcl_record &rec) {
            float slope;
            return diff_alpha(cover(rec), rec.fill_rule, slope);
        };
        for_each_fill<kClips>(a, t, [&](uint32_t pi, const tk_ptcl_record &rec, float mask) {
            record(pi, rec, cover(rec), mask);
        }, path_clip);
        return;
    }
    // The walk of for_each_fill, software-pipelined: the next record and its
    // coverage are loaded while the current record is processed.
    const uint32_t draws = a.cfg.draw_count;
    for (uint32_t b = t.v.q0; b < t.v.q1; ++b) {
        const size_t idx = table_index(t.v, b);
        const uint32_t kind_word = a.in.tile_kinds[idx];
        if ((kind_word & TK_TILE_KIND_MASK) == TK_TILE_KIND_EMPTY) continue;
        const tk_segment_range r = a.in.tile_ptcl_ranges[idx];
        const uint32_t pend = min(r.end, a.cfg.ptcl_capacity);
        uint32_t pi = r.start + (kind_word >> TK_TILE_PREFIX_SHIFT);
        if (pi >= pend) continue;
        tk_ptcl_record rec = a.in.ptcl[pi];
        float cov = coverage(pi);
        ClipMask clip{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 0u};
        while (rec.tag != TK_PTCL_END) {
            const uint32_t next = pi + 1u;
            const bool more = next < pend;
            tk_ptcl_record nrec{};
            if (more) nrec = a.in.ptcl[next];
            const float ncov = more ? coverage(next) : 0.0f;
            if (rec.tag == TK_PTCL_FILL && rec.color < draws) record(pi, rec, cov, clip.m);
            else if ((!kClips || !clip_record<true>(a, t, pi, rec, clip)) && lane == 0u && a.status)
                atomicOr(a.status, 1u);
            if (!more) break;
            pi = next;
            rec = nrec;
            cov = ncov;
        }
    }
}

}  // namespace

extern "C" __global__ __launch_bounds__(256) void tk_k_draw_tiles(tk_args_draw_tiles a) {
    const uint32_t d = blockIdx.x * 256u + threadIdx.x;
    if (d >= a.draw_count) return;
    const uint32_t path = a.draws[d].path_id;
    uint4 e = make_uint4(0u, 0u, 0u, 0u);
    if (path < a.path_count) {
        const tk_path_record &p = a.paths[path];
        e = make_uint4(p.data_offset, (p.tile_x0 & 0xFFFFu) | (p.tile_y0 << 16), (p.tile_x1 & 0xFFFFu) | (p.tile_y1 << 16),
                       1u);
    }
    reinterpret_cast<uint4 *>(a.table)[d] = e;
}

// One canvas or a batch of scenes (_scenes), without or with SDF clip layers
// (_clips: the clip evaluation raises the register count, so scenes without
// layers run the build that has none).
// _keep: the forward keeps every coverage for the backward; otherwise the
// backward recomputes it.
#define TK_DIFF_KERNELS(suffix, scenes, clips)                                                                  \
    extern "C" __global__ __launch_bounds__(256) void tk_k_fine_diff##suffix(tk_args_fine_diff a) {             \
        fine_diff_forward_tile<scenes, clips, false>(a);                                                         \
    }                                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void tk_k_fine_backward##suffix(tk_args_fine_diff a) {         \
        fine_diff_backward_tile<scenes, clips, false>(a);                                                        \
    }                                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void tk_k_fine_diff##suffix##_keep(tk_args_fine_diff a) {      \
        fine_diff_forward_tile<scenes, clips, true>(a);                                                          \
    }                                                                                                           \
    extern "C" __global__ __launch_bounds__(256) void tk_k_fine_backward##suffix##_keep(tk_args_fine_diff a) {  \
        fine_diff_backward_tile<scenes, clips, true>(a);                                                         \
    }
TK_DIFF_KERNELS(, false, false)
TK_DIFF_KERNELS(_scenes, true, false)
TK_DIFF_KERNELS(_clips, false, true)
TK_DIFF_KERNELS(_scenes_clips, true, true)
#undef TK_DIFF_KERNELS

// One thread per line: walks the pixels the line crosses inside its path's
// tile box, row slab by row slab and then column by column, and integrates
// dL/dcoverage along it.  For the part of the line with parameter s in
// [s_a, s_b] inside a pixel, the pixel's coverage moves by
//   -n . (dp0 * int (1 - s) ds + dp1 * int s ds),   n = (p0.y - p1.y, p1.x - p0.x),
// n pointing to the side where the winding number is one higher.
extern "C" __global__ __launch_bounds__(256) void tk_k_line_backward(tk_args_line_backward a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.line_count) return;
    const tk_line ln = a.lines[i];
    const float x0 = ln.p0[0], y0 = ln.p0[1], x1 = ln.p1[0], y1 = ln.p1[1];
    const float dx = x1 - x0, dy = y1 - y0;
    float g0 = 0.0f, g1 = 0.0f;
    if (ln.path_id < a.path_count) {
        const tk_path_record p = a.paths[ln.path_id];
        if (p.tile_x1 > p.tile_x0 && p.tile_y1 > p.tile_y0) {
            const int cx0 = (int)p.tile_x0 * 16, cx1 = (int)p.tile_x1 * 16 - 1;
            const int cy0 = (int)p.tile_y0 * 16, cy1 = (int)p.tile_y1 * 16 - 1;
            const uint32_t bw = p.tile_x1 - p.tile_x0;
            // Clamped in float before the conversions, so huge coordinates stay in range.
            const float ylo = fminf(y0, y1), yhi = fmaxf(y0, y1);
            const int r0 = max((int)floorf(fmaxf(ylo, (float)cy0 - 1.0f)), cy0);
            const int r1 = min(dy == 0.0f ? (int)floorf(fminf(ylo, (float)cy1 + 1.0f))
                                          : (int)ceilf(fminf(yhi, (float)cy1 + 2.0f)) - 1,
                               cy1);
            for (int r = r0; r <= r1; ++r) {
                float sa = 0.0f, sb = 1.0f;
                if (dy != 0.0f) {
                    const float ta = ((float)r - y0) / dy, tb = ((float)(r + 1) - y0) / dy;
                    sa = fmaxf(fminf(ta, tb), 0.0f);
                    sb = fminf(fmaxf(ta, tb), 1.0f);
                }
                if (!(sb > sa)) continue;
                const float xa = x0 + sa * dx, xb = x0 + sb * dx;
                const float xlo = fminf(xa, xb), xhi = fmaxf(xa, xb);
                const int c0 = max((int)floorf(fmaxf(xlo, (float)cx0 - 1.0f)), cx0);
                const int c1 = min(dx == 0.0f ? (int)floorf(fminf(xlo, (float)cx1 + 1.0f))
                                              : (int)ceilf(fminf(xhi, (float)cx1 + 2.0f)) - 1,
                                   cx1);
                const uint32_t row_base = p.data_offset + ((uint32_t)r / 16u - p.tile_y0) * bw;
                for (int c = c0; c <= c1; ++c) {
                    float ua = sa, ub = sb;
                    if (dx != 0.0f) {
                        const float ta = ((float)c - x0) / dx, tb = ((float)(c + 1) - x0) / dx;
                        ua = fmaxf(fminf(ta, tb), sa);
                        ub = fminf(fmaxf(ta, tb), sb);
                    }
                    if (!(ub > ua)) continue;
                    const uint32_t bd = row_base + (uint32_t)c / 16u - p.tile_x0;
                    if (bd >= a.rec_of_count) continue;
                    const uint32_t k = a.rec_of[bd];
                    if (k >= a.adj_records) continue;  // no record there (kNone), or beyond the buffer
                    const float gv = a.adj[(size_t)k * 256u + ((uint32_t)r & 15u) * 16u + ((uint32_t)c & 15u)];
                    const float i1 = 0.5f * (ub * ub - ua * ua);
                    g0 += gv * ((ub - ua) - i1);
                    g1 += gv * i1;
                }
            }
        }
    }
    const float nx = y0 - y1, ny = dx;
    float *out = a.line_grad + (size_t)i * 4u;
    out[0] = -nx * g0;
    out[1] = -ny * g0;
    out[2] = -nx * g1;
    out[3] = -ny * g1;
}

// The GPU front end of polygon scenes (tk_frame_set_polygons).
#include "frontend.hip"
