/*
 * tileink_hip.h — C ABI of the MI355X (gfx950) compute-raster backend.
 *
 * This library replaces the GPU half of Tileink's native renderer: the scan
 * (path_count / segment emit), backdrop prefix, coarse per-tile command list
 * (PTCL) and fine analytic-coverage stages, plus the host-side scene recording
 * that feeds them.  Every entry point takes plain pointers and sizes; device
 * pointers are HIP device addresses owned by the caller unless stated.
 *
 * Record layouts are byte-for-byte the reference's GPU ABI so that a Rust
 * `Device::Hip` adapter can hand its existing upload buffers straight through:
 *   Line                 src/shared/line.rs:4-9,           scene_records.hlsli:10-12
 *   PathRecord           src/shared/path.rs:6-23,          scene_records.hlsli:5-15
 *   DrawRecord           src/shared/draw_record.rs:54-84,  draw_records.hlsli:5-21
 *   LineSegment          src/shared/line_seg.rs:3-9
 *   GpuScanChunk(+Range) src/shared/gpu_plan/path_plans.rs:15-29
 *   LayerStackRecord     src/shared/gpu_coarse.rs:49-53
 *   PtclRecord           src/shared/gpu_coarse.rs:57-64, coarse_records.hlsli:14
 *   TileDrawRecord       src/shared/gpu_coarse.rs:42-46
 *   ScanConfig           src/shaders/hlsl/scan/config.hlsli:4-8
 *   CumsumConfig         src/shaders/hlsl/cumsum.hlsl:5
 *
 * Errors: every function returns tk_status; on failure tk_last_error() returns
 * a thread-local message.  No C++ exception crosses this boundary.
 */
#ifndef TILEINK_HIP_H
#define TILEINK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status                                                                     */
/* ------------------------------------------------------------------------- */

typedef enum tk_status {
    TK_OK = 0,
    TK_ERR_INVALID = 1,     /* bad argument / inconsistent scene metadata        */
    TK_ERR_HIP = 2,         /* a HIP runtime call failed                         */
    TK_ERR_OOM = 3,         /* device or host allocation failed                  */
    TK_ERR_UNAVAILABLE = 4, /* no gfx950 device at the requested ordinal         */
    TK_ERR_OVERFLOW = 5,    /* a capacity bound was exceeded; writes were dropped */
} tk_status;

/* Thread-local description of the last failure on this thread.
 * Replaces NativeError's Display (src/native/context.rs:16-25). */
const char *tk_last_error(void);

/* Library version word: (major << 16) | minor. */
uint32_t tk_version(void);

/* ------------------------------------------------------------------------- */
/* Constants (src/shaders/hlsl/constants.hlsli and shared/ tag headers)        */
/* ------------------------------------------------------------------------- */

#define TK_TILE_SIZE 16u
#define TK_SCAN_CHUNK_SIZE 256u
#define TK_CUMSUM_CHUNK_SIZE 256u
#define TK_SDF_RECORD_WORDS 17u
#define TK_INVALID_INDEX 0xFFFFFFFFu

#define TK_DRAW_BRUSH 0u
#define TK_DRAW_CLIP 1u
#define TK_DRAW_PATH_GLYPH 5u
#define TK_LAYER_CLIP 0u
#define TK_FILL_NONZERO 0u
#define TK_FILL_EVEN_ODD 1u

#define TK_PTCL_END 0u
#define TK_PTCL_FILL 1u
#define TK_PTCL_COLOR 2u
#define TK_PTCL_BEGIN_CLIP 3u
#define TK_PTCL_END_CLIP 4u
#define TK_PTCL_SDF 9u
#define TK_PTCL_BEGIN_SDF_CLIP 12u

#define TK_TILE_KIND_INTERPRETER 0u
#define TK_TILE_KIND_EMPTY 1u
#define TK_TILE_KIND_COLOR 2u
#define TK_TILE_KIND_SDF 3u
#define TK_TILE_KIND_MIXED 4u
#define TK_TILE_KIND_MASK 7u
#define TK_TILE_PREFIX_SHIFT 3u

#define TK_BRUSH_SOLID 1u
#define TK_BRUSH_LINEAR 2u
#define TK_BRUSH_RADIAL 3u
#define TK_BRUSH_SWEEP 4u
#define TK_BRUSH_FOUR_CORNER 5u
#define TK_BRUSH_PATTERN 6u            /* a pattern without an image: transparent */
#define TK_BRUSH_IMAGE 7u              /* BRUSH_PATTERN_RESOURCE: an image of the scene's atlas */
#define TK_PATTERN_NEAREST 0u
#define TK_PATTERN_BILINEAR 1u
#define TK_BRUSH_HEADER_WORDS 9u
#define TK_BRUSH_PARAM_WORDS 12u
#define TK_SDF_KIND_RECT 1u

/* ------------------------------------------------------------------------- */
/* GPU record ABI                                                              */
/* ------------------------------------------------------------------------- */

typedef struct tk_affine { float a, b, c, d, e, f; } tk_affine; /* (x,y)->(a x + c y + e, b x + d y + f) */

typedef struct tk_line {           /* 24 B */
    uint32_t path_id;
    float pad;
    float p0[2];
    float p1[2];
} tk_line;

typedef struct tk_path_record {    /* 76 B */
    uint32_t path_id, line_count, line_start, flags;
    uint32_t data_offset, data_len;
    uint32_t tile_x0, tile_y0, tile_x1, tile_y1;
    uint32_t segment_start, segment_capacity, segment_count;
    tk_affine transform;
} tk_path_record;

typedef struct tk_pixel_bounds { int32_t x0, y0, x1, y1; } tk_pixel_bounds;

typedef struct tk_draw_record {    /* 124 B */
    uint32_t path_id, glyph_run_id, sdf_offset, sdf_len;
    uint32_t sdf_shadow_offset, sdf_shadow_len, brush_offset, brush_len;
    uint32_t tag, fill_rule;
    tk_pixel_bounds pixel_bounds, local_pixel_bounds;
    uint32_t solid_rect;
    tk_affine transform, inverse_transform;
} tk_draw_record;

typedef struct tk_line_segment { float p0x, p0y, p1x, p1y, y_edge; } tk_line_segment; /* 20 B */
typedef struct tk_segment_range { uint32_t start, end; } tk_segment_range;          /* 8 B  */
typedef struct tk_scan_chunk { uint32_t path_id, backdrop_offset, segment_start, len; } tk_scan_chunk;
typedef struct tk_scan_chunk_range { uint32_t start, end; } tk_scan_chunk_range;
typedef struct tk_layer_record { uint32_t tag, draw, payload; } tk_layer_record;   /* 12 B */
typedef struct tk_ptcl_record {                                                     /* 24 B */
    uint32_t tag;
    int32_t backdrop;
    uint32_t fill_rule, segment_start, segment_end, color;
} tk_ptcl_record;
typedef struct tk_tile_draw_record { uint32_t start, len; } tk_tile_draw_record;   /* 8 B  */

/* One fused DrawBatch: a painter-ordered run of draws sharing one layer stack
 * snapshot (src/shared/execution.rs:106-113).  The draws of batch b are those
 * whose draw_batch_ids entry equals b. */
typedef struct tk_batch_record { uint32_t layer_start, layer_end; } tk_batch_record;

typedef struct tk_scan_config {    /* scan/config.hlsli:4-8 */
    uint32_t clear_len, backdrop_len, path_count, scan_chunk_count;
    uint32_t line_count, segment_capacity, incremental, line_base;
    uint32_t path_base, chunk_base, backdrop_base;
} tk_scan_config;

typedef struct tk_cumsum_config {  /* cumsum.hlsl:5 */
    uint32_t row_count, chunk_count, pad0, pad1;
} tk_cumsum_config;

/* ------------------------------------------------------------------------- */
/* Scene: everything the GPU stages read, as produced by the recording layer. */
/* The host view is owned by whoever filled it (a tk_canvas, or the caller).  */
/* ------------------------------------------------------------------------- */

typedef struct tk_scene {
    uint32_t width, height;          /* physical pixels */
    uint32_t tiles_width, tiles_height;
    uint32_t clear_color;            /* premultiplied RGBA8, little-endian u32 */

    const tk_line *lines;            uint32_t line_count;
    const tk_path_record *paths;     uint32_t path_count;
    const tk_draw_record *draws;     uint32_t draw_count;
    const uint32_t *paint;           uint32_t paint_words;   /* sdf || sdf_shadow || brush */
    uint32_t paint_sdf_shadow_base, paint_brush_base;

    const uint32_t *draw_batch_ids;  /* draw_count entries; TK_INVALID_INDEX outside batches */
    const tk_layer_record *layers;   uint32_t layer_count;
    const tk_batch_record *batches;  uint32_t batch_count;

    const tk_tile_draw_record *tile_draws;   /* tiles_width*tiles_height */
    const uint32_t *tile_draw_indices;       uint32_t tile_draw_index_count;

    const tk_scan_chunk *scan_chunks;        uint32_t scan_chunk_count;
    const tk_scan_chunk_range *scan_chunk_ranges; /* path_count */
    const uint32_t *cumsum_chunk_offsets;    /* backdrop offset of each cumsum chunk */
    const uint32_t *cumsum_chunk_lens;
    uint32_t cumsum_chunk_count;
    const uint32_t *cumsum_row_starts;       /* first chunk of each backdrop row */
    const uint32_t *cumsum_row_ends;
    uint32_t cumsum_row_count;

    uint32_t backdrop_len;           /* B = sum of path data_len */
    uint32_t segment_capacity;       /* S bound = sum of path segment_capacity */
    uint32_t max_clip_depth;         /* deepest fused clip stack of any batch */
    uint64_t ptcl_capacity;          /* PTCL records bound over all batches */

    /* The image atlas image brushes sample (the Texture2DArray the reference
     * samples, src/shaders/hlsl/shared/brush/pattern.hlsli:18-55): atlas_pages
     * pages of atlas_width x atlas_height premultiplied RGBA8 texels, page-major,
     * then row-major; an image brush's header holds its placement (x, y, page,
     * width, height in words 2-6, as resource_brush_header patches it,
     * src/shared/gpu_brush.rs:92-127).  NULL with zero pages when no draw has one. */
    const uint32_t *atlas;
    uint32_t atlas_width, atlas_height, atlas_pages, atlas_reserved;
} tk_scene;

/* ------------------------------------------------------------------------- */
/* Recording (host).  Mirrors the Canvas API (src/canvas/primitives/ and       */
/* src/canvas/layers.rs) for the primitives on the compute-raster path.        */
/* ------------------------------------------------------------------------- */

typedef struct tk_canvas tk_canvas;

/* Path element verbs, kurbo::PathEl order.  Points per verb: 1,1,2,3,0. */
#define TK_PATH_MOVE 0u
#define TK_PATH_LINE 1u
#define TK_PATH_QUAD 2u
#define TK_PATH_CUBIC 3u
#define TK_PATH_CLOSE 4u

typedef struct tk_gradient_stop { float offset; float rgba[4]; } tk_gradient_stop; /* straight sRGB */

/* Brush (src/shared/brush.rs:33-80): solid colour, or a gradient over sorted stops
 * (peniko::Gradient), or a four-corner gradient.  Fields a kind does not use are ignored. */
typedef struct tk_brush {
    uint32_t kind;                 /* TK_BRUSH_SOLID, _LINEAR, _RADIAL, _SWEEP or _FOUR_CORNER */
    float rgba[4];                 /* solid: straight-alpha sRGB in [0,1] */
    float start[2], end[2];        /* linear: gradient line; radial: start and end centres;
                                      sweep: start[] is the centre (user space) */
    uint32_t extend;               /* 0 pad, 1 repeat, 2 reflect */
    const tk_gradient_stop *stops; uint32_t stop_count;
    float start_radius, end_radius;  /* radial */
    float start_angle, end_angle;    /* sweep, radians (peniko SweepGradientPosition) */
    float bounds[4];                 /* four-corner: x0, y0, x1, y1 (user space) */
    float corners[4][4];             /* four-corner: top-left, top-right, bottom-right,
                                        bottom-left, straight-alpha sRGB */
    /* image (TK_BRUSH_IMAGE, peniko's image brush as PatternBrush, src/shared/brush.rs:
     * 520-583): an image of the canvas (tk_canvas_add_image), the transform taking
     * user space to the image's unit square [a b c d e f] (x' = a x + c y + e,
     * y' = b x + d y + f), TK_PATTERN_NEAREST or _BILINEAR, opacity 0-255;
     * extend as above */
    uint32_t image;
    float pattern_transform[6];
    uint32_t sampling, opacity;
    /* linear and radial: gradient_flags bit 0 (TK_GRADIENT_TRANSFORM) takes
     * gradient_transform [a b c d e f] from user space to the gradient's own space
     * (peniko's brush transform, e.g. the inverse of an SVG gradientTransform; the
     * identity without the bit); bit 1 (TK_GRADIENT_UNPREMULTIPLIED) interpolates
     * the stops in straight alpha (InterpolationAlphaSpace::Unpremultiplied, as SVG
     * gradients do; premultiplied without it). */
    float gradient_transform[6];
    uint32_t gradient_flags;
} tk_brush;
#define TK_GRADIENT_TRANSFORM 1u
#define TK_GRADIENT_UNPREMULTIPLIED 2u

tk_status tk_canvas_create(uint32_t width, uint32_t height, double scale, tk_canvas **out);
void tk_canvas_destroy(tk_canvas *canvas);
/* Canvas::push_path (src/canvas/primitives/paths.rs:50-60). transform may be NULL (identity). */
tk_status tk_canvas_push_path(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count,
                              const double *points, size_t point_count, const tk_brush *brush,
                              const double *transform, uint32_t fill_rule, double tolerance,
                              uint32_t *draw_id);
/* Canvas::push_stroke (src/canvas/primitives/paths.rs:24-48): round joins and caps. */
tk_status tk_canvas_push_stroke(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count,
                                const double *points, size_t point_count, double width,
                                const tk_brush *brush, const double *transform, uint32_t fill_rule,
                                double tolerance, uint32_t *draw_id);
/* Many closed polygons with solid brushes in one call (no counterpart in the
 * reference, which records path by path): polygon i is counts[i] points of
 * `points` (x, y pairs, laid end to end) recorded as push_path of
 * move_to + line_to... + close_path with straight-alpha colour rgba[4i..4i+3]
 * and fill_rules[i]; draws get consecutive ids from *first_draw_id. */
tk_status tk_canvas_push_polygons(tk_canvas *canvas, const double *points, const uint32_t *counts,
                                  uint32_t polygon_count, const float *rgba, const uint32_t *fill_rules,
                                  uint32_t *first_draw_id);
/* Canvas::push_stroke with a full kurbo::Stroke (src/canvas/primitives/paths.rs:24-48):
 * join and cap styles in kurbo's order (Join {Bevel, Miter, Round}, Cap {Butt,
 * Square, Round}); a miter join longer than miter_limit x width becomes a bevel.
 * Dashes are not supported.  tk_canvas_push_stroke is this with round joins and
 * caps (kurbo::Stroke::new). */
#define TK_JOIN_BEVEL 0u
#define TK_JOIN_MITER 1u
#define TK_JOIN_ROUND 2u
#define TK_CAP_BUTT 0u
#define TK_CAP_SQUARE 1u
#define TK_CAP_ROUND 2u
typedef struct tk_stroke_style {
    double width, miter_limit;
    uint32_t join, start_cap, end_cap, reserved;
} tk_stroke_style;
tk_status tk_canvas_push_stroke_styled(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count,
                                       const double *points, size_t point_count, const tk_stroke_style *style,
                                       const tk_brush *brush, const double *transform, uint32_t fill_rule,
                                       double tolerance, uint32_t *draw_id);
/* The outlines of a laid-out text run as one path with text compositing
 * (Canvas::push_text_layout_as_path, src/canvas/primitives/text.rs:104-146, after
 * its shaping: a nonzero path draw tagged DRAW_PATH_GLYPH, composited by
 * src_over_mask_linear_auto_u8, fine/text/auto.hlsli:28-33).  The caller passes
 * the outline; an empty one records nothing (*draw_id = TK_INVALID_INDEX). */
tk_status tk_canvas_push_text_outline(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count,
                                      const double *points, size_t point_count, const tk_brush *brush,
                                      const double *transform, double tolerance, uint32_t *draw_id);
/* Canvas::push_rect (src/canvas/primitives/rectangles.rs:33-43); radii = tl,tr,bl,br. */
tk_status tk_canvas_push_rect(tk_canvas *canvas, double x0, double y0, double x1, double y1,
                              const float radii[4], const tk_brush *brush, uint32_t *draw_id);
/* Canvas::register_scene_image (src/canvas/primitives/images.rs:44-54): an image of
 * premultiplied RGBA8 pixels (Image, src/shared/image.rs:9-19; row-major, red in the
 * low byte), copied; *image_id names it in brushes.  A zero-sized image is refused. */
tk_status tk_canvas_add_image(tk_canvas *canvas, uint32_t width, uint32_t height, const uint32_t *premul_rgba8,
                              uint32_t *image_id);
/* Canvas::push_image (src/canvas/primitives/images.rs:9-26): the image scaled into the
 * rect (PatternBrush::for_rect_resource, src/shared/brush.rs:557-583) as an SDF rect
 * draw.  An empty or non-finite rect records nothing (*draw_id = TK_INVALID_INDEX). */
tk_status tk_canvas_push_image(tk_canvas *canvas, double x0, double y0, double x1, double y1, uint32_t image,
                               uint32_t extend, uint32_t sampling, uint32_t *draw_id);
/* Canvas::push_clip_sdf_rect_layer (src/canvas/layers.rs:34-40). */
tk_status tk_canvas_push_clip_sdf_rect_layer(tk_canvas *canvas, double x0, double y0, double x1,
                                             double y1, const float radii[4]);
/* Canvas::push_clip_layer (src/canvas/layers.rs:7-27): pixel-aligned axis rects become SDF clips. */
tk_status tk_canvas_push_clip_layer(tk_canvas *canvas, const uint8_t *verbs, size_t verb_count,
                                    const double *points, size_t point_count,
                                    const double *transform, uint32_t fill_rule, double tolerance);
/* Canvas::pop_layer (src/canvas/layers.rs:280-284). TK_ERR_INVALID when no layer is open. */
tk_status tk_canvas_pop_layer(tk_canvas *canvas);
/* Compile + plan: ExecPlan batches (src/canvas/compile.rs:39-70), tile bins
 * (src/shared/gpu_plan/tile_bins/build.rs:4-83), scan/cumsum chunk plans
 * (src/shared/gpu_plan/path_plans.rs:463-504).  The view stays valid until the
 * next push or destroy. */
tk_status tk_canvas_encode(tk_canvas *canvas, uint32_t clear_color, tk_scene *out);
/* Flattened line count of the canvas so far (diagnostic). */
uint32_t tk_canvas_line_count(const tk_canvas *canvas);

/* CPU restatement of line_scanned_tile_count (src/shared/scan_line.rs:17-31),
 * used for segment capacities. */
uint32_t tk_line_scanned_tile_count(const tk_line *line, uint32_t tile_x0, uint32_t tile_y0,
                                    uint32_t tile_x1, uint32_t tile_y1);

/* ------------------------------------------------------------------------- */
/* Device context                                                              */
/* ------------------------------------------------------------------------- */

typedef struct tk_context tk_context;

/* Explicit device ordinal; never falls back (NativeContextOptions::physical_adapter,
 * src/native/context.rs:4-13).  Loads the embedded gfx950 code object. */
tk_status tk_context_create(int device, tk_context **out);
void tk_context_destroy(tk_context *ctx);
/* Stream every entry point of this context is ordered on (a hipStream_t). */
void *tk_context_stream(tk_context *ctx);
tk_status tk_context_synchronize(tk_context *ctx);

/* Limits (src/native/runtime/renderer/recording.rs:17-23). */
typedef struct tk_limits {
    uint32_t image_dimension;
    uint32_t max_tiles;
    uint64_t device_memory_bytes;
    uint32_t compute_units;
} tk_limits;
tk_status tk_context_limits(tk_context *ctx, tk_limits *out);

/* Device buffers (Adapter::allocate_buffer / readback, src/native/runtime/adapter.rs:159-251).
 * Copies are ordered on the context stream and complete before returning. */
tk_status tk_device_alloc(tk_context *ctx, size_t bytes, void **out);
tk_status tk_device_free(tk_context *ctx, void *ptr);
tk_status tk_copy_to_device(tk_context *ctx, void *dst, const void *src, size_t bytes);
tk_status tk_copy_to_host(tk_context *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain 16-byte-per-lane copy kernel (tk_k_copy);
 * bytes and both pointers must be multiples of 16.  Its measured HBM traffic
 * calibrates the FETCH_SIZE / WRITE_SIZE counters against a known byte count. */
tk_status tk_copy_device(tk_context *ctx, void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------------------- */
/* Whole-frame path (NativeRenderer::render_to_image, src/native/renderer.rs:177-188) */
/* ------------------------------------------------------------------------- */

typedef struct tk_frame tk_frame;

/* Allocates every device buffer for the scene and uploads its inputs. */
tk_status tk_frame_create(tk_context *ctx, const tk_scene *scene, tk_frame **out);
/* Re-uploads inputs of a scene with identical sizes (animation frames). */
tk_status tk_frame_update(tk_frame *frame, const tk_scene *scene);
void tk_frame_destroy(tk_frame *frame);
/* Enqueues scan -> backdrop prefix -> coarse -> fine on the context stream.
 * Inputs and outputs stay in device memory. */
tk_status tk_frame_render(tk_frame *frame);
/* Incremental render of the listed tiles only (row-major tile indices, as the
 * reference's DamageTiles worklist, src/render/damage_tiles.rs:14-20): the
 * active_tile_list / incremental dispatch of scan, coarse and fine
 * (scan/index.hlsli:5-8, coarse/active_tiles.hlsli:6-9, fine/main.hlsl:26-27).
 * Scan takes only the paths whose tile box meets a listed tile; coarse and fine
 * take only the listed tiles.  Pixels of other tiles keep the values of earlier
 * renders, so after tk_frame_update with a scene that differs from the rendered
 * one only inside the listed tiles the image equals a full render of the new
 * scene.  Duplicates are ignored; an index >= tiles_width * tiles_height is
 * TK_ERR_INVALID.  Afterwards the per-(batch, tile) tables read back through
 * tk_frame_read_buffer hold one column per listed tile, in list order. */
tk_status tk_frame_render_tiles(tk_frame *frame, const uint32_t *tiles, uint32_t tile_count);
/* Launch parameters of a frame's kernels (no counterpart in the reference, whose
 * dispatch shapes are fixed).
 *  - *_xcd_run: how many consecutive workgroups (scan: line blocks; coarse:
 *    workgroups of coarse_block tiles; fine and the differentiable kernels:
 *    tiles) one XCD takes before the next XCD's run: a power of two up to 1024,
 *    or 0 for the plain launch order.
 *  - persistent: bit 0 (TK_PERSISTENT_COARSE) runs coarse count and emit, bit 1
 *    (TK_PERSISTENT_FINE) runs fine, as persistent launches: as many workgroups
 *    as are resident at once, each wave (coarse) or workgroup (fine) taking
 *    tiles from eight per-XCD work queues (each XCD's queue holds the runs the
 *    plain launch would deal it; an XCD whose queue is used up takes from the
 *    others) until all are used up.
 *  - scan_block: lines per scan count / emit workgroup: 64, 128, 256, 512, 1024.
 *  - coarse_block: tiles (one 64-lane wave each) per coarse workgroup: 1, 2, 4.
 *  - coarse_scan_items: counts per workgroup of the coarse exclusive scan
 *    (decoupled lookback): 1024, 2048, 4096, 8192.
 *  - diff_keep_coverage: 1: the differentiable forward keeps every (record,
 *    pixel) coverage for the backward (4 bytes per pixel per PTCL record
 *    written, then read); 0 (the default): the backward recomputes the
 *    coverages from the tile segments, exactly as the forward computed them.
 * The tile is 16x16 pixels: the record formats the reference's GPU ABI fixes
 * (tile-local segment coordinates in [0, 16], 256-tile scan chunks, the PTCL)
 * and the bit-exact contract with its results depend on it, so it is not a
 * tuning choice; nor is fine's workgroup, one lane per pixel of a tile.
 * tk_frame_autotune renders the frame's own scene with candidate values (one
 * warm-up and `repeats` timed renders each, stage times from HIP events), and
 * searches one parameter at a time with the others held at the best so far;
 * it returns the result in *chosen.  No choice changes a result.             */
#define TK_PERSISTENT_COARSE 1u
#define TK_PERSISTENT_FINE 2u
typedef struct tk_tuning {
    uint32_t scan_xcd_run, coarse_xcd_run, fine_xcd_run;
    uint32_t persistent;
    uint32_t scan_block, coarse_block, coarse_scan_items;
    uint32_t diff_keep_coverage;
} tk_tuning;
/* GPU front end for polygon scenes.  The reference records every path on the
 * CPU (Canvas::push_path_inner_with_tag, src/canvas/paths.rs:65-156, with the
 * scan plans of src/shared/gpu_plan/path_plans.rs:156-228,463-504 and the tile
 * bins of src/shared/gpu_plan/tile_bins/build.rs:4-83 and transient.rs:7-146);
 * this entry moves the geometry-dependent part of that work onto the device for
 * scenes whose paths are closed polygons (recorded with tk_canvas_push_polygons,
 * or move/line/close paths of one contour each).
 * It replaces the vertices of every path of the frame (or batch frame) and
 * rebuilds on the device exactly what recording the same polygons and encoding
 * would: the lines (degenerate edges dropped), the path records (bounds, tile
 * boxes, backdrop slabs, segment capacities), the path draws' pixel bounds, the
 * painter-ordered tile bins minus the tiles early occlusion hides, the scan and
 * cumsum chunk plans and the PTCL bound.  Draws, paints, layers and batches stay
 * as recorded.  The frame's buffers grow when the new geometry needs more.
 *   points:      device, (x, y) FP32 pairs in physical pixels, path after path;
 *   starts:      device, path_count + 1 offsets into the vertices;
 *   point_count: the number of vertices;
 *   line_vertex: optional device output, point_count entries: for each line
 *                (in order), the index of the vertex it starts from.
 * One host synchronisation sizes the buffers.  A non-finite vertex is
 * TK_ERR_INVALID, after which the frame must receive valid vertices before it
 * renders.  Any captured graph is dropped. */
tk_status tk_frame_set_polygons(tk_frame *frame, const float *points, const uint32_t *starts, uint64_t point_count,
                                uint32_t *line_vertex);
/* The same for scenes whose paths are each one closed contour of cubics (a move,
 * cubics, a close: tk_canvas_push_path with those verbs), flattened on the
 * device: PathFlatten::flatten (src/shared/path_flatten.rs:20-56) through the
 * recorder's own flattener (flatten.h, the same FP64 arithmetic on both sides),
 * with the curves' exact bounds as the paths' bounds.  ctrl: device, FP64 (x, y)
 * pairs in physical pixels; path p's are ctrl[starts[p] .. starts[p+1]): the
 * start point, then the two control points and the end point of each cubic.
 * tolerance: the flattening tolerance in physical pixels (the recorder's
 * tolerance times the canvas scale).  Two host synchronisations. */
tk_status tk_frame_set_cubic_paths(tk_frame *frame, const double *ctrl, const uint32_t *starts, double tolerance);

tk_status tk_frame_get_tuning(tk_frame *frame, tk_tuning *out);
tk_status tk_frame_set_tuning(tk_frame *frame, const tk_tuning *tuning);
tk_status tk_frame_autotune(tk_frame *frame, uint32_t repeats, tk_tuning *chosen);

/* The differentiable form (no counterpart in the reference, which only renders).
 * Its forward pass is the fine stage (fine/main.hlsl:23-40) with the RGBA8 round
 * trips between batches left out: an FP32 premultiplied source-over composite of
 * the scene's FILL records.  Scenes: path draws (fills and strokes) inside any
 * nesting of clip layers, SDF or path; a clip multiplies the alpha of the draws
 * inside it by its FP32 coverage (a path clip's: its exact-area alpha under its
 * fill rule).  Each draw's paint is paint[32 * draw ..] (FP32), not its brush:
 *   {kind: 0 solid, 1 linear, 2 radial, 3 sweep, 4 four-corner gradient;
 *    x0, y0, x1, y1: the gradient line in pixels; n: stop count 1-4; 4 stop
 *    offsets; 6 unused; 4 stops of premultiplied RGBA},
 * a solid paint being its first stop.  The first three kinds interpolate the
 * stops linearly in t, clamped to [0, 1]: t = the projection of the pixel centre
 * on the line (linear), its distance from (x0, y0) over the line's length
 * (radial), or its angle about (x0, y0) from the line's direction, turning from
 * +x to +y, over a full turn (sweep).  A four-corner gradient (n = 4) is
 * bilinear over the box (x0, y0)-(x1, y1), stops at (x0, y0), (x1, y0),
 * (x0, y1), (x1, y1).
 * All float pointers are device pointers.
 * tk_frame_render_diff: the whole pipeline; writes `image` (4 floats per pixel,
 *   row-major; a batch's scenes at their image offsets, one after another).
 * tk_frame_backward: after tk_frame_render_diff; from image_grad = dL/d(image)
 *   writes line_grad = dL/d(p0.x, p0.y, p1.x, p1.y) for every line of the scene
 *   and paint_grad = dL/d(paint record) for every draw: 32 floats, its stops'
 *   colours at 16-31 (both overwritten).
 * tk_frame_diff_status: bit 0 when a record the form does not take was met.   */
tk_status tk_frame_render_diff(tk_frame *frame, const float *paint, float *image);
tk_status tk_frame_backward(tk_frame *frame, const float *paint, const float *image_grad, float *line_grad,
                            float *paint_grad);
tk_status tk_frame_diff_status(tk_frame *frame, uint32_t *flags);

/* enable != 0: tk_frame_render captures its launch sequence (both streams) into a
 * hipGraph on first use and replays it as one launch afterwards; renders with
 * timing or raw-backdrop capture enabled, and tk_frame_render_tiles, still launch
 * kernel by kernel.  tk_frame_update drops the graph (it is captured again).  No
 * counterpart in the reference: its passes are recorded into one command buffer
 * per frame (src/native/runtime/compute.rs:33-331). */
tk_status tk_frame_set_graph(tk_frame *frame, int enable);

/* A batch of independent scenes rendered as one frame: the scenes (each with its
 * own canvas size, clear colour, draws, clip layers and tile bins) are laid end
 * to end and go through one launch sequence of scan, coarse and fine, whose
 * coarse and fine kernels read each tile's canvas from a per-scene slot.  Each
 * scene's image equals its own tk_frame_render's.  The batched form of the
 * pipeline, for many small scenes; no counterpart in the reference, which
 * renders one canvas per submission. */
typedef struct tk_batch tk_batch;
tk_status tk_batch_create(tk_context *ctx, const tk_scene *scenes, uint32_t scene_count, tk_batch **out);
tk_status tk_batch_render(tk_batch *batch);
/* Copies scene `scene`'s RGBA8 image (its width x height pixels, row-major). */
tk_status tk_batch_read_image(tk_batch *batch, uint32_t scene, uint32_t *rgba_out, size_t pixel_count);
/* The frame behind the batch: stats, stage times, timing and graph switches.
 * tk_frame_update and tk_frame_render_tiles do not apply to it. */
tk_frame *tk_batch_frame(tk_batch *batch);
void tk_batch_destroy(tk_batch *batch);
/* Synchronises and copies the RGBA8 image (premultiplied, row-major, tightly
 * packed little-endian u32, src/native/submission.rs:125-146). */
tk_status tk_frame_read_image(tk_frame *frame, uint32_t *rgba_out, size_t pixel_count);

/* Asynchronous submission with a completion receipt: NativeRenderer's submit and
 * NativeImageSubmission::{is_complete, readback} (src/native/submission.rs:100-146)
 * over the adapter's Receipt (src/native/runtime/adapter.rs:17-47, 159-250).
 * tk_frame_submit enqueues a render of the frame's current scene and a copy of its
 * image and status words into pinned host memory owned by the submission, then
 * returns without waiting.  Later renders of the same frame do not disturb an
 * earlier submission's image.  A batch's frame is refused (read its scenes with
 * tk_batch_read_image). */
typedef struct tk_submission tk_submission;
tk_status tk_frame_submit(tk_frame *frame, tk_submission **out);
/* *complete = 1 once the submission's work has finished, else 0; never waits.
 * A device error is TK_ERR_HIP (the reference's Unconfirmed). */
tk_status tk_submission_is_complete(tk_submission *submission, uint32_t *complete);
/* Waits for the submission, then copies its image.  A frame that exceeded a
 * capacity is TK_ERR_OVERFLOW (the reference's Rejected); the image is not copied. */
tk_status tk_submission_readback(tk_submission *submission, uint32_t *rgba_out, size_t pixel_count);
/* Waits for the submission if it is still running, then frees it. */
void tk_submission_destroy(tk_submission *submission);
/* Device address of the RGBA8 image. */
void *tk_frame_image_device(tk_frame *frame);

/* Intermediate buffers, for parity checks. */
typedef enum tk_frame_buffer {
    TK_BUF_BACKDROPS = 0,      /* i32 x B: after the row prefix                     */
    TK_BUF_SEGMENT_COUNTS = 1, /* u32 x B: scan_count per backdrop tile            */
    TK_BUF_SEGMENT_RANGES = 2, /* {start,end} x B                                   */
    TK_BUF_SEGMENT_CURSORS = 3,/* u32 x B: after emit = range.end                   */
    TK_BUF_SEGMENT_BUMPS = 4,  /* u32 x P                                           */
    TK_BUF_SEGMENTS = 5,       /* tk_line_segment x S                               */
    TK_BUF_TILE_PTCL_COUNTS = 6,   /* u32 x Q*T (batch-major)                       */
    TK_BUF_TILE_PTCL_RANGES = 7,   /* {start,end} x Q*T, per-batch offsets          */
    TK_BUF_TILE_KINDS = 8,         /* u32 x Q*T: kind | skip << 3                   */
    TK_BUF_BATCH_PTCL_BASES = 9,   /* u32 x Q: first record of each batch           */
    TK_BUF_PTCL = 10,              /* tk_ptcl_record x K                            */
    TK_BUF_RAW_BACKDROPS = 11,     /* i32 x B: backdrop deltas before the row prefix */
    /* The scene as the frame holds it (as uploaded, or as tk_frame_set_polygons rebuilt it): */
    TK_BUF_LINES = 12,             /* tk_line x L                                   */
    TK_BUF_PATHS = 13,             /* tk_path_record x P                            */
    TK_BUF_DRAWS = 14,             /* tk_draw_record x D                            */
    TK_BUF_TILE_DRAWS = 15,        /* tk_tile_draw_record x T                       */
    TK_BUF_TILE_DRAW_INDICES = 16, /* u32 x R                                       */
    TK_BUF_SCAN_CHUNKS = 17,       /* tk_scan_chunk x scan_chunk_count              */
    TK_BUF_SCAN_CHUNK_RANGES = 18, /* tk_scan_chunk_range x P                       */
    TK_BUF_CUMSUM_CHUNK_OFFSETS = 19, /* u32 x cumsum_chunk_count                   */
    TK_BUF_CUMSUM_CHUNK_LENS = 20,    /* u32 x cumsum_chunk_count                   */
    TK_BUF_CUMSUM_ROW_STARTS = 21,    /* u32 x cumsum_row_count                     */
    TK_BUF_CUMSUM_ROW_ENDS = 22,      /* u32 x cumsum_row_count                     */
} tk_frame_buffer;
tk_status tk_frame_read_buffer(tk_frame *frame, tk_frame_buffer which, void *out, size_t bytes);

typedef struct tk_frame_stats {
    uint32_t paths, lines, draws, batches, tiles;   /* P, L, D, Q, T */
    uint32_t backdrop_len;                          /* B */
    uint32_t segment_capacity;                      /* S bound */
    uint32_t tile_draw_refs;                        /* R */
    uint64_t ptcl_capacity;                         /* K bound */
    uint64_t ptcl_records;                          /* K after the last render */
    uint32_t segments_emitted;                      /* S after the last render */
    uint32_t overflow;                              /* nonzero when a bound was exceeded */
    uint64_t device_bytes;                          /* device memory held by the frame */
    uint32_t scan_chunk_count, cumsum_chunk_count, cumsum_row_count, pad;
    uint64_t ptcl_bound;                            /* the scene's PTCL bound (<= ptcl_capacity) */
} tk_frame_stats;
/* Synchronises; ptcl_records / segments_emitted / overflow come from the device. */
tk_status tk_frame_stats_read(tk_frame *frame, tk_frame_stats *out);

/* When enabled, later renders keep a copy of the backdrop deltas before the
 * row prefix (TK_BUF_RAW_BACKDROPS). */
tk_status tk_frame_set_capture(tk_frame *frame, int enable);

/* Per-stage event timings of the last tk_frame_render, milliseconds. */
typedef struct tk_stage_times {
    float scan_clear, scan_count, scan_prefix, scan_emit, cumsum;
    float coarse_count, coarse_prefix, coarse_emit, fine, total;
} tk_stage_times;
/* Enables hipEvent timestamps around each stage of later renders. */
tk_status tk_frame_enable_timing(tk_frame *frame, int enable);
tk_status tk_frame_stage_times(tk_frame *frame, tk_stage_times *out);

/* One-call convenience: upload, render, read back, release. */
tk_status tk_render_to_image(tk_context *ctx, const tk_scene *scene, uint32_t *rgba_out,
                             size_t pixel_count);

/* ------------------------------------------------------------------------- */
/* Stage entry points (device pointers, ordered on the context stream).       */
/* Each restates one reference dispatch contract so the reference's raw        */
/* ComputeBatch known-answer cases can be replayed unchanged.                   */
/* ------------------------------------------------------------------------- */

/* scan_clear (scan/clear.hlsl:16-32), non-incremental form. Any pointer may be NULL. */
tk_status tk_scan_clear(tk_context *ctx, const tk_scan_config *config, void *backdrops,
                        void *segment_ranges, void *segment_counts, void *segment_cursors,
                        void *segment_bumps, void *chunk_totals, void *chunk_offsets);
/* scan_count (scan/count.hlsl:13-40): per-line DDA, atomic tile counts and
 * backdrop deltas. path_record_count bounds path ids (GetDimensions guard). */
tk_status tk_scan_count(tk_context *ctx, const tk_scan_config *config, const tk_line *lines,
                        const tk_path_record *paths, uint32_t path_record_count,
                        int32_t *backdrops, uint32_t *segment_counts);
/* scan_prefix_chunks + scan_chunk_offsets + scan_apply_chunk_offsets
 * (scan/prefix_chunks.hlsl:16-38, chunk_offsets.hlsl:14-26,
 * apply_chunk_offsets.hlsl:15-27) as one decoupled-lookback pass.  Writes the
 * same five outputs the three reference passes leave behind. */
tk_status tk_scan_prefix(tk_context *ctx, const tk_scan_config *config,
                         const tk_scan_chunk *chunks, const tk_scan_chunk_range *chunk_ranges,
                         const tk_path_record *paths, const uint32_t *segment_counts,
                         tk_segment_range *segment_ranges, uint32_t *segment_cursors,
                         uint32_t *segment_bumps, uint32_t *chunk_totals, uint32_t *chunk_offsets);
/* scan_emit (scan/emit.hlsl:14-28): writes tile-clipped segments at atomically
 * claimed cursor slots. */
tk_status tk_scan_emit(tk_context *ctx, const tk_scan_config *config, const tk_line *lines,
                       const tk_path_record *paths, uint32_t path_record_count,
                       uint32_t *segment_cursors, tk_line_segment *segments);
/* cumsum_prefix_chunks / chunk_offsets / apply_chunk_offsets (cumsum.hlsl:16-71):
 * in-place inclusive row prefix of backdrop deltas. */
tk_status tk_cumsum(tk_context *ctx, const tk_cumsum_config *config,
                    const uint32_t *chunk_backdrop_offsets, const uint32_t *chunk_lens,
                    const uint32_t *row_chunk_starts, const uint32_t *row_chunk_ends,
                    int32_t *backdrops, uint32_t *chunk_totals, uint32_t *chunk_offsets);

/* Coarse, all batches at once (replaces Q runs of coarse_count_bins ->
 * coarse_prefix_* -> coarse_emit_bins, coarse/count.hlsl:48-73,
 * coarse/prefix.hlsl:11-63, coarse/emit.hlsl:82-142). */
typedef struct tk_coarse_config {
    uint32_t tile_count, tiles_width, tiles_height, batch_count;
    uint32_t draw_count, path_count, paint_words, paint_brush_base;
    uint32_t layer_count, tile_draw_index_count, pad0, pad1;
    uint64_t ptcl_capacity;
} tk_coarse_config;
typedef struct tk_coarse_inputs {
    const tk_draw_record *draws;
    const tk_path_record *paths;
    const int32_t *backdrops;             /* after the row prefix */
    const tk_segment_range *segment_ranges;
    const tk_layer_record *layers;
    const tk_batch_record *batches;
    const uint32_t *paint;
    const uint32_t *draw_batch_ids;
    const tk_tile_draw_record *tile_draws;
    const uint32_t *tile_draw_indices;
} tk_coarse_inputs;
/* tile_ptcl_ranges are global record ranges: the stream of (batch b, tile t)
 * is ptcl[start, end).  The reference's per-batch tile offsets
 * (ptcl_start/end of TileCoarseRecord) are start - batch_ptcl_bases[b]. */
typedef struct tk_coarse_outputs {
    uint32_t *tile_ptcl_counts;           /* Q*T, batch-major (b*T + t) */
    tk_segment_range *tile_ptcl_ranges;   /* Q*T, global wrapping exclusive prefix */
    uint32_t *tile_kinds;                 /* Q*T: kind | opaque-prefix skip << 3 */
    uint32_t *batch_ptcl_bases;           /* Q+1: first record of each batch, total K last */
    tk_ptcl_record *ptcl;                 /* ptcl_capacity records */
    uint32_t *status;                     /* 4 words, zeroed by the caller: [0] overflow,
                                             [1] K, [3] tile list not in batch order */
} tk_coarse_outputs;
tk_status tk_coarse(tk_context *ctx, const tk_coarse_config *config, const tk_coarse_inputs *in,
                    const tk_coarse_outputs *out);

/* tk_coarse without device allocation.  It has no counterpart in the reference;
 * it is for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Inputs,
 * outputs and results are those of tk_coarse.  The coarse scan's lookback ticket
 * and status words (tk_coarse's only scratch, which it grows with hipMalloc)
 * live instead in one fixed array of TK_NO_ALLOC_SCRATCH_BYTES per device,
 * allocated by the HIP runtime when the library is loaded and held until the
 * process ends; each call clears the part it uses with a fill kernel (no
 * hipMemsetAsync).  batch_count * tile_count may not exceed
 * TK_NO_ALLOC_MAX_TABLE; larger calls fail with TK_ERR_INVALID and launch
 * nothing.  Every call on one GPU shares that GPU's array: calls that overlap
 * on one device (other streams, other contexts, other threads) race. */
#define TK_NO_ALLOC_SCRATCH_BYTES (1u << 20)
#define TK_NO_ALLOC_TICKET_BYTES 256u
#define TK_NO_ALLOC_MAX_TABLE 536739840ull /* ((1 MiB - 256) / 8) scan workgroups x 4096 counts */
tk_status tk_coarse_no_alloc(tk_context *ctx, const tk_coarse_config *config, const tk_coarse_inputs *in,
                             const tk_coarse_outputs *out);

/* Fine, all batches at once: one 256-lane workgroup per 16x16 tile walks the
 * batches in order, re-quantising to RGBA8 at each batch boundary exactly where
 * the reference stores and re-loads its RGBA8 target (fine/main.hlsl:23-40,
 * fine/interpreter.hlsli:21-219, fine/specialized.hlsli:19-111).  Coverage keeps
 * fill_alpha_at's three sums and their FP32 combination (base + running) +
 * partial (shared/coverage.hlsli:34-51), but each sum is exact: its FP32 terms
 * are added as 32.32 fixed point, so the result does not depend on the atomic
 * slot order in which emit stored a tile's segments. */
typedef struct tk_fine_config {
    uint32_t width, height, clear_color, tile_count;
    uint32_t tiles_width, tiles_height, batch_count, clip_spill_depth;
    uint32_t paint_words, paint_sdf_shadow_base, paint_brush_base, draw_count;
    uint32_t ptcl_capacity, segment_capacity;   /* records beyond these are never read */
} tk_fine_config;
typedef struct tk_fine_inputs {
    const tk_draw_record *draws;
    const uint32_t *paint;
    const tk_line_segment *segments;
    const tk_segment_range *tile_ptcl_ranges;
    const uint32_t *tile_kinds;
    const uint32_t *batch_ptcl_bases;
    const tk_ptcl_record *ptcl;
    uint32_t *status;  /* optional; bit 1 set when a tile's clip stack went deeper than
                          4 + clip_spill_depth levels (the extra levels were dropped) */
    const uint32_t *atlas;  /* tk_scene's image atlas (NULL: image brushes sample 0) */
    uint32_t atlas_width, atlas_height, atlas_pages, atlas_reserved;
} tk_fine_inputs;
tk_status tk_fine(tk_context *ctx, const tk_fine_config *config, const tk_fine_inputs *in,
                  uint32_t *rgba_out);

/* Validation helpers (validation/geometry.hlsl:22-29, validation/pixel.hlsl:9-34).
 * requests: 8 words each {segment_start, segment_end, backdrop, x, y, rule, 0, 0}.
 * alpha_out: u8 alpha per request; coverage_out (optional): FP32 winding area
 * backdrop + row sweep before the fill rule, with the three sums exact as in
 * tk_fine, evaluated per pixel (independently of tk_fine's row accumulation). */
tk_status tk_fill_coverage_words(tk_context *ctx, uint32_t count, const uint32_t *requests,
                                 const tk_line_segment *segments, uint32_t *alpha_out,
                                 float *coverage_out);
/* requests: 8 words {a, b, src, dst, c_bits, t_bits, rule, 0}; 14 output words each. */
tk_status tk_pixel_math_words(tk_context *ctx, uint32_t count, const uint32_t *requests,
                              uint32_t *out);
/* brush_words (validation/brush.hlsl:14-18): sample_brush of the encoded brush at
 * word data_base of the brush blob (paint + paint_brush_base) at point (x, y).
 * requests: 4 words each {data_base, x_bits, y_bits, 0}; one premultiplied RGBA8
 * word out per request. */
tk_status tk_brush_words(tk_context *ctx, const uint32_t *paint, uint32_t paint_words,
                         uint32_t paint_brush_base, uint32_t count, const uint32_t *requests,
                         uint32_t *out);

#ifdef __cplusplus
}
#endif

#endif /* TILEINK_HIP_H */
