/*
 * oracle.h — CPU restatement of Tileink's compute-raster stages.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the gfx950
 * pipeline in libtileink_hip.so: only tests, the smoke check and the CPU
 * baseline leg of the benchmark load it.  Nothing in the product links it.
 *
 * Each function restates one reference shader entry or Rust routine, cited at
 * its definition in oracle.c.  It processes DrawBatches one at a time exactly as
 * the reference frame does (coarse -> fine per batch over an RGBA8 target).
 *
 * Records are passed as raw little-endian u32 words with the reference's GPU
 * ABI (Line 6 words, PathRecord 19, DrawRecord 31, LineSegment 5, PtclRecord 6,
 * LayerStackRecord 3, TileCoarseRecord 6).
 */
#ifndef TILEINK_ORACLE_H
#define TILEINK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scan_config {
    uint32_t clear_len, backdrop_len, path_count, scan_chunk_count;
    uint32_t line_count, segment_capacity, incremental, line_base;
    uint32_t path_base, chunk_base, backdrop_base;
} or_scan_config;

/* Scan stages (src/shaders/hlsl/scan/).  paths: 19 words each; lines: 6 words each. */
void or_scan_count(const or_scan_config *cfg, const uint32_t *lines, const uint32_t *paths,
                   uint32_t path_record_count, int32_t *backdrops, uint32_t *counts);
void or_scan_prefix_chunks(const or_scan_config *cfg, const uint32_t *chunks, const uint32_t *counts,
                           uint32_t *ranges, uint32_t *chunk_totals);
void or_scan_chunk_offsets(const or_scan_config *cfg, const uint32_t *paths, const uint32_t *chunk_ranges,
                           uint32_t *bumps, const uint32_t *chunk_totals, uint32_t *chunk_offsets);
void or_scan_apply_chunk_offsets(const or_scan_config *cfg, const uint32_t *chunks, uint32_t *ranges,
                                 uint32_t *cursors, const uint32_t *chunk_offsets);
void or_scan_emit(const or_scan_config *cfg, const uint32_t *lines, const uint32_t *paths,
                  uint32_t path_record_count, uint32_t *cursors, uint32_t *segments);
/* cumsum.hlsl: prefix_chunks, chunk_offsets, apply (all three). */
void or_cumsum(uint32_t row_count, uint32_t chunk_count, const uint32_t *chunk_backdrop_offsets,
               const uint32_t *chunk_lens, const uint32_t *row_starts, const uint32_t *row_ends,
               int32_t *backdrops);
/* line_scanned_tile_count (src/shared/scan_line.rs:17-31), Rust f32 semantics. */
uint32_t or_line_scanned_tile_count(const float p0[2], const float p1[2], uint32_t x0, uint32_t y0,
                                    uint32_t x1, uint32_t y1);

/* Coarse (src/shaders/hlsl/coarse/), one DrawBatch.  `work` holds TileCoarseRecord[6] x T
 * followed by PtclRecord[6] x ptcl_capacity; kinds: u32 x T. */
typedef struct or_coarse_config {
    uint32_t tile_count, tiles_width, tiles_height, batch_id;
    uint32_t layer_stack_start, layer_stack_end, ptcl_capacity, paint_brush_base;
    uint32_t draw_count, path_count, paint_words, layer_count, tile_draw_index_count;
} or_coarse_config;
typedef struct or_coarse_inputs {
    const uint32_t *draws, *paths;
    const int32_t *backdrops;
    const uint32_t *segment_ranges, *layers, *paint, *draw_batch_ids, *tile_draws, *tile_draw_indices;
} or_coarse_inputs;
void or_coarse_count_bins(const or_coarse_config *cfg, const or_coarse_inputs *in, uint32_t *work,
                          uint32_t *kinds);
/* coarse_prefix_chunks / chunk_offsets / apply over TileCoarseRecords (wrapping u32). */
void or_coarse_prefix(uint32_t tile_count, uint32_t *work, uint32_t *chunk_records /* 4 x chunks */);
void or_coarse_emit_bins(const or_coarse_config *cfg, const or_coarse_inputs *in, uint32_t *work, uint32_t *kinds);

/* Fine (src/shaders/hlsl/fine/), one DrawBatch over the whole target (load_target). */
typedef struct or_fine_config {
    uint32_t width, height, tile_count, tiles_width, tiles_height;
    uint32_t clip_spill_depth, paint_words, paint_sdf_shadow_base, paint_brush_base, draw_count;
    uint32_t ptcl_base_words; /* word offset of the PTCL array inside `work` */
    /* the image atlas image brushes sample: pages x height x width premultiplied
     * RGBA8 texels (NULL: image brushes sample 0) */
    const uint32_t *atlas; uint32_t atlas_width, atlas_height, atlas_pages;
} or_fine_config;
void or_fine(const or_fine_config *cfg, const uint32_t *draws, const uint32_t *paint, const uint32_t *segments,
             const uint32_t *work, const uint32_t *kinds, uint32_t *target);

/* Whole frame: scan, cumsum, then coarse + fine per batch. */
typedef struct or_scene {
    uint32_t width, height, tiles_width, tiles_height, clear_color;
    const uint32_t *lines; uint32_t line_count;
    const uint32_t *paths; uint32_t path_count;
    const uint32_t *draws; uint32_t draw_count;
    const uint32_t *paint; uint32_t paint_words, paint_sdf_shadow_base, paint_brush_base;
    const uint32_t *draw_batch_ids;
    const uint32_t *layers; uint32_t layer_count;
    const uint32_t *batches; uint32_t batch_count;
    const uint32_t *tile_draws, *tile_draw_indices; uint32_t tile_draw_index_count;
    const uint32_t *scan_chunks; uint32_t scan_chunk_count;
    const uint32_t *scan_chunk_ranges;
    const uint32_t *cumsum_chunk_offsets, *cumsum_chunk_lens; uint32_t cumsum_chunk_count;
    const uint32_t *cumsum_row_starts, *cumsum_row_ends; uint32_t cumsum_row_count;
    uint32_t backdrop_len, segment_capacity, max_clip_depth;
    const uint32_t *atlas; uint32_t atlas_width, atlas_height, atlas_pages;
} or_scene;

/* Optional outputs of or_render_frame; every pointer may be NULL.
 * Batch-major (b * T + t) tables; ptcl receives each batch's records back to
 * back (batch b at bases[b]), bases has batch_count + 1 entries. */
typedef struct or_frame_outputs {
    int32_t *raw_backdrops, *backdrops;
    uint32_t *segment_counts, *segment_ranges, *segment_cursors, *segment_bumps, *segments;
    uint32_t *tile_ptcl_counts, *tile_ptcl_starts, *tile_ptcl_ends, *tile_kinds, *batch_bases;
    uint32_t *ptcl; uint64_t ptcl_capacity_words;
    double *stage_ms; /* 6 wall-clock times: scan count, segment prefix, scan emit, backdrop cumsum,
                         coarse (all batches), fine (all batches) */
} or_frame_outputs;
/* Threads for the stages whose work items are independent (default 1); the
 * results do not depend on the count. */
void or_set_threads(uint32_t n);
uint32_t or_get_threads(void);

/* Returns 0 on success, nonzero when a buffer bound was exceeded. */
int or_render_frame(const or_scene *scene, uint32_t *image, or_frame_outputs *out);

/* Painter-ordered tile draw lists of an immediate canvas from its draws, batch
 * ids, batches, layers and paint (TileDrawBins::reset_transient), with early
 * occlusion when `cull` is nonzero.  tile_draws: 2 words {start, len} per tile
 * (start NONE for an empty list); indices: up to `capacity` draw ids.  Returns
 * the number of references, or -1 when they exceed `capacity`. */
int64_t or_tile_draw_bins(const or_scene *scene, int cull, uint32_t *tile_draws, uint32_t *indices,
                          uint64_t capacity);

/* Validation entries (validation/geometry.hlsl:22-29, validation/pixel.hlsl:9-34). */
/* or_fill_coverage_words: fill_alpha_at with each of its three sums exact (32.32
 * fixed point), independent of segment order; the fine stage uses this form.
 * or_fill_coverage_words_sequential: the FP32 sums in stored segment order, as
 * the reference's shader adds them. */
void or_fill_coverage_words(uint32_t count, const uint32_t *requests, const uint32_t *segments,
                            uint32_t *alpha_out, float *coverage_out);
void or_fill_coverage_words_sequential(uint32_t count, const uint32_t *requests, const uint32_t *segments,
                                       uint32_t *alpha_out, float *coverage_out);
void or_pixel_math_words(uint32_t count, const uint32_t *requests, uint32_t *out);
/* Text compositing of path glyphs, per request {destination, source, coverage,
 * foreground f32 bits}: four words out {src_over_mask_linear_auto_u8,
 * text_correcting_coverage(coverage, foreground) bits, srgb_to_linear(foreground)
 * bits, linear_to_srgb(foreground) bits}. */
void or_text_words(uint32_t count, const uint32_t *requests, uint32_t *out);
/* brush_words: sample_brush of the brush at word q[0] of the brush blob
 * (paint + paint_brush_base) at (bits q[1], bits q[2]); 4 request words each. */
void or_brush_words(const uint32_t *paint, uint32_t paint_words, uint32_t paint_brush_base, uint32_t count,
                    const uint32_t *requests, uint32_t *out);

/* FP64 area oracles.
 * or_fill_coverage_f64: the row-sweep quantity of fill_coverage_words in exact
 * FP64 — backdrop + left-edge terms + the exact signed area of each tile-local
 * segment inside pixel (x, y) of its tile.
 * or_polygon_pixel_area_f64: exact integral of the winding number of a closed
 * polygon (lines, 6 words each, absolute pixel coordinates) over the unit
 * pixel whose top-left corner is (px, py). */
void or_fill_coverage_f64(uint32_t count, const uint32_t *requests, const uint32_t *segments, double *out);
double or_polygon_pixel_area_f64(const uint32_t *lines, uint32_t line_count, double px, double py);

/* The differentiable form in FP64 (checker of tk_frame_render_diff and
 * tk_frame_backward).  lines: 4 doubles each (p0.x, p0.y, p1.x, p1.y); draw d
 * fills lines [line_start[d], line_start[d] + line_count[d]) with rule[d] and
 * the premultiplied colour colors[4d..4d+3]; images are 4 doubles per pixel.
 * or_diff_area_f64: the box (x0, y0, w, h) of one draw's lines on the canvas
 * and, when area is given, the integral of the winding number over each pixel
 * of the box (w * h doubles, row-major). */
void or_diff_area_f64(const double *lines, uint32_t start, uint32_t count, uint32_t width, uint32_t height,
                      int32_t *box, double *area);
void or_diff_render_f64(uint32_t width, uint32_t height, const double *clear, uint32_t draw_count,
                        const uint32_t *line_start, const uint32_t *line_count, const uint32_t *rule,
                        const double *colors, const double *lines, double *image);
void or_diff_backward_f64(uint32_t width, uint32_t height, const double *clear, uint32_t draw_count,
                          const uint32_t *line_start, const uint32_t *line_count, const uint32_t *rule,
                          const double *colors, const double *lines, const double *image_grad, double *line_grad,
                          double *color_grad);

/* The line pass alone: dL/d(line endpoints) of lines [start, start + count) of
 * one draw from dL/dcoverage over the draw's box (or_diff_area_f64's box, w * h
 * doubles), written to line_grad[4 * i ..] for each of those lines. */
void or_diff_line_grad_f64(const double *lines, uint32_t start, uint32_t count, uint32_t width, uint32_t height,
                           const double *coverage_grad, double *line_grad);

/* The FP32 SDF coverage of clip draw `draw_index` (draw records: 31 words each)
 * at every pixel centre of the canvas: the mask its clip layer applies. */
void or_sdf_draw_coverage(const uint32_t *draws, uint32_t draw_count, uint32_t draw_index, const uint32_t *paint,
                          uint32_t paint_words, uint32_t paint_sdf_shadow_base, uint32_t width, uint32_t height,
                          float *out);

#ifdef __cplusplus
}
#endif

#endif
