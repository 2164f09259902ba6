// Kernel argument blocks shared by the gfx950 code object and the host launcher.
// Each kernel takes exactly one of these structs by value; the host passes it as
// the single entry of hipModuleLaunchKernel's kernelParams array.
#pragma once

#include <stdint.h>

#include "tileink_hip.h"

#define TK_FINE_MAX_CLIP_SPILL 60u    // clip stack levels beyond the 4 register slots, held in LDS
#define TK_COARSE_SCAN_ITEMS 4096u    // counts per workgroup of the coarse exclusive scan

// Runs of consecutive workgroups kept on one XCD (0: launch order).  Fine: tiles;
// coarse: 4-tile workgroups; scan: 256-line workgroups.
#ifndef TK_XCD_CHUNK_FINE
#define TK_XCD_CHUNK_FINE 256u
#endif
#ifndef TK_XCD_CHUNK_COARSE
#define TK_XCD_CHUNK_COARSE 64u
#endif
#ifndef TK_XCD_CHUNK_SCAN
#define TK_XCD_CHUNK_SCAN 0u
#endif

struct tk_args_fill {
    uint32_t *dst;
    uint64_t count;
    uint32_t value;
};

struct tk_args_copy {
    const uint4 *src;
    uint4 *dst;
    uint64_t count;  // 16-byte words
};

struct tk_args_scan_clear {
    tk_scan_config cfg;
    int32_t *backdrops;
    tk_segment_range *ranges;
    uint32_t *counts;
    uint32_t *cursors;
    uint32_t *bumps;
    uint32_t *chunk_totals;
    uint32_t *chunk_offsets;
};

struct tk_args_scan_lines {
    tk_scan_config cfg;
    const tk_line *lines;
    const tk_path_record *paths;
    uint32_t path_record_count;
    int32_t *backdrops;          // count only
    uint32_t *counts_or_cursors; // count: segment counts, emit: cursors
    tk_line_segment *segments;   // emit only
    const uint32_t *path_active; // incremental: lines of paths whose flag is 0 are skipped; null: all
    uint32_t xcd_run;            // XCD run length in blocks (a power of two; 0: launch order)
    uint32_t hash_slots;         // emit: workgroup key table entries (a power of two), in dynamic LDS
    uint32_t stage_slots;        // emit: staged segments per workgroup, in dynamic LDS
};

// A batch of independent scenes rendered as one frame: each scene keeps its own
// canvas and tile grid, its tiles take the table columns [column_start,
// column_start + tiles_width * tiles_height), its draws use the batches
// [batch_start, batch_end), and its (batch, tile) tables form one block of
// (batch_end - batch_start) x columns entries at table_base, batch-major.
struct tk_scene_slot {
    uint32_t width, height, tiles_width, tiles_height;
    uint32_t clear_color, batch_start, batch_end, column_start;
    uint64_t table_base, image_offset;  // image_offset: the scene's first pixel in the RGBA8 output
};
struct tk_scene_slots {
    const tk_scene_slot *slots;      // null: one scene, described by the stage configs
    const uint32_t *scene_of_column;
    uint32_t count;
    uint64_t table_entries;          // sum of the scenes' table blocks
};

// Incremental rendering over a list of damaged tiles (the reference's
// active_tile_list): the per-(batch, tile) tables hold one column per listed
// tile, in list order, instead of one per tile of the canvas.
struct tk_tile_list {
    const uint32_t *tiles;  // column -> tile index; null: column == tile
    uint32_t columns;       // list length, or tile_count
};

// Per path: 1 when the path's tile box meets a damaged tile (damage_sat is the
// summed-area table of the damage bitmap, (tiles_width + 1) x (tiles_height + 1)).
struct tk_args_active_paths {
    const tk_path_record *paths;
    uint32_t path_count, tiles_width, tiles_height;
    const uint32_t *damage_sat;
    uint32_t *path_active;
};

struct tk_args_scan_prefix {
    tk_scan_config cfg;
    const tk_scan_chunk *chunks;
    const tk_scan_chunk_range *chunk_ranges;
    const tk_path_record *paths;
    const uint32_t *counts;
    tk_segment_range *ranges;
    uint32_t *cursors;
    uint32_t *bumps;
    uint32_t *chunk_totals;
    uint32_t *chunk_offsets;
    uint32_t wave_count;  // pass A: max(chunks, ceil(paths / 64)) waves
};

struct tk_args_cumsum {
    tk_cumsum_config cfg;
    const uint32_t *chunk_offsets_in;  // backdrop offset per chunk
    const uint32_t *chunk_lens;
    const uint32_t *row_starts;
    const uint32_t *row_ends;
    int32_t *backdrops;
    uint32_t *chunk_totals;
    uint32_t *chunk_carries;
};

// Per-tile classes of the distinct (draw, tag) pairs among the layer records, one
// byte per (table column, pair), column-major: bit 0 the layer rejects the tile,
// bit 1 it counts as a wrapper, bit 2 it emits BEGIN/END records, bit 3 it is an
// SDF clip.  Null when the coarse kernels evaluate each layer themselves.
struct tk_layer_classes {
    const uint8_t *bits;        // columns * pair_count (one row of pair_count bytes per table column)
    const uint32_t *pair_of;    // per layer record: its pair
    uint32_t pair_count;
};

struct tk_args_coarse {
    tk_coarse_config cfg;
    tk_coarse_inputs in;
    tk_coarse_outputs out;
    unsigned long long *scan_status;  // per scan block
    uint32_t *ticket;
    uint32_t epoch;
    uint32_t scan_blocks;
    tk_layer_classes classes;
    tk_tile_list list;
    tk_scene_slots scenes;
    uint32_t xcd_run;  // XCD run length in four-tile workgroups (a power of two; 0: launch order)
    uint32_t *queue;   // 16 per-XCD work counters (count, emit), zero at launch: persistent count and emit
                       // launches, each wave taking tiles; null: one four-tile workgroup per launch item
};

struct tk_args_layer_classes {
    tk_coarse_config cfg;
    tk_coarse_inputs in;
    const uint32_t *pair_layer;  // per pair: a layer record with that (draw, tag)
    uint32_t pair_count;
    uint8_t *bits;               // columns * pair_count, column-major
    tk_tile_list list;
};

// Per-draw brush words for fine: two words per draw, {1, colour} when the draw's
// brush colour does not depend on the pixel, {0, 0} when it must be sampled.
struct tk_args_draw_brush {
    const tk_draw_record *draws;
    uint32_t draw_count;
    const uint32_t *paint;
    uint32_t paint_words, paint_brush_base;
    uint32_t *table;
};

struct tk_args_fine {
    tk_fine_config cfg;
    tk_fine_inputs in;
    uint32_t *rgba;
    const uint32_t *draw_brush;  // tk_args_draw_brush::table, draw_count entries
    tk_tile_list list;
    tk_scene_slots scenes;
    uint32_t xcd_run;  // XCD run length in tiles (a power of two; 0: launch order)
    uint32_t *queue;   // eight per-XCD work counters, zero at launch: a persistent launch; null: one workgroup per tile
};

struct tk_args_fill_coverage {
    uint32_t count;
    const uint32_t *requests;
    const tk_line_segment *segments;
    uint32_t *alpha_out;
    float *coverage_out;
};

struct tk_args_pixel_math {
    uint32_t count;
    const uint32_t *requests;
    uint32_t *out;
};

struct tk_args_brush_words {
    uint32_t count;
    const uint32_t *requests;  // {data_base, x_bits, y_bits, 0}
    const uint32_t *paint;
    uint32_t paint_words, paint_brush_base;
    uint32_t *out;
};

// The differentiable form: an FP32 composite of FILL records with per-draw
// premultiplied colours and no RGBA8 quantization, and its backward pass.
struct tk_args_fine_diff {
    tk_fine_config cfg;
    tk_fine_inputs in;
    const tk_path_record *paths;
    uint32_t path_count;
    uint32_t rec_of_count;      // backdrop tiles (the length of rec_of)
    const float *colors;        // 32-float paint record per draw (solid or linear gradient)
    float *image;               // forward: premultiplied RGBA per pixel
    const float *image_grad;    // backward: dL/d(image) per pixel
    float *color_grad;          // backward: dL/d(paint record) per draw (its stop colours), accumulated
    float *adj;                 // 256 floats per PTCL record: the forward's coverage, then dL/d(coverage)
    uint64_t adj_records;
    uint32_t *trans;            // 4 words per pixel (256 per column): the forward's transmittance product
                                // over the pixel's whole stream {mantissa bits, exponent, zero factors, 0}
    uint32_t *rec_of;           // forward: per backdrop tile, the PTCL record of its path there
    const uint32_t *draw_tiles; // forward: 4 words per draw, tk_args_draw_tiles::table
    uint32_t *status;           // bit 0: a record kind the differentiable form does not take; bit 1: adj too small
    tk_tile_list list;
    tk_scene_slots scenes;
    uint32_t xcd_run;           // as tk_args_fine
};

struct tk_args_line_backward {
    const tk_line *lines;
    uint32_t line_count, path_count;
    const tk_path_record *paths;
    const uint32_t *rec_of;
    uint32_t rec_of_count;
    const float *adj;
    uint64_t adj_records;
    float *line_grad;           // dL/d(p0.x, p0.y, p1.x, p1.y) per line
};

// Per draw, where its path's backdrop tiles start and its tile box:
// {data_offset, tile_x0 | tile_y0 << 16, tile_x1 | tile_y1 << 16, 1} (0 in the
// last word: no path).
struct tk_args_draw_tiles {
    const tk_draw_record *draws;
    uint32_t draw_count, path_count;
    const tk_path_record *paths;
    uint32_t *table;
};

// GPU front end of polygon scenes (tk_frame_set_polygons): one scene of the
// frame as the front end sees it.
struct tk_fe_scene {
    uint32_t tiles_width, tiles_height;
    uint32_t column_start;  // its first column of the frame's tile tables
    uint32_t row_start;     // its first tile row among all scenes' rows
    uint32_t draw_start, draw_end;
};

// Per-path values the front end counts, then scans in place (column-major,
// path_count words per column).
enum tk_fe_column : uint32_t {
    TK_FE_LINES = 0,
    TK_FE_BACKDROPS = 1,
    TK_FE_SEGMENTS = 2,
    TK_FE_SCAN_CHUNKS = 3,
    TK_FE_CUMSUM_ROWS = 4,
    TK_FE_CUMSUM_CHUNKS = 5,
    TK_FE_PATH_COLUMNS = 6,
};
// Totals of a front-end pass (u64 words): the path columns, then these.
enum tk_fe_total : uint32_t {
    TK_FE_TILE_REFS = 6,
    TK_FE_ROW_ENTRIES = 7,
    TK_FE_PTCL = 8,
    TK_FE_TOTALS = 9,
};

struct tk_args_frontend {
    const float *points;              // x, y of every vertex, physical pixels (FP32 input)
    const double *points64;           // or FP64 input (flattened curves); used when set
    const double *bounds64;           // optional: per path x0, y0, x1, y1 (exact curve bounds) instead of the vertices'
    const uint32_t *starts;           // path_count + 1 vertex offsets
    uint32_t path_count, draw_count, batch_count, layer_count;
    uint32_t scene_count, row_count, column_count, pad;
    const tk_fe_scene *scenes;
    const uint32_t *scene_of_path;    // null: every path in scene 0
    const uint32_t *scene_of_draw;
    const uint32_t *scene_of_batch;
    const uint32_t *scene_of_row;
    const uint32_t *covers;           // 4 words per draw: the tiles early occlusion hides
    uint32_t *path_vals;              // TK_FE_PATH_COLUMNS x path_count
    int32_t *path_pb;                 // 4 words per path: pixel bounds
    uint32_t *batch_clip;             // 4 words per batch: its clip stack's tile box
    unsigned long long *batch_area;   // per batch: its draws' tile boxes within its clip box
    int32_t *diff;                    // per scene row y, tiles_width + 1 entries at scene column_start +
                                      // row_start + y * (tiles_width + 1): +1 / -1 at the ends of visible runs
    uint32_t *draw_box;               // 4 words per draw: its tile box (zeros: no tiles or not batched)
    uint32_t *tile_counts;            // per column: list length, then (scanned) list start
    uint32_t *row_counts;             // per row: entries, then (scanned) first entry
    uint4 *row_entries;               // per row, its draws: {draw, x0, x1, hidden x0 | hidden x1 << 16}
    unsigned long long *totals;       // TK_FE_TOTALS
    uint32_t *status;                 // [0] bit 0: a non-finite vertex
    // frame buffers the front end rewrites
    tk_line *lines;
    uint32_t *line_vertex;            // optional: per line, the global index of its first vertex
    tk_path_record *paths;
    tk_draw_record *draws;
    const uint32_t *batch_ids;
    const tk_layer_record *layers;
    const tk_batch_record *batches;
    tk_scan_chunk *scan_chunks;
    tk_scan_chunk_range *chunk_ranges;
    uint32_t *cs_offsets, *cs_lens, *cs_row_starts, *cs_row_ends;
    tk_tile_draw_record *tile_draws;
    uint32_t *tile_draw_indices;
    // capacities of the rewritten buffers (entries); writes past them are dropped
    uint64_t line_capacity, chunk_capacity, cs_chunk_capacity, cs_row_capacity, index_capacity, row_capacity;
};

// Exclusive scan of `columns` columns of `count` u32 each (column-major), in
// place, in blocks of 1024; block sums and column totals are u64.
struct tk_args_fe_scan {
    uint32_t *values;
    uint32_t count, columns;
    unsigned long long *block_sums;   // columns x blocks
    uint32_t blocks;
    unsigned long long *totals;       // columns entries
    uint32_t *overflow;               // set when a prefix exceeds u32
};

// GPU flattening of closed cubic contours, one per path: path p's control
// points are ctrl[starts[p] .. starts[p+1]) = the start point, then three per
// cubic (the two control points and the end point).
struct tk_args_flatten {
    const double *ctrl;               // x, y pairs (FP64), physical pixels
    const uint32_t *starts;           // path_count + 1 offsets
    uint32_t path_count;
    double tolerance;
    uint32_t *counts;                 // per path: flattened points, then (scanned) their first index
    double *bounds;                   // per path: exact curve bounds x0, y0, x1, y1
    double *points;                   // the flattened points (x, y pairs)
    uint64_t point_capacity;
    uint32_t *status;                 // bit 0: a non-finite point; bit 1: a malformed contour
};
