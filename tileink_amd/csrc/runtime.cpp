// Device runtime of libtileink_hip: context (explicit gfx950 device, embedded
// code object loaded with hipModuleLoadData), whole-frame pipeline, and the
// stage entry points of the C ABI.
//
// The frame pipeline replaces the reference's per-frame dispatch list
// (src/native/runtime/program/scene_scan.rs:73-325, program/cumsum.rs:171-269,
// program/coarse.rs:29-129, program/fine.rs:26-67, one coarse+fine pair per
// DrawBatch) with a fixed sequence of gfx950 kernels on one stream:
//   clear -> scan_count -> scan_prefix -> scan_emit -> cumsum
//   -> coarse_count -> coarse_scan -> coarse_emit -> fine
// where coarse and fine process every DrawBatch in a single launch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <cmath>
#include <initializer_list>
#include <cstring>
#include <new>
#include <vector>

#include "tk_internal.h"
#include "tk_kernel_args.h"

extern "C" const unsigned char tk_code_object_begin[];
extern "C" const unsigned char tk_code_object_end[];

namespace tk {

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

namespace {

#define TK_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                      __LINE__);                                                       \
            return TK_ERR_HIP;                                                         \
        }                                                                              \
    } while (0)

#define TK_TRY(expr)                  \
    do {                              \
        tk_status s_ = (expr);        \
        if (s_ != TK_OK) return s_;   \
    } while (0)

}  // namespace
}  // namespace tk

using namespace tk;

struct tk_context {
    int device = -1;
    hipModule_t module = nullptr;
    hipStream_t stream = nullptr;
    // The backdrop prefix and coarse run on `side`, overlapping scan_emit on
    // `stream` (they need the segment ranges and backdrops, not the segments);
    // fork/join events order the two.  Both streams have the default priority.
    hipStream_t side = nullptr;
    hipStream_t cur = nullptr;  // where launches go; null: `stream`
    hipEvent_t fork_a = nullptr, fork_b = nullptr, join = nullptr;
    hipFunction_t f_fill = nullptr, f_copy = nullptr, f_scan_clear = nullptr, f_scan_count = nullptr, f_scan_prefix = nullptr, f_scan_chunk_apply = nullptr,
                  f_scan_emit = nullptr, f_cumsum = nullptr, f_coarse_count = nullptr,
                  f_coarse_scan = nullptr, f_coarse_emit = nullptr, f_layer_classes = nullptr, f_active_paths = nullptr,
                  f_draw_brush = nullptr,
                  f_fine = nullptr, f_coarse_count_scenes = nullptr, f_coarse_emit_scenes = nullptr,
                  f_fine_scenes = nullptr, f_fine_diff = nullptr, f_fine_diff_scenes = nullptr,
                  f_fine_backward = nullptr, f_fine_backward_scenes = nullptr, f_line_backward = nullptr,
                  f_draw_tiles = nullptr, f_fine_diff_clips = nullptr, f_fine_diff_scenes_clips = nullptr,
                  f_fine_backward_clips = nullptr, f_fine_backward_scenes_clips = nullptr,
                  // the same kernels keeping the forward's coverages for the backward
                  f_fine_diff_keep[4] = {}, f_fine_backward_keep[4] = {},
                  f_fill_coverage = nullptr, f_pixel_math = nullptr, f_brush_words = nullptr;
    // Persistent forms (work queues) and the coarse scan's other block sizes.
    hipFunction_t f_fine_persistent = nullptr, f_fine_scenes_persistent = nullptr,
                  f_coarse_count_persistent = nullptr, f_coarse_count_scenes_persistent = nullptr,
                  f_coarse_emit_persistent = nullptr, f_coarse_emit_scenes_persistent = nullptr,
                  f_coarse_scan_1k = nullptr, f_coarse_scan_2k = nullptr, f_coarse_scan_8k = nullptr;
    // The GPU front end of polygon scenes.
    hipFunction_t f_fe_paths = nullptr, f_fe_draw_bounds = nullptr, f_fe_batch_clips = nullptr,
                  f_fe_draw_count = nullptr, f_fe_batch_caps = nullptr, f_fe_tile_counts = nullptr,
                  f_fe_scan_blocks = nullptr, f_fe_scan_top = nullptr, f_fe_scan_apply = nullptr,
                  f_fe_path_fill = nullptr, f_fe_tile_records = nullptr, f_fe_rows = nullptr, f_fe_tiles = nullptr,
                  f_flatten_count = nullptr, f_flatten_emit = nullptr;
    uint32_t compute_units = 0;
    // Scratch for the stage entry points (frames own their own).
    uint32_t *tickets = nullptr;  // [0] scan prefix, [1] coarse scan
    unsigned long long *status = nullptr;
    size_t status_words = 0;
    uint32_t *draw_brush = nullptr;  // fine's per-draw brush words (tk_fine)
    size_t draw_brush_words = 0;
    // This device's copy of the fixed scratch array of tk_coarse_no_alloc (not
    // owned: the HIP runtime holds it until the process ends).
    uint8_t *no_alloc_scratch = nullptr;
    uint32_t epoch = 0;
    uint32_t next_epoch() {
        epoch = (epoch + 1u) & 0x3FFFFFFFu;
        if (epoch == 0u) epoch = 1u;
        return epoch;
    }
};

namespace tk {
namespace {

uint32_t div_up(uint64_t n, uint32_t d) { return uint32_t((n + d - 1) / d); }
// Work counters of a frame's persistent launches: eight per-XCD counters each for
// coarse count, coarse emit and fine.
constexpr uint64_t kQueueWords = 24;
// Grid of a kernel that renumbers its workgroups per XCD: a multiple of 8 * chunk.
uint32_t xcd_grid(uint64_t blocks, uint32_t chunk) {
    const uint64_t unit = chunk ? 8ull * chunk : 1ull;
    return uint32_t((blocks + unit - 1) / unit * unit);
}

// A 2-D grid of at least `blocks` workgroups with x <= 65535 and as few spare
// workgroups as possible: rows of equal width, so at most gy - 1 of them exit at
// once (65 536 tiles: 32 768 x 2, not 65 535 x 2).  Kernels number the
// workgroups linearly (x fastest) and drop those past `blocks`.
void tile_grid(uint32_t blocks, uint32_t &gx, uint32_t &gy) {
    gy = div_up(blocks, 65535u);
    gx = gy ? div_up(blocks, gy) : 0u;
}

hipStream_t launch_stream(const tk_context *ctx) { return ctx->cur ? ctx->cur : ctx->stream; }

// Sends the launches made while it lives to `stream`.
struct StreamScope {
    tk_context *ctx;
    StreamScope(tk_context *c, hipStream_t stream) : ctx(c) { ctx->cur = stream; }
    ~StreamScope() { ctx->cur = nullptr; }
};

template <class Args>
tk_status launch(tk_context *ctx, hipFunction_t f, uint32_t gx, uint32_t gy, uint32_t block, const Args &args,
                 uint32_t dynamic_lds = 0) {
    if (gx == 0 || gy == 0) return TK_OK;
    Args copy = args;
    void *params[] = {&copy};
    hipError_t e = hipModuleLaunchKernel(f, gx, gy, 1, block, 1, 1, dynamic_lds, launch_stream(ctx), params, nullptr);
    if (e != hipSuccess) {
        set_error("kernel launch failed: %s", hipGetErrorString(e));
        return TK_ERR_HIP;
    }
    return TK_OK;
}

tk_status ensure_status(tk_context *ctx, size_t words) {
    if (words <= ctx->status_words) return TK_OK;
    if (ctx->status) {
        TK_HIP(hipStreamSynchronize(ctx->stream));
        TK_HIP(hipFree(ctx->status));
        ctx->status = nullptr;
        ctx->status_words = 0;
    }
    size_t n = std::max<size_t>(words, 1024);
    TK_HIP(hipMalloc(&ctx->status, n * sizeof(unsigned long long)));
    TK_HIP(hipMemsetAsync(ctx->status, 0, n * sizeof(unsigned long long), ctx->stream));
    ctx->status_words = n;
    return TK_OK;
}

tk_status ensure_draw_brush(tk_context *ctx, size_t words) {
    if (words <= ctx->draw_brush_words) return TK_OK;
    if (ctx->draw_brush) {
        TK_HIP(hipStreamSynchronize(ctx->stream));
        TK_HIP(hipFree(ctx->draw_brush));
        ctx->draw_brush = nullptr;
        ctx->draw_brush_words = 0;
    }
    size_t n = std::max<size_t>(words, 1024);
    TK_HIP(hipMalloc(&ctx->draw_brush, n * sizeof(uint32_t)));
    ctx->draw_brush_words = n;
    return TK_OK;
}

tk_status fill_u32(tk_context *ctx, void *dst, uint64_t count, uint32_t value) {
    if (!dst || count == 0) return TK_OK;
    tk_args_fill a{static_cast<uint32_t *>(dst), count, value};
    uint32_t blocks = uint32_t(std::min<uint64_t>(div_up(count, 256), 65536));
    return launch(ctx, ctx->f_fill, blocks, 1, 256, a);
}

// --- stage launchers shared by the C ABI and the frame pipeline -------------

tk_status run_scan_count(tk_context *ctx, const tk_scan_config &cfg, const tk_line *lines, const tk_path_record *paths,
                         uint32_t path_record_count, int32_t *backdrops, uint32_t *counts,
                         const uint32_t *path_active = nullptr, uint32_t xcd_run = TK_XCD_CHUNK_SCAN,
                         uint32_t block = 256u) {
    tk_args_scan_lines a{cfg, lines, paths, path_record_count, backdrops, counts, nullptr, path_active, xcd_run, 0u, 0u};
    return launch(ctx, ctx->f_scan_count, xcd_grid(div_up(cfg.line_count, block), xcd_run), 1, block, a);
}

tk_status run_scan_emit(tk_context *ctx, const tk_scan_config &cfg, const tk_line *lines, const tk_path_record *paths,
                        uint32_t path_record_count, uint32_t *cursors, tk_line_segment *segments,
                        const uint32_t *path_active = nullptr, uint32_t xcd_run = TK_XCD_CHUNK_SCAN,
                        uint32_t block = 256u) {
    // The workgroup's key table (12 B an entry) and staging area (32 B), in dynamic
    // LDS of at most 60 KiB: twice as many entries as lines up to 512 lines,
    // 1024 and 1536 at 1024.
    const uint32_t hash_slots = block <= 512u ? 2u * block : 1024u;
    const uint32_t stage_slots = block <= 512u ? 2u * block : 1536u;
    tk_args_scan_lines a{cfg,     lines,   paths,   path_record_count, nullptr,   cursors,
                         segments, path_active, xcd_run, hash_slots,    stage_slots};
    const uint32_t lds = (3u * hash_slots + 4u + 8u * stage_slots) * 4u;
    return launch(ctx, ctx->f_scan_emit, xcd_grid(div_up(cfg.line_count, block), xcd_run), 1, block, a, lds);
}

tk_status run_scan_prefix(tk_context *ctx, const tk_scan_config &cfg, const tk_scan_chunk *chunks,
                          const tk_scan_chunk_range *chunk_ranges, const tk_path_record *paths, const uint32_t *counts,
                          tk_segment_range *ranges, uint32_t *cursors, uint32_t *bumps, uint32_t *chunk_totals,
                          uint32_t *chunk_offsets) {
    // Pass A: one wave per chunk; waves beyond the chunks zero the bumps of chunk-less paths.
    uint32_t waves = std::max<uint32_t>(cfg.scan_chunk_count, div_up(cfg.path_count, 64));
    if (waves == 0) return TK_OK;
    tk_args_scan_prefix a{cfg, chunks, chunk_ranges, paths, counts, ranges, cursors, bumps, chunk_totals,
                          chunk_offsets, waves};
    TK_TRY(launch(ctx, ctx->f_scan_prefix, div_up(waves, 4), 1, 256, a));
    // Pass B: chunks after the first of their path.
    return launch(ctx, ctx->f_scan_chunk_apply, div_up(cfg.scan_chunk_count, 4), 1, 256, a);
}

tk_status run_cumsum(tk_context *ctx, const tk_cumsum_config &cfg, const uint32_t *offs, const uint32_t *lens,
                     const uint32_t *row_starts, const uint32_t *row_ends, int32_t *backdrops, uint32_t *totals,
                     uint32_t *carries) {
    tk_args_cumsum a{cfg, offs, lens, row_starts, row_ends, backdrops, totals, carries};
    return launch(ctx, ctx->f_cumsum, div_up(cfg.row_count, 256), 1, 256, a);
}

// Scan blocks over the batch x column count table (columns: the listed tiles of an
// incremental render, else every tile).
// (items: counts per scan workgroup, 1024-8192.)
uint32_t coarse_scan_blocks_for(uint64_t entries, uint32_t items = TK_COARSE_SCAN_ITEMS) {
    return uint32_t(std::max<uint64_t>((entries + items - 1) / items, 1));
}
uint32_t coarse_scan_blocks(const tk_coarse_config &cfg, uint32_t columns, uint32_t items = TK_COARSE_SCAN_ITEMS) {
    return coarse_scan_blocks_for(uint64_t(cfg.batch_count) * columns, items);
}
uint32_t coarse_scan_blocks(const tk_coarse_config &cfg) { return coarse_scan_blocks(cfg, cfg.tile_count); }

// Workgroups of a persistent launch: as many as are resident at once (the
// occupancy the kernel's resources allow, on every CU), a multiple of the eight
// XCDs so that each XCD's queue gets the same number of workers.
tk_status resident_grid(tk_context *ctx, hipFunction_t fn, uint32_t block, uint32_t dynamic_lds, uint32_t items,
                        uint32_t *grid) {
    int per_cu = 0;
    TK_HIP(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, int(block), size_t(dynamic_lds)));
    uint64_t n = uint64_t(std::max(per_cu, 1)) * ctx->compute_units;
    n = std::max<uint64_t>(n / 8u * 8u, 8u);
    *grid = uint32_t(std::min<uint64_t>(n, std::max<uint64_t>(uint64_t(items + 7u) / 8u * 8u, 8u)));
    return TK_OK;
}

// Launch choices of the coarse and fine stages (tk_tuning); the stage entry
// points use the defaults.
struct TileLaunch {
    uint32_t xcd_run;
    uint32_t coarse_block = 4u;                      // tiles (waves) per coarse workgroup
    uint32_t scan_items = TK_COARSE_SCAN_ITEMS;      // counts per coarse scan workgroup
    uint32_t *queue = nullptr;                       // persistent launches' counters; null: grid launches
};

// Where the frame keeps its per-tile layer classes: pair_layer[pair] is a layer
// record of each distinct (draw, tag) pair, pair_of[layer] its pair, and bits
// holds tile_count * pair_count bytes.  bits == nullptr: no table.
struct LayerClassPlan {
    const uint32_t *pair_layer = nullptr, *pair_of = nullptr;
    uint32_t pair_count = 0;
    uint8_t *bits = nullptr;
};

tk_status run_coarse(tk_context *ctx, const tk_coarse_config &cfg, const tk_coarse_inputs &in,
                     const tk_coarse_outputs &out, unsigned long long *status, uint32_t *ticket, uint32_t epoch,
                     hipEvent_t *marks, const LayerClassPlan &plan = LayerClassPlan{},
                     tk_tile_list list = tk_tile_list{nullptr, 0u},
                     tk_scene_slots scenes = tk_scene_slots{nullptr, nullptr, 0u, 0ull},
                     TileLaunch shape = TileLaunch{TK_XCD_CHUNK_COARSE}) {
    if (!list.tiles) list.columns = cfg.tile_count;
    if (cfg.batch_count == 0 || list.columns == 0) return TK_OK;
    const uint32_t blocks = scenes.slots ? coarse_scan_blocks_for(scenes.table_entries, shape.scan_items)
                                         : coarse_scan_blocks(cfg, list.columns, shape.scan_items);
    tk_args_coarse a{cfg, in, out, status, ticket, epoch, blocks, tk_layer_classes{nullptr, nullptr, 0u}, list, scenes,
                     shape.xcd_run, shape.queue};
    if (plan.bits && plan.pair_count) {
        tk_args_layer_classes b{cfg, in, plan.pair_layer, plan.pair_count, plan.bits, list};
        // pair-major: div_up(columns, 256) workgroups per pair (the table is at most
        // 64 MiB, so the grid fits 32 bits)
        TK_TRY(launch(ctx, ctx->f_layer_classes, uint32_t(div_up(list.columns, 256) * uint64_t(plan.pair_count)), 1,
                      256, b));
        a.classes = tk_layer_classes{plan.bits, plan.pair_of, plan.pair_count};
    }
    const uint32_t block = 64u * shape.coarse_block;
    hipFunction_t count_fn, emit_fn;
    uint32_t count_grid, emit_grid;
    if (shape.queue) {
        count_fn = scenes.slots ? ctx->f_coarse_count_scenes_persistent : ctx->f_coarse_count_persistent;
        emit_fn = scenes.slots ? ctx->f_coarse_emit_scenes_persistent : ctx->f_coarse_emit_persistent;
        const uint32_t groups = div_up(list.columns, shape.coarse_block);
        TK_TRY(resident_grid(ctx, count_fn, block, 0u, groups, &count_grid));
        TK_TRY(resident_grid(ctx, emit_fn, block, 0u, groups, &emit_grid));
    } else {
        count_fn = scenes.slots ? ctx->f_coarse_count_scenes : ctx->f_coarse_count;
        emit_fn = scenes.slots ? ctx->f_coarse_emit_scenes : ctx->f_coarse_emit;
        count_grid = emit_grid = xcd_grid(div_up(list.columns, shape.coarse_block), shape.xcd_run);
    }
    hipFunction_t scan_fn = shape.scan_items == 1024u   ? ctx->f_coarse_scan_1k
                            : shape.scan_items == 2048u ? ctx->f_coarse_scan_2k
                            : shape.scan_items == 8192u ? ctx->f_coarse_scan_8k
                                                        : ctx->f_coarse_scan;
    TK_TRY(launch(ctx, count_fn, count_grid, 1, block, a));
    if (marks) TK_HIP(hipEventRecord(marks[0], launch_stream(ctx)));
    TK_TRY(launch(ctx, scan_fn, a.scan_blocks, 1, 256, a));
    if (marks) TK_HIP(hipEventRecord(marks[1], launch_stream(ctx)));
    TK_TRY(launch(ctx, emit_fn, emit_grid, 1, block, a));
    return TK_OK;
}

// `draw_brush` holds 2 * cfg.draw_count words; it is rebuilt here from the draws
// and paint before fine reads it.
tk_status run_fine(tk_context *ctx, const tk_fine_config &cfg, const tk_fine_inputs &in, uint32_t *rgba,
                   uint32_t *draw_brush, tk_tile_list list = tk_tile_list{nullptr, 0u},
                   tk_scene_slots scenes = tk_scene_slots{nullptr, nullptr, 0u, 0ull},
                   uint32_t xcd_run = TK_XCD_CHUNK_FINE, uint32_t *queue = nullptr) {
    if (!list.tiles) list.columns = cfg.tile_count;
    if (cfg.clip_spill_depth > TK_FINE_MAX_CLIP_SPILL) {
        set_error("clip nesting deeper than %u is not supported", 4u + TK_FINE_MAX_CLIP_SPILL);
        return TK_ERR_INVALID;
    }
    if (cfg.draw_count) {
        tk_args_draw_brush b{in.draws, cfg.draw_count, in.paint, cfg.paint_words, cfg.paint_brush_base, draw_brush};
        TK_TRY(launch(ctx, ctx->f_draw_brush, div_up(cfg.draw_count, 256), 1, 256, b));
    }
    tk_args_fine a{cfg, in, rgba, draw_brush, list, scenes, xcd_run, queue};
    if (list.columns == 0) return TK_OK;
    const uint32_t lds = cfg.clip_spill_depth * 256u;
    if (queue) {
        hipFunction_t fn = scenes.slots ? ctx->f_fine_scenes_persistent : ctx->f_fine_persistent;
        uint32_t grid;
        TK_TRY(resident_grid(ctx, fn, 256u, lds, list.columns, &grid));
        return launch(ctx, fn, grid, 1, 256, a, lds);
    }
    uint32_t gx, gy;
    tile_grid(xcd_grid(list.columns, xcd_run), gx, gy);
    return launch(ctx, scenes.slots ? ctx->f_fine_scenes : ctx->f_fine, gx, gy, 256, a, lds);
}

}  // namespace
}  // namespace tk

// ---------------------------------------------------------------------------
// Frame
// ---------------------------------------------------------------------------

struct tk_frame {
    tk_context *ctx = nullptr;
    tk_scene meta{};  // sizes and scalars only; pointers are cleared
    uint64_t device_bytes = 0;
    std::vector<void *> allocations;

    // inputs
    tk_line *lines = nullptr;
    tk_path_record *paths = nullptr;
    tk_draw_record *draws = nullptr;
    uint32_t *paint = nullptr, *batch_ids = nullptr;
    uint32_t *atlas = nullptr;  // the scene's image atlas (atlas_pages pages), or null
    tk_layer_record *layers = nullptr;
    tk_batch_record *batches = nullptr;
    tk_tile_draw_record *tile_draws = nullptr;
    uint32_t *tile_draw_indices = nullptr;
    uint32_t index_capacity = 0;  // tile_draw_indices entries allocated
    tk_scan_chunk *scan_chunks = nullptr;
    tk_scan_chunk_range *scan_chunk_ranges = nullptr;
    uint32_t *cs_offsets = nullptr, *cs_lens = nullptr, *cs_row_starts = nullptr, *cs_row_ends = nullptr;
    // scan outputs
    int32_t *backdrops = nullptr, *raw_backdrops = nullptr;
    uint32_t *seg_counts = nullptr, *cursors = nullptr, *bumps = nullptr, *chunk_totals = nullptr,
             *chunk_offsets = nullptr, *cs_totals = nullptr, *cs_carries = nullptr;
    tk_segment_range *seg_ranges = nullptr;
    tk_line_segment *segments = nullptr;
    // coarse outputs
    uint32_t *tile_counts = nullptr, *tile_kinds = nullptr, *batch_bases = nullptr, *status = nullptr;
    tk_segment_range *tile_ranges = nullptr;
    tk_ptcl_record *ptcl = nullptr;
    uint64_t ptcl_capacity = 0;
    uint32_t *draw_brush = nullptr;  // fine's per-draw brush words
    // per-tile layer classes (coarse): see LayerClassPlan
    uint32_t *layer_pair_of = nullptr, *pair_layer = nullptr;
    uint8_t *layer_bits = nullptr;
    uint32_t pair_capacity = 0, pair_count = 0;
    std::vector<uint32_t> host_pair_of, host_pair_layer;
    // incremental renders: the listed tiles, the damage summed-area table, per-path flags
    uint32_t *active_tiles = nullptr, *damage_sat = nullptr, *path_active = nullptr;
    std::vector<uint32_t> host_active, host_sat;
    // output
    uint32_t *image = nullptr;
    // scratch
    uint32_t *tickets = nullptr;
    unsigned long long *coarse_status = nullptr;
    uint32_t epoch = 0;
    bool capture_raw = false;
    bool timing = false;
    // a batch of independent scenes (tk_batch): per-scene slots, column -> scene
    tk_scene_slot *scene_slots = nullptr;
    uint32_t *scene_of_column = nullptr;
    uint32_t scene_count = 0;
    uint64_t table_entries = 0;  // batch: sum of the scenes' table blocks
    // hipGraph replay of a full render (tk_frame_set_graph)
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t ev[13] = {};
    bool have_times = false;
    // the differentiable form (tk_frame_render_diff, tk_frame_backward)
    bool diff_scene = false;    // path draws only, no layers
    bool diff_ready = false;    // a tk_frame_render_diff ran since the last upload
    float *diff_adj = nullptr;  // 256 floats per PTCL record, grown on demand
    uint64_t diff_adj_records = 0;
    uint32_t *diff_rec_of = nullptr, *diff_status = nullptr;
    uint32_t *diff_trans = nullptr;  // 16 bytes per pixel of every column
    uint32_t *diff_draw_tiles = nullptr;  // 16 bytes per draw
    // launch parameters (tk_frame_set_tuning, tk_frame_autotune)
    // The autotuner's choice on the 4096x4096, 200 000-path frame.
    tk_tuning tuning{TK_XCD_CHUNK_SCAN, 128u, TK_XCD_CHUNK_FINE, 0u, 64u, 1u, 8192u, 0u};
    bool diff_kept = false;     // the last differentiable forward kept its coverages
    uint32_t *queues = nullptr;  // kQueueWords work counters of the persistent launches
    // Allocated entries of the buffers whose sizes follow the geometry (the GPU
    // front end grows them): lines, backdrop slabs, segments, scan chunks,
    // cumsum chunks and rows.
    uint64_t cap_lines = 0, cap_backdrops = 0, cap_segments = 0, cap_chunks = 0, cap_cs_chunks = 0, cap_cs_rows = 0;
    // The GPU front end (tk_frame_set_polygons): host-side scene structure kept
    // from creation, and its device scratch (allocated on first use).
    std::vector<uint32_t> fe_cover_host;  // 4 words per draw (draw_covers)
    std::vector<tk_fe_scene> fe_scenes_host;
    std::vector<uint32_t> fe_scene_of_path_host, fe_scene_of_draw_host, fe_scene_of_batch_host, fe_scene_of_row_host;
    uint32_t fe_rows = 0;
    bool fe_ready = false;
    uint32_t *fe_cover = nullptr, *fe_scene_of_path = nullptr, *fe_scene_of_draw = nullptr,
             *fe_scene_of_batch = nullptr, *fe_scene_of_row = nullptr;
    tk_fe_scene *fe_scenes = nullptr;
    uint32_t *fe_path_vals = nullptr, *fe_batch_clip = nullptr, *fe_draw_box = nullptr, *fe_tile_counts = nullptr,
             *fe_row_counts = nullptr, *fe_status = nullptr;
    int32_t *fe_path_pb = nullptr, *fe_diff = nullptr;
    unsigned long long *fe_batch_area = nullptr, *fe_totals = nullptr, *fe_block_sums = nullptr;
    uint64_t fe_block_sum_words = 0;
    uint4 *fe_row_entries = nullptr;
    uint64_t fe_row_capacity = 0;
    unsigned long long *fe_host = nullptr;  // pinned: the totals and the status word
    // GPU flattening of cubic contours (tk_frame_set_cubic_paths): per-path point
    // counts (then starts), exact bounds, and the flattened FP64 points.
    uint32_t *fl_counts = nullptr;
    double *fl_bounds = nullptr, *fl_points = nullptr;
    unsigned long long *fl_totals = nullptr;
    uint64_t fl_point_capacity = 0;
};

namespace tk {
namespace {

template <class T>
tk_status dalloc(tk_frame *f, T **ptr, uint64_t count) {
    uint64_t bytes = std::max<uint64_t>(count * sizeof(T), 16);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%llu) failed: %s", (unsigned long long)bytes, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? TK_ERR_OOM : TK_ERR_HIP;
    }
    f->allocations.push_back(p);
    f->device_bytes += bytes;
    *ptr = static_cast<T *>(p);
    return TK_OK;
}

template <class T>
tk_status upload(tk_frame *f, T *dst, const T *src, uint64_t count) {
    if (count == 0) return TK_OK;
    if (!src) return fail(TK_ERR_INVALID, "scene array missing");
    TK_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, f->ctx->stream));
    return TK_OK;
}

tk_status validate_scene(const tk_scene *s) {
    if (!s) return fail(TK_ERR_INVALID, "null scene");
    if (s->width == 0 || s->height == 0) return fail(TK_ERR_INVALID, "empty canvas");
    if (s->tiles_width != (s->width + 15) / 16 || s->tiles_height != (s->height + 15) / 16)
        return fail(TK_ERR_INVALID, "tile grid does not match canvas size");
    if (uint64_t(s->tiles_width) * s->tiles_height > 0xFFFFFFFFull) return fail(TK_ERR_INVALID, "too many tiles");
    if (s->max_clip_depth > 4u + TK_FINE_MAX_CLIP_SPILL) return fail(TK_ERR_INVALID, "clip nesting too deep");
    if (s->ptcl_capacity >= 0xFFFFFFFFull) return fail(TK_ERR_OVERFLOW, "PTCL capacity exceeds u32 record indices");
    if (s->paint_sdf_shadow_base > s->paint_brush_base || s->paint_brush_base > s->paint_words)
        return fail(TK_ERR_INVALID, "paint section bases out of order");
    if (s->atlas && (s->atlas_width == 0 || s->atlas_height == 0 || s->atlas_pages == 0))
        return fail(TK_ERR_INVALID, "an atlas has at least one texel");
    if (uint64_t(s->atlas_width) * s->atlas_height * s->atlas_pages > (1ull << 31))
        return fail(TK_ERR_OVERFLOW, "atlas larger than 2^31 texels");
    return TK_OK;
}

// The distinct (draw, tag) pairs among the layer records: pair_of[layer] and one
// representative layer per pair.  The classes a coarse kernel derives from a layer
// depend on nothing else of the record.
uint32_t layer_pairs(const tk_scene *s, std::vector<uint32_t> &pair_of, std::vector<uint32_t> &pair_layer) {
    pair_of.assign(s->layer_count, 0u);
    pair_layer.clear();
    std::vector<std::pair<uint64_t, uint32_t>> keyed(s->layer_count);
    for (uint32_t i = 0; i < s->layer_count; ++i)
        keyed[i] = {(uint64_t(s->layers[i].draw) << 32) | s->layers[i].tag, i};
    std::sort(keyed.begin(), keyed.end());
    for (uint32_t k = 0; k < s->layer_count; ++k) {
        if (k == 0 || keyed[k].first != keyed[k - 1].first) pair_layer.push_back(keyed[k].second);
        pair_of[keyed[k].second] = uint32_t(pair_layer.size() - 1);
    }
    return uint32_t(pair_layer.size());
}

// Largest class table a frame allocates (tiles x pairs bytes); beyond it the
// coarse kernels evaluate each layer themselves.
constexpr uint64_t kMaxLayerClassBytes = 64ull << 20;

// The differentiable form takes scenes of path draws (fills and strokes) inside
// any nesting of clip layers, SDF (rounded-rectangle) or path: every PTCL record
// is then a FILL carrying its draw index, a BEGIN_SDF_CLIP, a BEGIN_CLIP or an
// END_CLIP.
bool diff_compatible(const tk_scene *s) {
    for (uint32_t l = 0; l < s->layer_count; ++l) {
        const tk_layer_record &layer = s->layers[l];
        if (layer.tag != 0u || layer.draw >= s->draw_count) return false;
        const tk_draw_record &r = s->draws[layer.draw];
        if (r.tag != TK_DRAW_CLIP) return false;
    }
    for (uint32_t d = 0; d < s->draw_count; ++d) {
        const tk_draw_record &r = s->draws[d];
        if (r.tag == TK_DRAW_CLIP && (r.sdf_offset != TK_INVALID_INDEX || r.path_id < s->path_count)) continue;
        if (r.tag != TK_DRAW_BRUSH || r.path_id >= s->path_count || r.sdf_offset != TK_INVALID_INDEX ||
            r.sdf_shadow_offset != TK_INVALID_INDEX)
            return false;
    }
    return true;
}

// Texels of a scene's image atlas (0 without one).
uint64_t atlas_words(const tk_scene &s) {
    return s.atlas ? uint64_t(s.atlas_width) * s.atlas_height * s.atlas_pages : 0ull;
}

tk_status frame_upload(tk_frame *f, const tk_scene *s) {
    f->diff_scene = diff_compatible(s);
    f->diff_ready = false;
    TK_TRY(upload(f, f->lines, s->lines, s->line_count));
    TK_TRY(upload(f, f->paths, s->paths, s->path_count));
    TK_TRY(upload(f, f->draws, s->draws, s->draw_count));
    TK_TRY(upload(f, f->paint, s->paint, s->paint_words));
    if (f->atlas) TK_TRY(upload(f, f->atlas, s->atlas, atlas_words(*s)));
    TK_TRY(upload(f, f->batch_ids, s->draw_batch_ids, s->draw_count));
    TK_TRY(upload(f, f->layers, s->layers, s->layer_count));
    TK_TRY(upload(f, f->batches, s->batches, s->batch_count));
    TK_TRY(upload(f, f->tile_draws, s->tile_draws, uint64_t(s->tiles_width) * s->tiles_height));
    TK_TRY(upload(f, f->tile_draw_indices, s->tile_draw_indices, s->tile_draw_index_count));
    TK_TRY(upload(f, f->scan_chunks, s->scan_chunks, s->scan_chunk_count));
    TK_TRY(upload(f, f->scan_chunk_ranges, s->scan_chunk_ranges, s->path_count));
    TK_TRY(upload(f, f->cs_offsets, s->cumsum_chunk_offsets, s->cumsum_chunk_count));
    TK_TRY(upload(f, f->cs_lens, s->cumsum_chunk_lens, s->cumsum_chunk_count));
    TK_TRY(upload(f, f->cs_row_starts, s->cumsum_row_starts, s->cumsum_row_count));
    TK_TRY(upload(f, f->cs_row_ends, s->cumsum_row_ends, s->cumsum_row_count));
    f->pair_count = 0;
    if (f->layer_bits && s->layer_count && s->layers) {
        const uint32_t pairs = layer_pairs(s, f->host_pair_of, f->host_pair_layer);
        if (pairs <= f->pair_capacity) {  // else: no table for this scene
            TK_TRY(upload(f, f->layer_pair_of, f->host_pair_of.data(), s->layer_count));
            TK_TRY(upload(f, f->pair_layer, f->host_pair_layer.data(), pairs));
            f->pair_count = pairs;
        }
    }
    TK_HIP(hipStreamSynchronize(f->ctx->stream));
    return TK_OK;
}

// Whether scene `a` fits the buffers of a frame created for `b`.  The tile draw
// index list may shrink or grow up to the frame's allocation (`index_capacity`):
// an opacity change can hide or reveal draws behind an opaque rectangle.
bool same_sizes(const tk_scene &a, const tk_scene &b, uint32_t index_capacity) {
    return a.atlas_width == b.atlas_width && a.atlas_height == b.atlas_height && a.atlas_pages == b.atlas_pages &&
           a.width == b.width && a.height == b.height && a.line_count == b.line_count && a.path_count == b.path_count &&
           a.draw_count == b.draw_count && a.paint_words == b.paint_words && a.layer_count == b.layer_count &&
           a.batch_count == b.batch_count && a.tile_draw_index_count <= index_capacity &&
           a.scan_chunk_count == b.scan_chunk_count && a.cumsum_chunk_count == b.cumsum_chunk_count &&
           a.cumsum_row_count == b.cumsum_row_count && a.backdrop_len == b.backdrop_len &&
           a.segment_capacity == b.segment_capacity && a.ptcl_capacity <= b.ptcl_capacity &&
           a.max_clip_depth <= b.max_clip_depth;
}

tk_scan_config frame_scan_config(const tk_frame *f) {
    const tk_scene &s = f->meta;
    tk_scan_config c{};
    c.clear_len = std::max({s.backdrop_len, s.path_count, s.scan_chunk_count});
    c.backdrop_len = s.backdrop_len;
    c.path_count = s.path_count;
    c.scan_chunk_count = s.scan_chunk_count;
    c.line_count = s.line_count;
    c.segment_capacity = s.segment_capacity;
    return c;
}

tk_coarse_config frame_coarse_config(const tk_frame *f) {
    const tk_scene &s = f->meta;
    tk_coarse_config c{};
    c.tile_count = s.tiles_width * s.tiles_height;
    c.tiles_width = s.tiles_width;
    c.tiles_height = s.tiles_height;
    c.batch_count = s.batch_count;
    c.draw_count = s.draw_count;
    c.path_count = s.path_count;
    c.paint_words = s.paint_words;
    c.paint_brush_base = s.paint_brush_base;
    c.layer_count = s.layer_count;
    c.tile_draw_index_count = s.tile_draw_index_count;
    c.ptcl_capacity = f->ptcl_capacity;
    return c;
}

tk_fine_config frame_fine_config(const tk_frame *f) {
    const tk_scene &s = f->meta;
    tk_fine_config c{};
    c.width = s.width;
    c.height = s.height;
    c.clear_color = s.clear_color;
    c.tile_count = s.tiles_width * s.tiles_height;
    c.tiles_width = s.tiles_width;
    c.tiles_height = s.tiles_height;
    c.batch_count = s.batch_count;
    c.clip_spill_depth = s.max_clip_depth > 4u ? s.max_clip_depth - 4u : 0u;
    c.paint_words = s.paint_words;
    c.paint_sdf_shadow_base = s.paint_sdf_shadow_base;
    c.paint_brush_base = s.paint_brush_base;
    c.draw_count = s.draw_count;
    c.ptcl_capacity = uint32_t(f->ptcl_capacity);
    c.segment_capacity = s.segment_capacity;
    return c;
}

// Entries of the per-(batch, tile) tables: one block per scene for a batch of
// scenes, else batches x columns (the listed tiles of an incremental render).
uint64_t frame_table_entries(const tk_frame *f, uint32_t active_count = 0) {
    if (f->scene_count) return f->table_entries;
    const uint64_t columns = active_count ? active_count : uint64_t(f->meta.tiles_width) * f->meta.tiles_height;
    return uint64_t(f->meta.batch_count) * columns;
}
// Status words of the coarse scan at its smallest block size (1024 counts), the
// most any tuning needs.
uint32_t frame_scan_blocks(const tk_frame *f, uint32_t active_count = 0) {
    return coarse_scan_blocks_for(frame_table_entries(f, active_count), 1024u);
}

// The coarse and fine launch choices of the frame's tuning.
TileLaunch coarse_launch(const tk_frame *f) {
    TileLaunch t{f->tuning.coarse_xcd_run};
    t.coarse_block = f->tuning.coarse_block;
    t.scan_items = f->tuning.coarse_scan_items;
    t.queue = (f->tuning.persistent & TK_PERSISTENT_COARSE) ? f->queues : nullptr;
    return t;
}
uint32_t *fine_queue(const tk_frame *f) { return (f->tuning.persistent & TK_PERSISTENT_FINE) ? f->queues + 16 : nullptr; }

tk_status record(tk_frame *f, int mark) {
    if (f->timing) TK_HIP(hipEventRecord(f->ev[mark], launch_stream(f->ctx)));
    return TK_OK;
}

// A full render, or (active_count != 0) an incremental one over f->active_tiles
// with f->damage_sat uploaded: only paths whose tile box meets a listed tile are
// scanned, and coarse and fine run on the listed tiles, whose tables take one
// column each in list order.  Other tiles of the image keep their pixels.
// Where a differentiable render sends its fine stage.
struct DiffTarget {
    const float *colors;
    float *image;
};

tk_args_fine_diff diff_args(tk_frame *f, const float *colors) {
    tk_args_fine_diff a{};
    a.cfg = frame_fine_config(f);
    a.in.draws = f->draws;
    a.in.paint = f->paint;
    a.in.segments = f->segments;
    a.in.tile_ptcl_ranges = f->tile_ranges;
    a.in.tile_kinds = f->tile_kinds;
    a.in.batch_ptcl_bases = f->batch_bases;
    a.in.ptcl = f->ptcl;
    a.paths = f->paths;
    a.path_count = f->meta.path_count;
    a.rec_of_count = f->meta.backdrop_len;
    a.colors = colors;
    a.adj = f->diff_adj;
    a.adj_records = f->diff_adj_records;
    a.trans = f->diff_trans;
    a.draw_tiles = f->diff_draw_tiles;
    a.rec_of = f->diff_rec_of;
    a.status = f->diff_status;
    a.list = tk_tile_list{nullptr, a.cfg.tile_count};
    a.scenes = tk_scene_slots{f->scene_slots, f->scene_of_column, f->scene_count, f->table_entries};
    a.xcd_run = f->tuning.fine_xcd_run;
    return a;
}

tk_status launch_tiles(tk_context *ctx, hipFunction_t fn, uint32_t columns, const tk_args_fine_diff &a) {
    if (columns == 0) return TK_OK;
    uint32_t gx, gy;
    tile_grid(xcd_grid(columns, a.xcd_run), gx, gy);
    return launch(ctx, fn, gx, gy, 256, a);
}

// The fine stage of a differentiable render: the per-record buffer (one 1 KiB
// slot per PTCL record: the coverage gradients, and the coverages when they are
// kept) is sized to the frame's PTCL capacity, the host-side bound coarse writes
// within (1.2-1.6x the records of typical scenes), so no host round trip waits
// for the record count; then composites.
tk_status run_fine_diff(tk_frame *f, const DiffTarget &t) {
    tk_context *ctx = f->ctx;
    const uint64_t need = f->ptcl_capacity;
    if (need > f->diff_adj_records || !f->diff_adj) {
        if (f->diff_adj) TK_HIP(hipFree(f->diff_adj));
        f->diff_adj = nullptr;
        f->diff_adj_records = 0;
        const uint64_t cap = std::max<uint64_t>(need + need / 4, 64);
        hipError_t e = hipMalloc(&f->diff_adj, cap * 256 * sizeof(float));
        if (e != hipSuccess) {
            f->diff_adj = nullptr;
            set_error("hipMalloc of the coverage buffer (%llu records) failed: %s", (unsigned long long)cap,
                      hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? TK_ERR_OOM : TK_ERR_HIP;
        }
        f->diff_adj_records = cap;
    }
    const tk_fine_config fc = frame_fine_config(f);
    if (!f->diff_trans) TK_TRY(dalloc(f, &f->diff_trans, uint64_t(fc.tile_count) * 256 * 4));
    if (!f->diff_rec_of) TK_TRY(dalloc(f, &f->diff_rec_of, std::max<uint64_t>(f->meta.backdrop_len, 1)));
    TK_TRY(fill_u32(ctx, f->diff_rec_of, f->meta.backdrop_len, TK_INVALID_INDEX));
    if (!f->diff_draw_tiles) TK_TRY(dalloc(f, &f->diff_draw_tiles, std::max<uint64_t>(f->meta.draw_count, 1) * 4));
    if (f->meta.draw_count) {
        tk_args_draw_tiles dt{f->draws, f->meta.draw_count, f->meta.path_count, f->paths, f->diff_draw_tiles};
        TK_TRY(launch(ctx, ctx->f_draw_tiles, div_up(f->meta.draw_count, 256), 1, 256, dt));
    }
    tk_args_fine_diff a = diff_args(f, t.colors);
    a.image = t.image;
    const bool clips = f->meta.layer_count != 0;
    hipFunction_t fn = f->scene_count ? (clips ? ctx->f_fine_diff_scenes_clips : ctx->f_fine_diff_scenes)
                                      : (clips ? ctx->f_fine_diff_clips : ctx->f_fine_diff);
    f->diff_kept = f->tuning.diff_keep_coverage != 0u;
    if (f->diff_kept) fn = ctx->f_fine_diff_keep[(f->scene_count ? 1 : 0) + (clips ? 2 : 0)];
    return launch_tiles(ctx, fn, a.cfg.tile_count, a);
}

tk_status frame_render(tk_frame *f, uint32_t active_count = 0, const DiffTarget *diff = nullptr) {
    tk_context *ctx = f->ctx;
    // A plain or incremental render rewrites the PTCL and tables a backward pass
    // walks; only a differentiable render leaves them matching its own state.
    f->diff_ready = false;
    const tk_scene &s = f->meta;
    const tk_tile_list list{active_count ? f->active_tiles : nullptr, active_count};
    const tk_scene_slots scenes{f->scene_slots, f->scene_of_column, f->scene_count, f->table_entries};
    const uint32_t *path_active = nullptr;
    tk_scan_config scfg = frame_scan_config(f);
    const uint32_t T = s.tiles_width * s.tiles_height;
    f->epoch = (f->epoch + 1u) & 0x3FFFFFFFu;
    if (f->epoch == 0u) f->epoch = 1u;

    TK_TRY(record(f, 0));
    // scan_clear: the deltas and counts accumulate; every other scan output is fully rewritten.
    TK_TRY(fill_u32(ctx, f->backdrops, s.backdrop_len, 0u));
    TK_TRY(fill_u32(ctx, f->seg_counts, s.backdrop_len, 0u));
    TK_TRY(fill_u32(ctx, f->status, 4, 0u));
    // The coarse scan's lookback words start each frame at zero, so a replayed
    // graph (whose epoch argument is fixed) never sees the previous frame's flags.
    TK_TRY(fill_u32(ctx, f->coarse_status, 2ull * frame_scan_blocks(f, active_count), 0u));
    // The persistent launches' work counters (coarse count, coarse emit, fine) start at zero.
    if (f->tuning.persistent) TK_TRY(fill_u32(ctx, f->queues, kQueueWords, 0u));
    if (active_count && s.path_count) {
        tk_args_active_paths ap{f->paths, s.path_count, s.tiles_width, s.tiles_height, f->damage_sat, f->path_active};
        TK_TRY(launch(ctx, ctx->f_active_paths, div_up(s.path_count, 256), 1, 256, ap));
        path_active = f->path_active;
    }
    TK_TRY(record(f, 1));
    TK_TRY(run_scan_count(ctx, scfg, f->lines, f->paths, s.path_count, f->backdrops, f->seg_counts, path_active,
                          f->tuning.scan_xcd_run, f->tuning.scan_block));
    TK_TRY(record(f, 2));
    // Fork A: the backdrop prefix needs only scan_count's deltas.
    TK_HIP(hipEventRecord(ctx->fork_a, ctx->stream));
    TK_HIP(hipStreamWaitEvent(ctx->side, ctx->fork_a, 0));
    {
        StreamScope on_side(ctx, ctx->side);
        TK_TRY(record(f, 5));
        if (f->capture_raw && s.backdrop_len)
            TK_HIP(hipMemcpyAsync(f->raw_backdrops, f->backdrops, uint64_t(s.backdrop_len) * 4, hipMemcpyDeviceToDevice,
                                  ctx->side));
        tk_cumsum_config cc{s.cumsum_row_count, s.cumsum_chunk_count, 0u, 0u};
        TK_TRY(run_cumsum(ctx, cc, f->cs_offsets, f->cs_lens, f->cs_row_starts, f->cs_row_ends, f->backdrops,
                          f->cs_totals, f->cs_carries));
        TK_TRY(record(f, 6));
    }
    TK_TRY(run_scan_prefix(ctx, scfg, f->scan_chunks, f->scan_chunk_ranges, f->paths, f->seg_counts, f->seg_ranges,
                           f->cursors, f->bumps, f->chunk_totals, f->chunk_offsets));
    TK_TRY(record(f, 3));
    // Fork B: coarse needs the segment ranges (and the backdrops, earlier on the
    // side stream), not the segments, so it overlaps scan_emit.
    TK_HIP(hipEventRecord(ctx->fork_b, ctx->stream));
    TK_HIP(hipStreamWaitEvent(ctx->side, ctx->fork_b, 0));
    {
        StreamScope on_side(ctx, ctx->side);
        TK_TRY(record(f, 11));
        tk_coarse_config cc = frame_coarse_config(f);
        // Fields are set by name: the structs are part of the public ABI and may grow.
        tk_coarse_inputs in{};
        in.draws = f->draws;
        in.paths = f->paths;
        in.backdrops = f->backdrops;
        in.segment_ranges = f->seg_ranges;
        in.layers = f->layers;
        in.batches = f->batches;
        in.paint = f->paint;
        in.draw_batch_ids = f->batch_ids;
        in.tile_draws = f->tile_draws;
        in.tile_draw_indices = f->tile_draw_indices;
        tk_coarse_outputs out{};
        out.tile_ptcl_counts = f->tile_counts;
        out.tile_ptcl_ranges = f->tile_ranges;
        out.tile_kinds = f->tile_kinds;
        out.batch_ptcl_bases = f->batch_bases;
        out.ptcl = f->ptcl;
        out.status = f->status;
        LayerClassPlan plan;
        if (f->pair_count) {
            plan.pair_layer = f->pair_layer;
            plan.pair_of = f->layer_pair_of;
            plan.pair_count = f->pair_count;
            plan.bits = f->layer_bits;
        }
        TK_TRY(run_coarse(ctx, cc, in, out, f->coarse_status, &f->tickets[1], f->epoch, f->timing ? &f->ev[7] : nullptr,
                          plan, list, scenes, coarse_launch(f)));
        if (cc.batch_count == 0 && f->timing) {
            TK_TRY(record(f, 7));
            TK_TRY(record(f, 8));
        }
        TK_TRY(record(f, 9));
        TK_HIP(hipEventRecord(ctx->join, ctx->side));
    }
    TK_TRY(run_scan_emit(ctx, scfg, f->lines, f->paths, s.path_count, f->cursors, f->segments, path_active,
                         f->tuning.scan_xcd_run, f->tuning.scan_block));
    TK_TRY(record(f, 4));
    // Join: fine reads the segments and the PTCL.
    TK_HIP(hipStreamWaitEvent(ctx->stream, ctx->join, 0));
    TK_TRY(record(f, 12));
    if (diff) {
        TK_TRY(fill_u32(ctx, f->diff_status, 1, 0u));
        TK_TRY(run_fine_diff(f, *diff));
    } else {
        tk_fine_config fc = frame_fine_config(f);
        tk_fine_inputs in{};
        in.draws = f->draws;
        in.paint = f->paint;
        in.segments = f->segments;
        in.tile_ptcl_ranges = f->tile_ranges;
        in.tile_kinds = f->tile_kinds;
        in.batch_ptcl_bases = f->batch_bases;
        in.ptcl = f->ptcl;
        in.status = f->status + 2;
        in.atlas = f->atlas;
        in.atlas_width = s.atlas_width;
        in.atlas_height = s.atlas_height;
        in.atlas_pages = f->atlas ? s.atlas_pages : 0u;
        (void)T;
        TK_TRY(run_fine(ctx, fc, in, f->image, f->draw_brush, list, scenes, f->tuning.fine_xcd_run, fine_queue(f)));
    }
    TK_TRY(record(f, 10));
    f->have_times = f->timing;
    return TK_OK;
}

void drop_graph(tk_frame *f) {
    if (f->graph_exec) (void)hipGraphExecDestroy(f->graph_exec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    f->graph_exec = nullptr;
    f->graph = nullptr;
}

// A full render through the frame's hipGraph: captured from the two streams on
// first use (fill kernels only, no memset nodes), then replayed as one launch.
tk_status frame_render_graph(tk_frame *f) {
    tk_context *ctx = f->ctx;
    if (!f->graph_exec) {
        TK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const tk_status st = frame_render(f);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (st != TK_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess) {
            set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
            return TK_ERR_HIP;
        }
        f->graph = graph;
        if (hipGraphInstantiate(&f->graph_exec, f->graph, nullptr, nullptr, 0) != hipSuccess) {
            drop_graph(f);
            return fail(TK_ERR_HIP, "hipGraphInstantiate failed");
        }
    }
    TK_HIP(hipGraphLaunch(f->graph_exec, ctx->stream));
    return TK_OK;
}

// How a batch of independent scenes sits in one frame (tk_batch_create).
struct BatchLayout {
    std::vector<tk_scene_slot> slots;
    std::vector<uint32_t> scene_of_column;
    uint64_t table_entries = 0;
    // for the GPU front end: each scene's draw covers and rows, and what belongs to which scene
    std::vector<uint32_t> covers;
    std::vector<tk_fe_scene> fe_scenes;
    std::vector<uint32_t> scene_of_path, scene_of_draw, scene_of_batch, scene_of_row;
};

tk_status frame_create(tk_context *ctx, const tk_scene *s, const BatchLayout *batch, tk_frame **out) {
    if (!ctx || !out) return fail(TK_ERR_INVALID, "null argument");
    *out = nullptr;
    TK_TRY(validate_scene(s));
    tk_frame *f = new (std::nothrow) tk_frame();
    if (!f) return fail(TK_ERR_OOM, "frame allocation failed");
    f->ctx = ctx;
    f->meta = *s;
    f->ptcl_capacity = std::max<uint64_t>(s->ptcl_capacity, 1);
    if (batch) {
        f->scene_count = uint32_t(batch->slots.size());
        f->table_entries = batch->table_entries;
    }
    auto bail = [&](tk_status st) {
        tk_frame_destroy(f);
        return st;
    };
    const uint64_t T = uint64_t(s->tiles_width) * s->tiles_height;
    const uint64_t QT = frame_table_entries(f);
    tk_status st = TK_OK;
#define ALLOC(ptr, n) \
    if ((st = dalloc(f, &f->ptr, (n))) != TK_OK) return bail(st)
    ALLOC(lines, s->line_count);
    ALLOC(paths, s->path_count);
    ALLOC(draws, s->draw_count);
    ALLOC(paint, s->paint_words);
    if (atlas_words(*s)) ALLOC(atlas, atlas_words(*s));
    ALLOC(batch_ids, s->draw_count);
    ALLOC(layers, s->layer_count);
    ALLOC(batches, s->batch_count);
    ALLOC(tile_draws, T);
    ALLOC(tile_draw_indices, s->tile_draw_index_count);
    f->index_capacity = s->tile_draw_index_count;
    ALLOC(scan_chunks, s->scan_chunk_count);
    ALLOC(scan_chunk_ranges, s->path_count);
    ALLOC(cs_offsets, s->cumsum_chunk_count);
    ALLOC(cs_lens, s->cumsum_chunk_count);
    ALLOC(cs_row_starts, s->cumsum_row_count);
    ALLOC(cs_row_ends, s->cumsum_row_count);
    ALLOC(backdrops, s->backdrop_len);
    ALLOC(raw_backdrops, s->backdrop_len);
    ALLOC(seg_counts, s->backdrop_len);
    ALLOC(seg_ranges, s->backdrop_len);
    ALLOC(cursors, s->backdrop_len);
    ALLOC(bumps, s->path_count);
    ALLOC(chunk_totals, s->scan_chunk_count);
    ALLOC(chunk_offsets, s->scan_chunk_count);
    ALLOC(cs_totals, s->cumsum_chunk_count);
    ALLOC(cs_carries, s->cumsum_chunk_count);
    ALLOC(segments, s->segment_capacity);
    ALLOC(tile_counts, QT);
    ALLOC(tile_ranges, QT);
    ALLOC(tile_kinds, QT);
    ALLOC(batch_bases, uint64_t(s->batch_count) + 1);
    ALLOC(status, 4);
    ALLOC(diff_status, 1);
    ALLOC(ptcl, f->ptcl_capacity);
    ALLOC(draw_brush, 2 * std::max<uint64_t>(s->draw_count, 1));
    ALLOC(active_tiles, T);
    ALLOC(damage_sat, uint64_t(s->tiles_width + 1) * (s->tiles_height + 1));
    ALLOC(path_active, std::max<uint64_t>(s->path_count, 1));
    if (batch) {
        ALLOC(scene_slots, batch->slots.size());
        ALLOC(scene_of_column, batch->scene_of_column.size());
        if ((st = upload(f, f->scene_slots, batch->slots.data(), batch->slots.size())) != TK_OK ||
            (st = upload(f, f->scene_of_column, batch->scene_of_column.data(), batch->scene_of_column.size())) != TK_OK)
            return bail(st);
    }
    if (s->layer_count && s->layers && !batch) {  // the class table indexes one canvas's tiles
        std::vector<uint32_t> pair_of, pair_layer;
        const uint64_t pairs = layer_pairs(s, pair_of, pair_layer);
        if (pairs * T <= kMaxLayerClassBytes) {
            ALLOC(layer_pair_of, s->layer_count);
            ALLOC(pair_layer, pairs);
            ALLOC(layer_bits, pairs * T);
            f->pair_capacity = uint32_t(pairs);
        }
    }
    ALLOC(image, uint64_t(s->width) * s->height);
    f->cap_lines = s->line_count;
    f->cap_backdrops = s->backdrop_len;
    f->cap_segments = s->segment_capacity;
    f->cap_chunks = s->scan_chunk_count;
    f->cap_cs_chunks = s->cumsum_chunk_count;
    f->cap_cs_rows = s->cumsum_row_count;
    if (batch) {
        f->fe_cover_host = batch->covers;
        f->fe_scenes_host = batch->fe_scenes;
        f->fe_scene_of_path_host = batch->scene_of_path;
        f->fe_scene_of_draw_host = batch->scene_of_draw;
        f->fe_scene_of_batch_host = batch->scene_of_batch;
        f->fe_scene_of_row_host = batch->scene_of_row;
        f->fe_rows = uint32_t(batch->scene_of_row.size());
    } else {
        draw_covers(*s, f->fe_cover_host);
        f->fe_scenes_host.push_back(tk_fe_scene{s->tiles_width, s->tiles_height, 0u, 0u, 0u, s->draw_count});
        f->fe_rows = s->tiles_height;
    }
    ALLOC(tickets, 4);
    ALLOC(coarse_status, frame_scan_blocks(f));
    ALLOC(queues, kQueueWords);
#undef ALLOC
    // batch_ptcl_bases: a scene without batches never runs coarse, so its one entry stays 0.
    if (hipMemsetAsync(f->tickets, 0, 16, ctx->stream) != hipSuccess ||
        hipMemsetAsync(f->batch_bases, 0, (uint64_t(s->batch_count) + 1) * 4, ctx->stream) != hipSuccess ||
        hipMemsetAsync(f->coarse_status, 0, uint64_t(frame_scan_blocks(f)) * 8, ctx->stream) != hipSuccess ||
        hipMemsetAsync(f->cs_carries, 0, std::max<uint64_t>(s->cumsum_chunk_count, 1) * 4, ctx->stream) != hipSuccess)
        return bail(fail(TK_ERR_HIP, "scratch clear failed"));
    for (auto &e : f->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(TK_ERR_HIP, "event creation failed"));
    if ((st = frame_upload(f, s)) != TK_OK) return bail(st);
    // Only sizes survive in the frame; host pointers belong to the caller.
    f->meta.lines = nullptr;
    f->meta.paths = nullptr;
    f->meta.draws = nullptr;
    f->meta.paint = nullptr;
    f->meta.draw_batch_ids = nullptr;
    f->meta.layers = nullptr;
    f->meta.batches = nullptr;
    f->meta.tile_draws = nullptr;
    f->meta.tile_draw_indices = nullptr;
    f->meta.scan_chunks = nullptr;
    f->meta.scan_chunk_ranges = nullptr;
    f->meta.cumsum_chunk_offsets = nullptr;
    f->meta.cumsum_chunk_lens = nullptr;
    f->meta.cumsum_row_starts = nullptr;
    f->meta.cumsum_row_ends = nullptr;
    *out = f;
    return TK_OK;
}

}  // namespace
}  // namespace tk

extern "C" {

const char *tk_last_error(void) { return tk::g_error.c_str(); }
uint32_t tk_version(void) { return (0u << 16) | 1u; }

tk_status tk_context_create(int device, tk_context **out) {
    if (!out) return fail(TK_ERR_INVALID, "null output");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(TK_ERR_UNAVAILABLE, "no HIP device");
    if (device < 0 || device >= count) return fail(TK_ERR_UNAVAILABLE, "device ordinal out of range");
    hipDeviceProp_t prop;
    TK_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s, this library is built for gfx950", device, prop.gcnArchName);
        return TK_ERR_UNAVAILABLE;
    }
    TK_HIP(hipSetDevice(device));
    tk_context *ctx = new (std::nothrow) tk_context();
    if (!ctx) return fail(TK_ERR_OOM, "context allocation failed");
    ctx->device = device;
    ctx->compute_units = uint32_t(std::max(prop.multiProcessorCount, 1));
    auto cleanup = [&](tk_status st) {
        tk_context_destroy(ctx);
        return st;
    };
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(TK_ERR_HIP, "stream creation failed"));
    for (hipEvent_t *e : {&ctx->fork_a, &ctx->fork_b, &ctx->join})
        if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(TK_ERR_HIP, "event creation failed"));
    if (hipModuleLoadData(&ctx->module, tk_code_object_begin) != hipSuccess)
        return cleanup(fail(TK_ERR_HIP, "loading the gfx950 code object failed"));
    struct {
        hipFunction_t *fn;
        const char *name;
    } table[] = {
        {&ctx->f_fill, "tk_k_fill"},
        {&ctx->f_copy, "tk_k_copy"},
        {&ctx->f_scan_clear, "tk_k_scan_clear"},
        {&ctx->f_scan_count, "tk_k_scan_count"},
        {&ctx->f_scan_prefix, "tk_k_scan_prefix"},
        {&ctx->f_scan_chunk_apply, "tk_k_scan_chunk_apply"},
        {&ctx->f_scan_emit, "tk_k_scan_emit"},
        {&ctx->f_cumsum, "tk_k_cumsum"},
        {&ctx->f_coarse_count, "tk_k_coarse_count"},
        {&ctx->f_coarse_scan, "tk_k_coarse_scan"},
        {&ctx->f_coarse_emit, "tk_k_coarse_emit"},
        {&ctx->f_layer_classes, "tk_k_layer_classes"},
        {&ctx->f_active_paths, "tk_k_active_paths"},
        {&ctx->f_draw_brush, "tk_k_draw_brush"},
        {&ctx->f_fine, "tk_k_fine"},
        {&ctx->f_coarse_count_scenes, "tk_k_coarse_count_scenes"},
        {&ctx->f_coarse_emit_scenes, "tk_k_coarse_emit_scenes"},
        {&ctx->f_fine_scenes, "tk_k_fine_scenes"},
        {&ctx->f_fine_diff, "tk_k_fine_diff"},
        {&ctx->f_fine_diff_scenes, "tk_k_fine_diff_scenes"},
        {&ctx->f_fine_backward, "tk_k_fine_backward"},
        {&ctx->f_fine_backward_scenes, "tk_k_fine_backward_scenes"},
        {&ctx->f_line_backward, "tk_k_line_backward"},
        {&ctx->f_draw_tiles, "tk_k_draw_tiles"},
        {&ctx->f_fine_diff_clips, "tk_k_fine_diff_clips"},
        {&ctx->f_fine_diff_scenes_clips, "tk_k_fine_diff_scenes_clips"},
        {&ctx->f_fine_backward_clips, "tk_k_fine_backward_clips"},
        {&ctx->f_fine_backward_scenes_clips, "tk_k_fine_backward_scenes_clips"},
        {&ctx->f_fine_diff_keep[0], "tk_k_fine_diff_keep"},
        {&ctx->f_fine_diff_keep[1], "tk_k_fine_diff_scenes_keep"},
        {&ctx->f_fine_diff_keep[2], "tk_k_fine_diff_clips_keep"},
        {&ctx->f_fine_diff_keep[3], "tk_k_fine_diff_scenes_clips_keep"},
        {&ctx->f_fine_backward_keep[0], "tk_k_fine_backward_keep"},
        {&ctx->f_fine_backward_keep[1], "tk_k_fine_backward_scenes_keep"},
        {&ctx->f_fine_backward_keep[2], "tk_k_fine_backward_clips_keep"},
        {&ctx->f_fine_backward_keep[3], "tk_k_fine_backward_scenes_clips_keep"},
        {&ctx->f_fill_coverage, "tk_k_fill_coverage"},
        {&ctx->f_pixel_math, "tk_k_pixel_math"},
        {&ctx->f_brush_words, "tk_k_brush_words"},
        {&ctx->f_fine_persistent, "tk_k_fine_persistent"},
        {&ctx->f_fine_scenes_persistent, "tk_k_fine_scenes_persistent"},
        {&ctx->f_coarse_count_persistent, "tk_k_coarse_count_persistent"},
        {&ctx->f_coarse_count_scenes_persistent, "tk_k_coarse_count_scenes_persistent"},
        {&ctx->f_coarse_emit_persistent, "tk_k_coarse_emit_persistent"},
        {&ctx->f_coarse_emit_scenes_persistent, "tk_k_coarse_emit_scenes_persistent"},
        {&ctx->f_coarse_scan_1k, "tk_k_coarse_scan_1k"},
        {&ctx->f_coarse_scan_2k, "tk_k_coarse_scan_2k"},
        {&ctx->f_coarse_scan_8k, "tk_k_coarse_scan_8k"},
        {&ctx->f_fe_paths, "tk_k_fe_paths"},
        {&ctx->f_fe_draw_bounds, "tk_k_fe_draw_bounds"},
        {&ctx->f_fe_batch_clips, "tk_k_fe_batch_clips"},
        {&ctx->f_fe_draw_count, "tk_k_fe_draw_count"},
        {&ctx->f_fe_batch_caps, "tk_k_fe_batch_caps"},
        {&ctx->f_fe_tile_counts, "tk_k_fe_tile_counts"},
        {&ctx->f_fe_scan_blocks, "tk_k_fe_scan_blocks"},
        {&ctx->f_fe_scan_top, "tk_k_fe_scan_top"},
        {&ctx->f_fe_scan_apply, "tk_k_fe_scan_apply"},
        {&ctx->f_fe_path_fill, "tk_k_fe_path_fill"},
        {&ctx->f_fe_tile_records, "tk_k_fe_tile_records"},
        {&ctx->f_fe_rows, "tk_k_fe_rows"},
        {&ctx->f_fe_tiles, "tk_k_fe_tiles"},
        {&ctx->f_flatten_count, "tk_k_flatten_count"},
        {&ctx->f_flatten_emit, "tk_k_flatten_emit"},
    };
    for (auto &entry : table) {
        if (hipModuleGetFunction(entry.fn, ctx->module, entry.name) != hipSuccess) {
            set_error("kernel %s missing from the code object", entry.name);
            return cleanup(TK_ERR_HIP);
        }
    }
    if (hipMalloc(&ctx->tickets, 16) != hipSuccess) return cleanup(fail(TK_ERR_OOM, "ticket allocation failed"));
    if (hipMemset(ctx->tickets, 0, 16) != hipSuccess) return cleanup(fail(TK_ERR_HIP, "ticket clear failed"));
    {
        void *scratch = nullptr;
        const tk_status st = scratch_address(device, &scratch);
        if (st != TK_OK) return cleanup(st);
        ctx->no_alloc_scratch = static_cast<uint8_t *>(scratch);
    }
    *out = ctx;
    return TK_OK;
}

void tk_context_destroy(tk_context *ctx) {
    if (!ctx) return;
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->status) (void)hipFree(ctx->status);
    if (ctx->tickets) (void)hipFree(ctx->tickets);
    if (ctx->draw_brush) (void)hipFree(ctx->draw_brush);
    if (ctx->module) (void)hipModuleUnload(ctx->module);
    for (hipEvent_t e : {ctx->fork_a, ctx->fork_b, ctx->join})
        if (e) (void)hipEventDestroy(e);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void *tk_context_stream(tk_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

tk_status tk_context_synchronize(tk_context *ctx) {
    if (!ctx) return fail(TK_ERR_INVALID, "null context");
    TK_HIP(hipStreamSynchronize(ctx->side));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TK_OK;
}

tk_status tk_context_limits(tk_context *ctx, tk_limits *out) {
    if (!ctx || !out) return fail(TK_ERR_INVALID, "null argument");
    hipDeviceProp_t prop;
    TK_HIP(hipGetDeviceProperties(&prop, ctx->device));
    out->image_dimension = 65536u;
    out->max_tiles = 4096u * 4096u;
    out->device_memory_bytes = prop.totalGlobalMem;
    out->compute_units = uint32_t(prop.multiProcessorCount);
    return TK_OK;
}

// --- device buffers ---------------------------------------------------------

tk_status tk_device_alloc(tk_context *ctx, size_t bytes, void **out) {
    if (!ctx || !out) return fail(TK_ERR_INVALID, "null argument");
    *out = nullptr;
    hipError_t e = hipMalloc(out, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? TK_ERR_OOM : TK_ERR_HIP;
    }
    return TK_OK;
}

tk_status tk_device_free(tk_context *ctx, void *ptr) {
    if (!ctx) return fail(TK_ERR_INVALID, "null context");
    if (!ptr) return TK_OK;
    TK_HIP(hipStreamSynchronize(ctx->stream));
    TK_HIP(hipFree(ptr));
    return TK_OK;
}

tk_status tk_copy_to_device(tk_context *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(TK_ERR_INVALID, "null argument");
    if (!bytes) return TK_OK;
    TK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TK_OK;
}

tk_status tk_copy_to_host(tk_context *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(TK_ERR_INVALID, "null argument");
    if (!bytes) return TK_OK;
    TK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TK_OK;
}

tk_status tk_copy_device(tk_context *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(TK_ERR_INVALID, "null argument");
    if ((bytes | (uintptr_t)dst | (uintptr_t)src) & 15u) return fail(TK_ERR_INVALID, "copy must be 16-byte aligned");
    if (!bytes) return TK_OK;
    tk_args_copy a{static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), bytes / 16u};
    const uint32_t blocks = uint32_t(std::min<uint64_t>((a.count + 255) / 256, 8192));
    TK_TRY(launch(ctx, ctx->f_copy, blocks, 1, 256, a));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TK_OK;
}

// --- frames -----------------------------------------------------------------

tk_status tk_frame_create(tk_context *ctx, const tk_scene *s, tk_frame **out) {
    return tk::frame_create(ctx, s, nullptr, out);
}

}  // extern "C"

namespace tk {
namespace {

// Frees a frame buffer (the streams are idle) and allocates `count` entries.
template <class T>
tk_status reallocate(tk_frame *f, T **ptr, uint64_t count) {
    if (*ptr) {
        auto it = std::find(f->allocations.begin(), f->allocations.end(), static_cast<void *>(*ptr));
        if (it != f->allocations.end()) f->allocations.erase(it);
        TK_HIP(hipFree(*ptr));
        *ptr = nullptr;
    }
    return dalloc(f, ptr, count);
}

// Capacity for `need` entries where `have` are allocated: unchanged when they
// fit, else a quarter more than needed.
bool grow(uint64_t &have, uint64_t need) {
    if (need <= have) return false;
    have = need + need / 4 + 16;
    return true;
}

tk_status fe_prepare(tk_frame *f) {
    if (f->fe_ready) return TK_OK;
    const tk_scene &s = f->meta;
    const uint64_t P = s.path_count, D = s.draw_count, Q = s.batch_count, R = f->fe_rows;
    const uint64_t C = uint64_t(s.tiles_width) * s.tiles_height;  // table columns (a batch: all scenes' tiles)
    tk_status st = TK_OK;
#define FE_ALLOC(ptr, n) \
    if ((st = dalloc(f, &f->ptr, (n))) != TK_OK) return st
    FE_ALLOC(fe_cover, std::max<uint64_t>(4 * D, 4));
    FE_ALLOC(fe_scenes, f->fe_scenes_host.size());
    if (!f->fe_scene_of_path_host.empty()) {
        FE_ALLOC(fe_scene_of_path, f->fe_scene_of_path_host.size());
        FE_ALLOC(fe_scene_of_draw, std::max<size_t>(f->fe_scene_of_draw_host.size(), 1));
        FE_ALLOC(fe_scene_of_batch, std::max<size_t>(f->fe_scene_of_batch_host.size(), 1));
        FE_ALLOC(fe_scene_of_row, std::max<size_t>(f->fe_scene_of_row_host.size(), 1));
        TK_TRY(upload(f, f->fe_scene_of_path, f->fe_scene_of_path_host.data(), f->fe_scene_of_path_host.size()));
        TK_TRY(upload(f, f->fe_scene_of_draw, f->fe_scene_of_draw_host.data(), f->fe_scene_of_draw_host.size()));
        TK_TRY(upload(f, f->fe_scene_of_batch, f->fe_scene_of_batch_host.data(), f->fe_scene_of_batch_host.size()));
        TK_TRY(upload(f, f->fe_scene_of_row, f->fe_scene_of_row_host.data(), f->fe_scene_of_row_host.size()));
    }
    TK_TRY(upload(f, f->fe_cover, f->fe_cover_host.data(), f->fe_cover_host.size()));
    TK_TRY(upload(f, f->fe_scenes, f->fe_scenes_host.data(), f->fe_scenes_host.size()));
    FE_ALLOC(fe_path_vals, std::max<uint64_t>(TK_FE_PATH_COLUMNS * P, 1));
    FE_ALLOC(fe_path_pb, std::max<uint64_t>(4 * P, 4));
    FE_ALLOC(fe_batch_clip, std::max<uint64_t>(4 * Q, 4));
    FE_ALLOC(fe_batch_area, std::max<uint64_t>(Q, 1));
    FE_ALLOC(fe_diff, C + R + 1);
    FE_ALLOC(fe_draw_box, std::max<uint64_t>(4 * D, 4));
    FE_ALLOC(fe_tile_counts, std::max<uint64_t>(C, 1));
    FE_ALLOC(fe_row_counts, std::max<uint64_t>(R, 1));
    FE_ALLOC(fe_totals, TK_FE_TOTALS);
    FE_ALLOC(fe_status, 4);
    const uint64_t blocks = std::max<uint64_t>({div_up(P, 1024), div_up(C, 1024), div_up(R, 1024), 1});
    f->fe_block_sum_words = blocks * TK_FE_PATH_COLUMNS;
    FE_ALLOC(fe_block_sums, f->fe_block_sum_words);
#undef FE_ALLOC
    if (hipHostMalloc(reinterpret_cast<void **>(&f->fe_host), (TK_FE_TOTALS + 1) * sizeof(unsigned long long)) !=
        hipSuccess)
        return fail(TK_ERR_OOM, "pinned host allocation failed");
    f->fe_ready = true;
    return TK_OK;
}

tk_status fe_scan(tk_context *ctx, tk_frame *f, uint32_t *values, uint32_t count, uint32_t columns,
                  unsigned long long *totals) {
    if (count == 0) {
        return fill_u32(ctx, totals, 2ull * columns, 0u);
    }
    const uint32_t blocks = div_up(count, 1024);
    tk_args_fe_scan a{values, count, columns, f->fe_block_sums, blocks, totals, f->fe_status + 1};
    TK_TRY(launch(ctx, ctx->f_fe_scan_blocks, blocks, columns, 256, a));
    TK_TRY(launch(ctx, ctx->f_fe_scan_top, columns, 1, 256, a));
    return launch(ctx, ctx->f_fe_scan_apply, blocks, columns, 256, a);
}

// The vertices of every path: FP32 (`points`) or FP64 (`points64`, with the
// paths' exact curve bounds in `bounds64`), path p's at [starts[p], starts[p+1]).
struct FeInput {
    const float *points = nullptr;
    const double *points64 = nullptr, *bounds64 = nullptr;
    const uint32_t *starts = nullptr;
    uint64_t point_count = 0;
    uint32_t *line_vertex = nullptr;
};

tk_status frame_set_geometry(tk_frame *f, const FeInput &in) {
    const float *points = in.points;
    const uint32_t *starts = in.starts;
    const uint64_t point_count = in.point_count;
    uint32_t *line_vertex = in.line_vertex;
    tk_context *ctx = f->ctx;
    TK_TRY(fe_prepare(f));
    tk_scene &s = f->meta;
    const uint32_t P = s.path_count, D = s.draw_count, Q = s.batch_count, R = f->fe_rows;
    const uint32_t C = s.tiles_width * s.tiles_height;  // table columns (a batch: all scenes' tiles)
    TK_HIP(hipStreamSynchronize(ctx->side));
    tk_args_frontend a{};
    a.points = points;
    a.points64 = in.points64;
    a.bounds64 = in.bounds64;
    a.starts = starts;
    a.path_count = P;
    a.draw_count = D;
    a.batch_count = Q;
    a.layer_count = s.layer_count;
    a.scene_count = uint32_t(f->fe_scenes_host.size());
    a.row_count = R;
    a.column_count = C;
    a.scenes = f->fe_scenes;
    a.scene_of_path = f->fe_scene_of_path;
    a.scene_of_draw = f->fe_scene_of_draw;
    a.scene_of_batch = f->fe_scene_of_batch;
    a.scene_of_row = f->fe_scene_of_row;
    a.covers = f->fe_cover;
    a.path_vals = f->fe_path_vals;
    a.path_pb = f->fe_path_pb;
    a.batch_clip = f->fe_batch_clip;
    a.batch_area = f->fe_batch_area;
    a.diff = f->fe_diff;
    a.draw_box = f->fe_draw_box;
    a.tile_counts = f->fe_tile_counts;
    a.row_counts = f->fe_row_counts;
    a.totals = f->fe_totals;
    a.status = f->fe_status;
    a.paths = f->paths;
    a.draws = f->draws;
    a.batch_ids = f->batch_ids;
    a.layers = f->layers;
    a.batches = f->batches;
    // Phase 1: counts, bounds and prefix sums.
    TK_TRY(fill_u32(ctx, f->fe_diff, uint64_t(C) + R + 1, 0u));
    TK_TRY(fill_u32(ctx, f->fe_row_counts, std::max<uint32_t>(R, 1), 0u));
    TK_TRY(fill_u32(ctx, f->fe_batch_area, 2ull * std::max<uint32_t>(Q, 1), 0u));
    TK_TRY(fill_u32(ctx, f->fe_totals, 2ull * TK_FE_TOTALS, 0u));
    TK_TRY(fill_u32(ctx, f->fe_status, 4, 0u));
    if (P) TK_TRY(launch(ctx, ctx->f_fe_paths, div_up(P, 4), 1, 256, a));
    if (D) TK_TRY(launch(ctx, ctx->f_fe_draw_bounds, div_up(D, 256), 1, 256, a));
    if (Q) TK_TRY(launch(ctx, ctx->f_fe_batch_clips, div_up(Q, 256), 1, 256, a));
    if (D) TK_TRY(launch(ctx, ctx->f_fe_draw_count, div_up(D, 256), 1, 256, a));
    if (Q) TK_TRY(launch(ctx, ctx->f_fe_batch_caps, div_up(Q, 256), 1, 256, a));
    if (R) TK_TRY(launch(ctx, ctx->f_fe_tile_counts, div_up(R, 256), 1, 256, a));
    TK_TRY(fe_scan(ctx, f, f->fe_path_vals, P, TK_FE_PATH_COLUMNS, f->fe_totals));
    TK_TRY(fe_scan(ctx, f, f->fe_tile_counts, C, 1, f->fe_totals + TK_FE_TILE_REFS));
    TK_TRY(fe_scan(ctx, f, f->fe_row_counts, R, 1, f->fe_totals + TK_FE_ROW_ENTRIES));
    TK_HIP(hipMemcpyAsync(f->fe_host, f->fe_totals, TK_FE_TOTALS * 8, hipMemcpyDeviceToHost, ctx->stream));
    TK_HIP(hipMemcpyAsync(f->fe_host + TK_FE_TOTALS, f->fe_status, 8, hipMemcpyDeviceToHost, ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    const unsigned long long *t = f->fe_host;
    const uint32_t flags = uint32_t(t[TK_FE_TOTALS]), overflow = uint32_t(t[TK_FE_TOTALS] >> 32);
    if (flags & 1u) return fail(TK_ERR_INVALID, "non-finite vertex: the frame needs valid vertices before it renders");
    if (overflow || t[TK_FE_TILE_REFS] > 0xFFFFFFFFull || t[TK_FE_PTCL] >= 0xFFFFFFFFull ||
        t[TK_FE_ROW_ENTRIES] > 0xFFFFFFFFull)
        return fail(TK_ERR_OVERFLOW, "geometry exceeds 32-bit record indices");
    if (t[TK_FE_LINES] > point_count) return fail(TK_ERR_INVALID, "more lines than vertices");

    // Grow what the new geometry needs (the streams are idle).
    drop_graph(f);
    const uint64_t lines = t[TK_FE_LINES], backdrops = t[TK_FE_BACKDROPS], segments = t[TK_FE_SEGMENTS];
    const uint64_t chunks = t[TK_FE_SCAN_CHUNKS], cs_rows = t[TK_FE_CUMSUM_ROWS], cs_chunks = t[TK_FE_CUMSUM_CHUNKS];
    if (grow(f->cap_lines, lines)) TK_TRY(reallocate(f, &f->lines, f->cap_lines));
    if (grow(f->cap_backdrops, backdrops)) {
        TK_TRY(reallocate(f, &f->backdrops, f->cap_backdrops));
        TK_TRY(reallocate(f, &f->raw_backdrops, f->cap_backdrops));
        TK_TRY(reallocate(f, &f->seg_counts, f->cap_backdrops));
        TK_TRY(reallocate(f, &f->seg_ranges, f->cap_backdrops));
        TK_TRY(reallocate(f, &f->cursors, f->cap_backdrops));
        if (f->diff_rec_of) TK_TRY(reallocate(f, &f->diff_rec_of, f->cap_backdrops));
    }
    if (grow(f->cap_segments, segments)) TK_TRY(reallocate(f, &f->segments, f->cap_segments));
    if (grow(f->cap_chunks, chunks)) {
        TK_TRY(reallocate(f, &f->scan_chunks, f->cap_chunks));
        TK_TRY(reallocate(f, &f->chunk_totals, f->cap_chunks));
        TK_TRY(reallocate(f, &f->chunk_offsets, f->cap_chunks));
    }
    if (grow(f->cap_cs_chunks, cs_chunks)) {
        TK_TRY(reallocate(f, &f->cs_offsets, f->cap_cs_chunks));
        TK_TRY(reallocate(f, &f->cs_lens, f->cap_cs_chunks));
        TK_TRY(reallocate(f, &f->cs_totals, f->cap_cs_chunks));
        TK_TRY(reallocate(f, &f->cs_carries, f->cap_cs_chunks));
        TK_TRY(fill_u32(ctx, f->cs_carries, f->cap_cs_chunks, 0u));
    }
    if (grow(f->cap_cs_rows, cs_rows)) {
        TK_TRY(reallocate(f, &f->cs_row_starts, f->cap_cs_rows));
        TK_TRY(reallocate(f, &f->cs_row_ends, f->cap_cs_rows));
    }
    uint64_t index_cap = f->index_capacity;
    if (grow(index_cap, t[TK_FE_TILE_REFS])) {
        TK_TRY(reallocate(f, &f->tile_draw_indices, index_cap));
        f->index_capacity = uint32_t(std::min<uint64_t>(index_cap, 0xFFFFFFFFull));
    }
    uint64_t ptcl_cap = f->ptcl_capacity;
    if (grow(ptcl_cap, std::max<uint64_t>(t[TK_FE_PTCL], 1))) {
        ptcl_cap = std::min<uint64_t>(ptcl_cap, 0xFFFFFFFEull);
        TK_TRY(reallocate(f, &f->ptcl, ptcl_cap));
        f->ptcl_capacity = ptcl_cap;
    }
    if (grow(f->fe_row_capacity, t[TK_FE_ROW_ENTRIES])) TK_TRY(reallocate(f, &f->fe_row_entries, f->fe_row_capacity));

    // Phase 2: the fills.
    a.lines = f->lines;
    a.line_vertex = line_vertex;
    a.scan_chunks = f->scan_chunks;
    a.chunk_ranges = f->scan_chunk_ranges;
    a.cs_offsets = f->cs_offsets;
    a.cs_lens = f->cs_lens;
    a.cs_row_starts = f->cs_row_starts;
    a.cs_row_ends = f->cs_row_ends;
    a.tile_draws = f->tile_draws;
    a.tile_draw_indices = f->tile_draw_indices;
    a.row_entries = f->fe_row_entries;
    a.line_capacity = f->cap_lines;
    a.chunk_capacity = f->cap_chunks;
    a.cs_chunk_capacity = f->cap_cs_chunks;
    a.cs_row_capacity = f->cap_cs_rows;
    a.index_capacity = f->index_capacity;
    a.row_capacity = f->fe_row_capacity;
    if (P) TK_TRY(launch(ctx, ctx->f_fe_path_fill, div_up(P, 4), 1, 256, a));
    if (C) TK_TRY(launch(ctx, ctx->f_fe_tile_records, div_up(C, 256), 1, 256, a));
    if (R) TK_TRY(launch(ctx, ctx->f_fe_rows, R, 1, 256, a));
    if (R) TK_TRY(launch(ctx, ctx->f_fe_tiles, R, 1, 256, a));
    s.line_count = uint32_t(lines);
    s.backdrop_len = uint32_t(backdrops);
    s.segment_capacity = uint32_t(segments);
    s.scan_chunk_count = uint32_t(chunks);
    s.cumsum_chunk_count = uint32_t(cs_chunks);
    s.cumsum_row_count = uint32_t(cs_rows);
    s.tile_draw_index_count = uint32_t(t[TK_FE_TILE_REFS]);
    s.ptcl_capacity = t[TK_FE_PTCL];
    f->diff_ready = false;
    return TK_OK;
}

// GPU flattening of one closed cubic contour per path, then the front end on
// the flattened FP64 points and the curves' exact bounds.
tk_status frame_set_cubic_paths(tk_frame *f, const double *ctrl, const uint32_t *starts, double tolerance) {
    tk_context *ctx = f->ctx;
    TK_TRY(fe_prepare(f));
    const uint32_t P = f->meta.path_count;
    tk_status st = TK_OK;
    if (!f->fl_counts) {
        if ((st = dalloc(f, &f->fl_counts, uint64_t(P) + 1)) != TK_OK ||
            (st = dalloc(f, &f->fl_bounds, 4 * std::max<uint64_t>(P, 1))) != TK_OK ||
            (st = dalloc(f, &f->fl_totals, 2)) != TK_OK)
            return st;
    }
    TK_HIP(hipStreamSynchronize(ctx->side));
    tk_args_flatten a{ctrl, starts, P, tolerance, f->fl_counts, f->fl_bounds, f->fl_points, f->fl_point_capacity,
                      f->fe_status};
    TK_TRY(fill_u32(ctx, f->fe_status, 4, 0u));
    TK_TRY(fill_u32(ctx, f->fl_totals, 4, 0u));
    if (P) TK_TRY(launch(ctx, ctx->f_flatten_count, div_up(P, 256), 1, 256, a));
    TK_TRY(fe_scan(ctx, f, f->fl_counts, P, 1, f->fl_totals));
    // starts[P] = the total (the low word of the scan's u64 total)
    TK_HIP(hipMemcpyAsync(f->fl_counts + P, f->fl_totals, 4, hipMemcpyDeviceToDevice, ctx->stream));
    TK_HIP(hipMemcpyAsync(f->fe_host, f->fl_totals, 8, hipMemcpyDeviceToHost, ctx->stream));
    TK_HIP(hipMemcpyAsync(f->fe_host + 1, f->fe_status, 8, hipMemcpyDeviceToHost, ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t total = f->fe_host[0];
    const uint32_t flags = uint32_t(f->fe_host[1]), overflow = uint32_t(f->fe_host[1] >> 32);
    if (flags & 1u) return fail(TK_ERR_INVALID, "non-finite control point: the frame needs valid geometry before it renders");
    if (flags & 2u) return fail(TK_ERR_INVALID, "a path's points are not a start point and three per cubic");
    if (overflow || total > 0x7FFFFFFFull) return fail(TK_ERR_OVERFLOW, "too many flattened points");
    uint64_t cap = f->fl_point_capacity;
    if (grow(cap, std::max<uint64_t>(total, 1))) {
        TK_TRY(reallocate(f, &f->fl_points, 2 * cap));
        f->fl_point_capacity = cap;
    }
    a.points = f->fl_points;
    a.point_capacity = f->fl_point_capacity;
    if (P) TK_TRY(launch(ctx, ctx->f_flatten_emit, div_up(P, 256), 1, 256, a));
    FeInput in;
    in.points64 = f->fl_points;
    in.bounds64 = f->fl_bounds;
    in.starts = f->fl_counts;
    in.point_count = total;
    return frame_set_geometry(f, in);
}

}  // namespace
}  // namespace tk

extern "C" {

tk_status tk_frame_set_polygons(tk_frame *f, const float *points, const uint32_t *starts, uint64_t point_count,
                                uint32_t *line_vertex) {
    if (!f || (point_count && !points) || !starts) return fail(TK_ERR_INVALID, "null argument");
    if (point_count > 0x7FFFFFFFull) return fail(TK_ERR_INVALID, "too many vertices");
    tk::FeInput in;
    in.points = points;
    in.starts = starts;
    in.point_count = point_count;
    in.line_vertex = line_vertex;
    return tk::frame_set_geometry(f, in);
}

tk_status tk_frame_set_cubic_paths(tk_frame *f, const double *ctrl, const uint32_t *starts, double tolerance) {
    if (!f || !ctrl || !starts) return fail(TK_ERR_INVALID, "null argument");
    if (!(tolerance > 0.0) || !std::isfinite(tolerance)) return fail(TK_ERR_INVALID, "tolerance must be positive");
    return tk::frame_set_cubic_paths(f, ctrl, starts, tolerance);
}




// --- batches of independent scenes -------------------------------------------

}  // extern "C"

struct tk_batch {
    tk_frame *frame = nullptr;
    std::vector<tk_scene_slot> slots;
};

namespace tk {
namespace {

constexpr uint32_t kNoIndex = 0xFFFFFFFFu;
uint32_t shifted(uint32_t v, uint64_t by) { return v == kNoIndex ? v : uint32_t(v + by); }

// The scenes laid end to end as one scene: every index into another array is
// offset by the lengths of the earlier scenes' arrays, the paint blob keeps its
// three regions (SDF, shadow, brush) each concatenated over the scenes, and the
// tile bins of the scenes' tile grids follow one another (one table column per
// tile).  The merged canvas is a single row of tiles; the slots carry each
// scene's real canvas.
struct MergedScenes {
    std::vector<tk_line> lines;
    std::vector<tk_path_record> paths;
    std::vector<tk_draw_record> draws;
    std::vector<uint32_t> paint, batch_ids, tile_draw_indices, cs_offsets, cs_lens, cs_row_starts, cs_row_ends;
    std::vector<tk_layer_record> layers;
    std::vector<tk_batch_record> batches;
    std::vector<tk_tile_draw_record> tile_draws;
    std::vector<tk_scan_chunk> chunks;
    std::vector<tk_scan_chunk_range> chunk_ranges;
    std::vector<uint32_t> atlas;  // the scenes' atlas pages, each re-laid to the widest and tallest page
    tk_scene scene{};
    BatchLayout layout;
};

template <class T>
void append(std::vector<T> &dst, const T *src, uint64_t n) {
    if (n) dst.insert(dst.end(), src, src + n);
}

tk_status merge_scenes(const tk_scene *scenes, uint32_t count, MergedScenes &m) {
    uint64_t sdf_total = 0, shadow_total = 0, columns = 0;
    for (uint32_t i = 0; i < count; ++i) {
        sdf_total += scenes[i].paint_sdf_shadow_base;
        shadow_total += scenes[i].paint_brush_base - scenes[i].paint_sdf_shadow_base;
    }
    uint64_t line_off = 0, path_off = 0, draw_off = 0, sdf_off = 0, shadow_off = 0, brush_off = 0, batch_off = 0,
             layer_off = 0, tdi_off = 0, chunk_off = 0, cs_chunk_off = 0, backdrop_off = 0, segment_off = 0,
             entries = 0, ptcl = 0;
    uint32_t depth = 0;
    std::vector<uint32_t> sdf, shadow, brush;
    // The merged atlas: every scene's pages in turn, as pages of the widest width
    // and the tallest height; an image brush's page word moves by the pages of
    // the scenes before its own.
    uint32_t atlas_w = 0, atlas_h = 0;
    uint64_t atlas_pages = 0, page_off = 0;
    for (uint32_t i = 0; i < count; ++i)
        if (scenes[i].atlas && scenes[i].atlas_pages) {
            atlas_w = std::max(atlas_w, scenes[i].atlas_width);
            atlas_h = std::max(atlas_h, scenes[i].atlas_height);
            atlas_pages += scenes[i].atlas_pages;
        }
    if (uint64_t(atlas_w) * atlas_h * atlas_pages > (1ull << 31)) return fail(TK_ERR_OVERFLOW, "batch atlas too large");
    m.atlas.assign(size_t(atlas_w) * atlas_h * atlas_pages, 0u);
    std::vector<std::pair<uint64_t, uint32_t>> image_brushes;  // (offset in the brush region, page shift)
    for (uint32_t i = 0; i < count; ++i) {
        const tk_scene &s = scenes[i];
        if (s.paint_sdf_shadow_base > s.paint_brush_base || s.paint_brush_base > s.paint_words)
            return fail(TK_ERR_INVALID, "scene paint regions out of order");
        if (s.atlas && s.atlas_pages) {
            for (uint64_t p = 0; p < s.atlas_pages; ++p)
                for (uint32_t y = 0; y < s.atlas_height; ++y)
                    std::memcpy(&m.atlas[((page_off + p) * atlas_h + y) * atlas_w],
                                &s.atlas[(p * s.atlas_height + y) * s.atlas_width], size_t(s.atlas_width) * 4);
            for (uint32_t k = 0; k < s.draw_count; ++k) {
                const uint32_t b = s.draws[k].brush_offset;
                if (b != kNoIndex && uint64_t(s.paint_brush_base) + b < s.paint_words &&
                    s.paint[s.paint_brush_base + b] == TK_BRUSH_IMAGE)
                    image_brushes.emplace_back(brush_off + b, uint32_t(page_off));
            }
            page_off += s.atlas_pages;
        }
        const uint64_t T = uint64_t(s.tiles_width) * s.tiles_height;
        tk_scene_slot slot{};
        slot.width = s.width;
        slot.height = s.height;
        slot.tiles_width = s.tiles_width;
        slot.tiles_height = s.tiles_height;
        slot.clear_color = s.clear_color;
        slot.batch_start = uint32_t(batch_off);
        slot.batch_end = uint32_t(batch_off + s.batch_count);
        slot.column_start = uint32_t(columns);
        slot.table_base = entries;
        slot.image_offset = columns * 256u;  // the scene's tiles cover its pixels
        m.layout.slots.push_back(slot);
        m.layout.scene_of_column.insert(m.layout.scene_of_column.end(), T, i);
        {
            std::vector<uint32_t> cover;
            draw_covers(s, cover);
            m.layout.covers.insert(m.layout.covers.end(), cover.begin(), cover.end());
            m.layout.fe_scenes.push_back(tk_fe_scene{s.tiles_width, s.tiles_height, uint32_t(columns),
                                                     uint32_t(m.layout.scene_of_row.size()), uint32_t(draw_off),
                                                     uint32_t(draw_off + s.draw_count)});
            m.layout.scene_of_path.insert(m.layout.scene_of_path.end(), s.path_count, i);
            m.layout.scene_of_draw.insert(m.layout.scene_of_draw.end(), s.draw_count, i);
            m.layout.scene_of_batch.insert(m.layout.scene_of_batch.end(), s.batch_count, i);
            m.layout.scene_of_row.insert(m.layout.scene_of_row.end(), s.tiles_height, i);
        }

        for (uint32_t k = 0; k < s.line_count; ++k) {
            tk_line l = s.lines[k];
            l.path_id = uint32_t(l.path_id + path_off);
            m.lines.push_back(l);
        }
        for (uint32_t k = 0; k < s.path_count; ++k) {
            tk_path_record p = s.paths[k];
            p.path_id = uint32_t(p.path_id + path_off);
            p.line_start = uint32_t(p.line_start + line_off);
            p.data_offset = uint32_t(p.data_offset + backdrop_off);
            p.segment_start = uint32_t(p.segment_start + segment_off);
            m.paths.push_back(p);
            tk_scan_chunk_range r = s.scan_chunk_ranges[k];
            r.start = uint32_t(r.start + chunk_off);
            r.end = uint32_t(r.end + chunk_off);
            m.chunk_ranges.push_back(r);
        }
        for (uint32_t k = 0; k < s.draw_count; ++k) {
            tk_draw_record d = s.draws[k];
            d.path_id = shifted(d.path_id, path_off);
            d.sdf_offset = shifted(d.sdf_offset, sdf_off);
            d.sdf_shadow_offset = shifted(d.sdf_shadow_offset, shadow_off);
            d.brush_offset = shifted(d.brush_offset, brush_off);
            m.draws.push_back(d);
            m.batch_ids.push_back(shifted(s.draw_batch_ids[k], batch_off));
        }
        append(sdf, s.paint, s.paint_sdf_shadow_base);
        append(shadow, s.paint + s.paint_sdf_shadow_base, s.paint_brush_base - s.paint_sdf_shadow_base);
        append(brush, s.paint + s.paint_brush_base, s.paint_words - s.paint_brush_base);
        for (uint32_t k = 0; k < s.layer_count; ++k) {
            tk_layer_record l = s.layers[k];
            l.draw = shifted(l.draw, draw_off);
            m.layers.push_back(l);
        }
        for (uint32_t k = 0; k < s.batch_count; ++k)
            m.batches.push_back(tk_batch_record{uint32_t(s.batches[k].layer_start + layer_off),
                                                uint32_t(s.batches[k].layer_end + layer_off)});
        for (uint64_t k = 0; k < T; ++k) {
            tk_tile_draw_record t = s.tile_draws[k];
            t.start = shifted(t.start, tdi_off);
            m.tile_draws.push_back(t);
        }
        for (uint32_t k = 0; k < s.tile_draw_index_count; ++k)
            m.tile_draw_indices.push_back(shifted(s.tile_draw_indices[k], draw_off));
        for (uint32_t k = 0; k < s.scan_chunk_count; ++k) {
            tk_scan_chunk c = s.scan_chunks[k];
            c.path_id = uint32_t(c.path_id + path_off);
            c.backdrop_offset = uint32_t(c.backdrop_offset + backdrop_off);
            c.segment_start = uint32_t(c.segment_start + segment_off);
            m.chunks.push_back(c);
        }
        for (uint32_t k = 0; k < s.cumsum_chunk_count; ++k) {
            m.cs_offsets.push_back(uint32_t(s.cumsum_chunk_offsets[k] + backdrop_off));
            m.cs_lens.push_back(s.cumsum_chunk_lens[k]);
        }
        for (uint32_t k = 0; k < s.cumsum_row_count; ++k) {
            m.cs_row_starts.push_back(uint32_t(s.cumsum_row_starts[k] + cs_chunk_off));
            m.cs_row_ends.push_back(uint32_t(s.cumsum_row_ends[k] + cs_chunk_off));
        }
        line_off += s.line_count;
        path_off += s.path_count;
        draw_off += s.draw_count;
        sdf_off += s.paint_sdf_shadow_base;
        shadow_off += s.paint_brush_base - s.paint_sdf_shadow_base;
        brush_off += s.paint_words - s.paint_brush_base;
        batch_off += s.batch_count;
        layer_off += s.layer_count;
        tdi_off += s.tile_draw_index_count;
        chunk_off += s.scan_chunk_count;
        cs_chunk_off += s.cumsum_chunk_count;
        backdrop_off += s.backdrop_len;
        segment_off += s.segment_capacity;
        columns += T;
        entries += uint64_t(s.batch_count) * T;
        ptcl += s.ptcl_capacity;
        depth = std::max(depth, s.max_clip_depth);
    }
    const uint64_t limit = 0xFFFFFFF0ull;
    if (line_off > limit || path_off > limit || draw_off > limit || sdf_off + shadow_off + brush_off > limit ||
        batch_off > limit || layer_off > limit || tdi_off > limit || chunk_off > limit || cs_chunk_off > limit ||
        backdrop_off > limit || segment_off > limit || columns > (1ull << 27))
        return fail(TK_ERR_OVERFLOW, "batch exceeds 32-bit record indices");
    m.paint = std::move(sdf);
    m.paint.insert(m.paint.end(), shadow.begin(), shadow.end());
    m.paint.insert(m.paint.end(), brush.begin(), brush.end());
    // a header several draws share moves once
    std::sort(image_brushes.begin(), image_brushes.end());
    image_brushes.erase(std::unique(image_brushes.begin(), image_brushes.end()), image_brushes.end());
    for (const auto &ib : image_brushes) {
        const uint64_t word = sdf_total + shadow_total + ib.first + 4u;  // the page word
        if (word < m.paint.size() && !(m.paint[word] & 0x80000000u)) m.paint[word] += ib.second;
    }
    m.layout.table_entries = entries;
    tk_scene &o = m.scene;
    o.tiles_width = uint32_t(columns);
    o.tiles_height = 1;
    o.width = o.tiles_width * 16u;
    o.height = 16;
    o.clear_color = 0;
    o.lines = m.lines.data(), o.line_count = uint32_t(line_off);
    o.paths = m.paths.data(), o.path_count = uint32_t(path_off);
    o.draws = m.draws.data(), o.draw_count = uint32_t(draw_off);
    o.paint = m.paint.data(), o.paint_words = uint32_t(m.paint.size());
    o.paint_sdf_shadow_base = uint32_t(sdf_total);
    o.paint_brush_base = uint32_t(sdf_total + shadow_total);
    o.draw_batch_ids = m.batch_ids.data();
    o.layers = m.layers.data(), o.layer_count = uint32_t(layer_off);
    o.batches = m.batches.data(), o.batch_count = uint32_t(batch_off);
    o.tile_draws = m.tile_draws.data();
    o.tile_draw_indices = m.tile_draw_indices.data(), o.tile_draw_index_count = uint32_t(tdi_off);
    o.scan_chunks = m.chunks.data(), o.scan_chunk_count = uint32_t(chunk_off);
    o.scan_chunk_ranges = m.chunk_ranges.data();
    o.cumsum_chunk_offsets = m.cs_offsets.data();
    o.cumsum_chunk_lens = m.cs_lens.data();
    o.cumsum_chunk_count = uint32_t(cs_chunk_off);
    o.cumsum_row_starts = m.cs_row_starts.data();
    o.cumsum_row_ends = m.cs_row_ends.data();
    o.cumsum_row_count = uint32_t(m.cs_row_starts.size());
    o.backdrop_len = uint32_t(backdrop_off);
    o.segment_capacity = uint32_t(segment_off);
    o.max_clip_depth = depth;
    o.ptcl_capacity = ptcl;
    o.atlas = atlas_pages ? m.atlas.data() : nullptr;
    o.atlas_width = atlas_w;
    o.atlas_height = atlas_h;
    o.atlas_pages = uint32_t(atlas_pages);
    return TK_OK;
}

}  // namespace
}  // namespace tk

extern "C" {

tk_status tk_batch_create(tk_context *ctx, const tk_scene *scenes, uint32_t count, tk_batch **out) {
    if (!ctx || !out || (count && !scenes)) return fail(TK_ERR_INVALID, "null argument");
    *out = nullptr;
    if (count == 0) return fail(TK_ERR_INVALID, "empty batch");
    for (uint32_t i = 0; i < count; ++i) TK_TRY(validate_scene(&scenes[i]));
    MergedScenes m;
    TK_TRY(merge_scenes(scenes, count, m));
    tk_frame *f = nullptr;
    TK_TRY(tk::frame_create(ctx, &m.scene, &m.layout, &f));
    tk_batch *b = new (std::nothrow) tk_batch();
    if (!b) {
        tk_frame_destroy(f);
        return fail(TK_ERR_OOM, "batch allocation failed");
    }
    b->frame = f;
    b->slots = m.layout.slots;
    *out = b;
    return TK_OK;
}

tk_frame *tk_batch_frame(tk_batch *b) { return b ? b->frame : nullptr; }

tk_status tk_batch_render(tk_batch *b) {
    if (!b) return fail(TK_ERR_INVALID, "null batch");
    return tk_frame_render(b->frame);
}

tk_status tk_batch_read_image(tk_batch *b, uint32_t scene, uint32_t *rgba, size_t pixel_count) {
    if (!b || !rgba) return fail(TK_ERR_INVALID, "null argument");
    if (scene >= b->slots.size()) return fail(TK_ERR_INVALID, "scene index out of range");
    const tk_scene_slot &s = b->slots[scene];
    const uint64_t n = uint64_t(s.width) * s.height;
    if (pixel_count < n) return fail(TK_ERR_INVALID, "output smaller than the scene's image");
    TK_HIP(hipMemcpyAsync(rgba, b->frame->image + s.image_offset, n * 4, hipMemcpyDeviceToHost, b->frame->ctx->stream));
    TK_HIP(hipStreamSynchronize(b->frame->ctx->stream));
    return TK_OK;
}

void tk_batch_destroy(tk_batch *b) {
    if (!b) return;
    tk_frame_destroy(b->frame);
    delete b;
}

tk_status tk_frame_update(tk_frame *f, const tk_scene *s) {
    if (!f) return fail(TK_ERR_INVALID, "null frame");
    if (f->scene_count) return fail(TK_ERR_INVALID, "a batch of scenes cannot be updated in place");
    TK_TRY(validate_scene(s));
    if (!same_sizes(*s, f->meta, f->index_capacity)) return fail(TK_ERR_INVALID, "scene sizes differ from the frame's");
    if (f->graph_exec) {  // the captured arguments (layer pair count, clear colour, paint bases) may change
        TK_HIP(hipStreamSynchronize(f->ctx->stream));
        drop_graph(f);
    }
    TK_TRY(frame_upload(f, s));
    // Scalars the kernels take as arguments follow the new scene.
    f->meta.clear_color = s->clear_color;
    f->meta.paint_sdf_shadow_base = s->paint_sdf_shadow_base;
    f->meta.paint_brush_base = s->paint_brush_base;
    f->meta.tile_draw_index_count = s->tile_draw_index_count;
    return TK_OK;
}

void tk_frame_destroy(tk_frame *f) {
    if (!f) return;
    if (f->ctx && f->ctx->side) (void)hipStreamSynchronize(f->ctx->side);
    if (f->ctx && f->ctx->stream) (void)hipStreamSynchronize(f->ctx->stream);
    drop_graph(f);
    for (void *p : f->allocations) (void)hipFree(p);
    if (f->diff_adj) (void)hipFree(f->diff_adj);
    if (f->fe_host) (void)hipHostFree(f->fe_host);
    for (auto &e : f->ev)
        if (e) (void)hipEventDestroy(e);
    delete f;
}

tk_status tk_frame_render(tk_frame *f) {
    if (!f) return fail(TK_ERR_INVALID, "null frame");
    f->diff_ready = false;  // graph replays do not pass through frame_render
    if (f->use_graph && !f->timing && !f->capture_raw) return frame_render_graph(f);
    return frame_render(f);
}

namespace {
bool valid_run(uint32_t v) { return v == 0u || (v <= 1024u && (v & (v - 1u)) == 0u); }
}  // namespace

tk_status tk_frame_get_tuning(tk_frame *f, tk_tuning *out) {
    if (!f || !out) return fail(TK_ERR_INVALID, "null argument");
    *out = f->tuning;
    return TK_OK;
}

tk_status tk_frame_set_tuning(tk_frame *f, const tk_tuning *t) {
    if (!f || !t) return fail(TK_ERR_INVALID, "null argument");
    if (!valid_run(t->scan_xcd_run) || !valid_run(t->coarse_xcd_run) || !valid_run(t->fine_xcd_run))
        return fail(TK_ERR_INVALID, "XCD run lengths are 0 or powers of two up to 1024");
    if (t->persistent > (TK_PERSISTENT_COARSE | TK_PERSISTENT_FINE))
        return fail(TK_ERR_INVALID, "persistent takes TK_PERSISTENT_COARSE and TK_PERSISTENT_FINE bits only");
    auto pow2_in = [](uint32_t v, uint32_t lo, uint32_t hi) { return v >= lo && v <= hi && (v & (v - 1u)) == 0u; };
    if (!pow2_in(t->scan_block, 64u, 1024u)) return fail(TK_ERR_INVALID, "scan_block is 64, 128, 256, 512 or 1024");
    if (!pow2_in(t->coarse_block, 1u, 4u)) return fail(TK_ERR_INVALID, "coarse_block is 1, 2 or 4");
    if (!pow2_in(t->coarse_scan_items, 1024u, 8192u))
        return fail(TK_ERR_INVALID, "coarse_scan_items is 1024, 2048, 4096 or 8192");
    if (t->diff_keep_coverage > 1u) return fail(TK_ERR_INVALID, "diff_keep_coverage is 0 or 1");
    if (f->graph_exec) {  // the captured launches carry the old values
        TK_HIP(hipStreamSynchronize(f->ctx->stream));
        tk::drop_graph(f);
    }
    f->tuning = *t;
    return TK_OK;
}

// Coordinate search: fine first (the largest stage), then coarse, then scan; one
// parameter at a time over its candidates, the others held at the best so far,
// each candidate scored by the time of its own stage.
tk_status tk_frame_autotune(tk_frame *f, uint32_t repeats, tk_tuning *chosen) {
    if (!f) return fail(TK_ERR_INVALID, "null frame");
    repeats = std::max<uint32_t>(repeats, 1u);
    const bool was_timing = f->timing;
    f->timing = true;
    tk_tuning best = f->tuning;
    enum Stage { kFine, kCoarse, kScan };
    auto cost = [&](const tk_tuning &t, Stage stage, float *out) -> tk_status {
        TK_TRY(tk_frame_set_tuning(f, &t));
        TK_TRY(tk::frame_render(f));  // warm-up
        float sum = 0.0f;
        for (uint32_t r = 0; r < repeats; ++r) {
            TK_TRY(tk::frame_render(f));
            tk_stage_times st;
            TK_TRY(tk_frame_stage_times(f, &st));
            sum += stage == kFine     ? st.fine
                   : stage == kCoarse ? st.coarse_count + st.coarse_prefix + st.coarse_emit
                                      : st.scan_count + st.scan_emit;
        }
        *out = sum / float(repeats);
        return TK_OK;
    };
    // Candidate values of each parameter; the persistent axes toggle one bit.
    struct Axis {
        Stage stage;
        void (*apply)(tk_tuning &, uint32_t);
        std::vector<uint32_t> values;
    };
    const Axis axes[] = {
        {kFine, [](tk_tuning &t, uint32_t v) { t.fine_xcd_run = v; }, {0u, 32u, 64u, 128u, 256u, 512u}},
        {kFine,
         [](tk_tuning &t, uint32_t v) { t.persistent = (t.persistent & ~TK_PERSISTENT_FINE) | (v ? TK_PERSISTENT_FINE : 0u); },
         {0u, 1u}},
        {kCoarse, [](tk_tuning &t, uint32_t v) { t.coarse_xcd_run = v; }, {0u, 16u, 32u, 64u, 128u}},
        {kCoarse, [](tk_tuning &t, uint32_t v) { t.coarse_block = v; }, {1u, 2u, 4u}},
        {kCoarse,
         [](tk_tuning &t, uint32_t v) {
             t.persistent = (t.persistent & ~TK_PERSISTENT_COARSE) | (v ? TK_PERSISTENT_COARSE : 0u);
         },
         {0u, 1u}},
        {kCoarse, [](tk_tuning &t, uint32_t v) { t.coarse_scan_items = v; }, {1024u, 2048u, 4096u, 8192u}},
        {kScan, [](tk_tuning &t, uint32_t v) { t.scan_xcd_run = v; }, {0u, 16u, 64u}},
        {kScan, [](tk_tuning &t, uint32_t v) { t.scan_block = v; }, {64u, 128u, 256u, 512u, 1024u}},
    };
    tk_status st = TK_OK;
    for (const Axis &axis : axes) {
        if (st != TK_OK) break;
        float best_ms = 0.0f;
        tk_tuning axis_best = best;
        for (size_t i = 0; i < axis.values.size() && st == TK_OK; ++i) {
            tk_tuning t = best;
            axis.apply(t, axis.values[i]);
            float ms = 0.0f;
            st = cost(t, axis.stage, &ms);
            if (st == TK_OK && (i == 0 || ms < best_ms)) {
                best_ms = ms;
                axis_best = t;
            }
        }
        best = axis_best;
    }
    f->timing = was_timing;
    if (st != TK_OK) return st;
    TK_TRY(tk_frame_set_tuning(f, &best));
    TK_TRY(tk::frame_render(f));  // leave the scene's image in the frame
    if (chosen) *chosen = best;
    return TK_OK;
}

tk_status tk_frame_render_diff(tk_frame *f, const float *colors, float *image) {
    if (!f || !colors || !image) return fail(TK_ERR_INVALID, "null argument");
    if (!f->diff_scene)
        return fail(TK_ERR_INVALID,
                    "the differentiable form takes path draws (fills, strokes) and SDF clip layers only");
    const tk::DiffTarget t{colors, image};
    TK_TRY(tk::frame_render(f, 0, &t));
    f->diff_ready = true;
    return TK_OK;
}

tk_status tk_frame_backward(tk_frame *f, const float *colors, const float *image_grad, float *line_grad,
                            float *color_grad) {
    if (!f || !colors || !image_grad || !line_grad || !color_grad) return fail(TK_ERR_INVALID, "null argument");
    if (!f->diff_ready) return fail(TK_ERR_INVALID, "tk_frame_backward needs a tk_frame_render_diff of the scene first");
    tk_context *ctx = f->ctx;
    const tk_scene &s = f->meta;
    TK_TRY(tk::fill_u32(ctx, color_grad, uint64_t(s.draw_count) * 32, 0u));
    tk_args_fine_diff a = tk::diff_args(f, colors);
    a.image_grad = image_grad;
    a.color_grad = color_grad;
    const bool clips = s.layer_count != 0;
    hipFunction_t fn = f->scene_count ? (clips ? ctx->f_fine_backward_scenes_clips : ctx->f_fine_backward_scenes)
                                      : (clips ? ctx->f_fine_backward_clips : ctx->f_fine_backward);
    if (f->diff_kept) fn = ctx->f_fine_backward_keep[(f->scene_count ? 1 : 0) + (clips ? 2 : 0)];
    TK_TRY(tk::launch_tiles(ctx, fn, a.cfg.tile_count, a));
    if (s.line_count) {
        tk_args_line_backward b{f->lines, s.line_count, s.path_count, f->paths, f->diff_rec_of, s.backdrop_len,
                                f->diff_adj, f->diff_adj_records, line_grad};
        TK_TRY(tk::launch(ctx, ctx->f_line_backward, div_up(s.line_count, 256), 1, 256, b));
    }
    // The tile pass replaced the kept coverages with their gradients: another
    // backward needs another forward.
    f->diff_ready = false;
    return TK_OK;
}

tk_status tk_frame_diff_status(tk_frame *f, uint32_t *flags) {
    if (!f || !flags) return fail(TK_ERR_INVALID, "null argument");
    TK_HIP(hipMemcpyAsync(flags, f->diff_status, 4, hipMemcpyDeviceToHost, f->ctx->stream));
    TK_HIP(hipStreamSynchronize(f->ctx->stream));
    return TK_OK;
}

tk_status tk_frame_set_graph(tk_frame *f, int enable) {
    if (!f) return fail(TK_ERR_INVALID, "null frame");
    f->use_graph = enable != 0;
    if (!f->use_graph) {
        TK_HIP(hipStreamSynchronize(f->ctx->stream));
        drop_graph(f);
    }
    return TK_OK;
}

tk_status tk_frame_render_tiles(tk_frame *f, const uint32_t *tiles, uint32_t tile_count) {
    if (!f || (tile_count && !tiles)) return fail(TK_ERR_INVALID, "null argument");
    if (f->scene_count) return fail(TK_ERR_INVALID, "incremental rendering does not apply to a batch of scenes");
    const tk_scene &s = f->meta;
    const uint32_t tw = s.tiles_width, th = s.tiles_height, T = tw * th;
    // The host arrays are reused: wait until earlier uploads from them have completed.
    TK_HIP(hipStreamSynchronize(f->ctx->stream));
    std::vector<uint8_t> damaged(T, 0u);
    f->host_active.clear();
    for (uint32_t i = 0; i < tile_count; ++i) {
        const uint32_t t = tiles[i];
        if (t >= T) {
            set_error("tile index %u out of range (%u tiles)", t, T);
            return TK_ERR_INVALID;
        }
        if (!damaged[t]) {
            damaged[t] = 1u;
            f->host_active.push_back(t);
        }
    }
    const uint32_t active = uint32_t(f->host_active.size());
    if (active == 0) return TK_OK;
    const uint32_t w = tw + 1;
    f->host_sat.assign(uint64_t(w) * (th + 1), 0u);
    for (uint32_t y = 0; y < th; ++y)
        for (uint32_t x = 0; x < tw; ++x)
            f->host_sat[(y + 1) * w + x + 1] =
                damaged[y * tw + x] + f->host_sat[y * w + x + 1] + f->host_sat[(y + 1) * w + x] - f->host_sat[y * w + x];
    TK_TRY(upload(f, f->active_tiles, f->host_active.data(), active));
    TK_TRY(upload(f, f->damage_sat, f->host_sat.data(), f->host_sat.size()));
    return frame_render(f, active);
}

tk_status tk_frame_set_capture(tk_frame *f, int enable) {
    if (!f) return fail(TK_ERR_INVALID, "null frame");
    f->capture_raw = enable != 0;
    return TK_OK;
}

tk_status tk_frame_enable_timing(tk_frame *f, int enable) {
    if (!f) return fail(TK_ERR_INVALID, "null frame");
    f->timing = enable != 0;
    f->have_times = false;
    return TK_OK;
}

tk_status tk_frame_stage_times(tk_frame *f, tk_stage_times *out) {
    if (!f || !out) return fail(TK_ERR_INVALID, "null argument");
    if (!f->have_times) return fail(TK_ERR_INVALID, "no timed render recorded");
    TK_HIP(hipEventSynchronize(f->ev[10]));
    // Each stage from its own start event; the side-stream stages (cumsum, coarse)
    // overlap scan_prefix and scan_emit, so the stages add up to more than total.
    auto span = [&](int from, int to, float &ms) -> tk_status {
        TK_HIP(hipEventElapsedTime(&ms, f->ev[from], f->ev[to]));
        return TK_OK;
    };
    TK_TRY(span(0, 1, out->scan_clear));
    TK_TRY(span(1, 2, out->scan_count));
    TK_TRY(span(2, 3, out->scan_prefix));
    TK_TRY(span(3, 4, out->scan_emit));
    TK_TRY(span(5, 6, out->cumsum));
    TK_TRY(span(11, 7, out->coarse_count));
    TK_TRY(span(7, 8, out->coarse_prefix));
    TK_TRY(span(8, 9, out->coarse_emit));
    TK_TRY(span(12, 10, out->fine));
    float total = 0;
    TK_HIP(hipEventElapsedTime(&total, f->ev[0], f->ev[10]));
    out->total = total;
    return TK_OK;
}

tk_status tk_frame_read_image(tk_frame *f, uint32_t *rgba, size_t pixel_count) {
    if (!f || !rgba) return fail(TK_ERR_INVALID, "null argument");
    uint64_t n = uint64_t(f->meta.width) * f->meta.height;
    if (pixel_count < n) return fail(TK_ERR_INVALID, "output buffer too small");
    TK_HIP(hipMemcpyAsync(rgba, f->image, n * 4, hipMemcpyDeviceToHost, f->ctx->stream));
    TK_HIP(hipStreamSynchronize(f->ctx->stream));
    return TK_OK;
}

void *tk_frame_image_device(tk_frame *f) { return f ? f->image : nullptr; }

// --- asynchronous submissions -------------------------------------------------

struct tk_submission {
    hipEvent_t done = nullptr;
    uint32_t *pixels = nullptr;  // pinned: the image as of this submission
    uint32_t *status = nullptr;  // pinned: the frame's four status words
    uint64_t pixel_count = 0;
    uint64_t ptcl_capacity = 0;
};

tk_status tk_frame_submit(tk_frame *f, tk_submission **out) {
    if (!f || !out) return fail(TK_ERR_INVALID, "null argument");
    *out = nullptr;
    if (f->scene_count) return fail(TK_ERR_INVALID, "a batch's frame is read with tk_batch_read_image");
    TK_TRY(tk_frame_render(f));
    std::unique_ptr<tk_submission, void (*)(tk_submission *)> sub(new tk_submission, tk_submission_destroy);
    sub->pixel_count = uint64_t(f->meta.width) * f->meta.height;
    sub->ptcl_capacity = f->ptcl_capacity;
    TK_HIP(hipHostMalloc((void **)&sub->pixels, std::max<uint64_t>(sub->pixel_count, 1) * 4, hipHostMallocDefault));
    TK_HIP(hipHostMalloc((void **)&sub->status, 16, hipHostMallocDefault));
    TK_HIP(hipEventCreateWithFlags(&sub->done, hipEventDisableTiming));
    hipStream_t st = f->ctx->stream;
    if (sub->pixel_count)
        TK_HIP(hipMemcpyAsync(sub->pixels, f->image, sub->pixel_count * 4, hipMemcpyDeviceToHost, st));
    TK_HIP(hipMemcpyAsync(sub->status, f->status, 16, hipMemcpyDeviceToHost, st));
    TK_HIP(hipEventRecord(sub->done, st));
    *out = sub.release();
    return TK_OK;
}

tk_status tk_submission_is_complete(tk_submission *s, uint32_t *complete) {
    if (!s || !complete) return fail(TK_ERR_INVALID, "null argument");
    const hipError_t e = hipEventQuery(s->done);
    if (e == hipErrorNotReady) {
        *complete = 0u;
        return TK_OK;
    }
    TK_HIP(e);
    *complete = 1u;
    return TK_OK;
}

tk_status tk_submission_readback(tk_submission *s, uint32_t *rgba, size_t pixel_count) {
    if (!s || (!rgba && s->pixel_count)) return fail(TK_ERR_INVALID, "null argument");
    if (pixel_count < s->pixel_count) return fail(TK_ERR_INVALID, "output buffer too small");
    TK_HIP(hipEventSynchronize(s->done));
    const uint32_t *w = s->status;
    if (w[0] | w[2] | w[3] || w[1] > s->ptcl_capacity)
        return fail(TK_ERR_OVERFLOW, "a pipeline capacity was exceeded");
    if (s->pixel_count) std::memcpy(rgba, s->pixels, s->pixel_count * 4);
    return TK_OK;
}

void tk_submission_destroy(tk_submission *s) {
    if (!s) return;
    if (s->done) {
        (void)hipEventSynchronize(s->done);  // the copies into the pinned buffers end first
        (void)hipEventDestroy(s->done);
    }
    if (s->pixels) (void)hipHostFree(s->pixels);
    if (s->status) (void)hipHostFree(s->status);
    delete s;
}

tk_status tk_frame_read_buffer(tk_frame *f, tk_frame_buffer which, void *out, size_t bytes) {
    if (!f || !out) return fail(TK_ERR_INVALID, "null argument");
    const tk_scene &s = f->meta;
    const uint64_t QT = uint64_t(s.batch_count) * s.tiles_width * s.tiles_height;
    const void *src = nullptr;
    uint64_t n = 0;
    switch (which) {
    case TK_BUF_BACKDROPS: src = f->backdrops; n = uint64_t(s.backdrop_len) * 4; break;
    case TK_BUF_RAW_BACKDROPS: src = f->raw_backdrops; n = uint64_t(s.backdrop_len) * 4; break;
    case TK_BUF_SEGMENT_COUNTS: src = f->seg_counts; n = uint64_t(s.backdrop_len) * 4; break;
    case TK_BUF_SEGMENT_RANGES: src = f->seg_ranges; n = uint64_t(s.backdrop_len) * 8; break;
    case TK_BUF_SEGMENT_CURSORS: src = f->cursors; n = uint64_t(s.backdrop_len) * 4; break;
    case TK_BUF_SEGMENT_BUMPS: src = f->bumps; n = uint64_t(s.path_count) * 4; break;
    case TK_BUF_SEGMENTS: src = f->segments; n = uint64_t(s.segment_capacity) * 20; break;
    case TK_BUF_TILE_PTCL_COUNTS: src = f->tile_counts; n = QT * 4; break;
    case TK_BUF_TILE_PTCL_RANGES: src = f->tile_ranges; n = QT * 8; break;
    case TK_BUF_TILE_KINDS: src = f->tile_kinds; n = QT * 4; break;
    case TK_BUF_BATCH_PTCL_BASES: src = f->batch_bases; n = (uint64_t(s.batch_count) + 1) * 4; break;
    case TK_BUF_PTCL: src = f->ptcl; n = f->ptcl_capacity * 24; break;
    case TK_BUF_LINES: src = f->lines; n = uint64_t(s.line_count) * sizeof(tk_line); break;
    case TK_BUF_PATHS: src = f->paths; n = uint64_t(s.path_count) * sizeof(tk_path_record); break;
    case TK_BUF_DRAWS: src = f->draws; n = uint64_t(s.draw_count) * sizeof(tk_draw_record); break;
    case TK_BUF_TILE_DRAWS: src = f->tile_draws; n = uint64_t(s.tiles_width) * s.tiles_height * 8; break;
    case TK_BUF_TILE_DRAW_INDICES: src = f->tile_draw_indices; n = uint64_t(s.tile_draw_index_count) * 4; break;
    case TK_BUF_SCAN_CHUNKS: src = f->scan_chunks; n = uint64_t(s.scan_chunk_count) * sizeof(tk_scan_chunk); break;
    case TK_BUF_SCAN_CHUNK_RANGES: src = f->scan_chunk_ranges; n = uint64_t(s.path_count) * 8; break;
    case TK_BUF_CUMSUM_CHUNK_OFFSETS: src = f->cs_offsets; n = uint64_t(s.cumsum_chunk_count) * 4; break;
    case TK_BUF_CUMSUM_CHUNK_LENS: src = f->cs_lens; n = uint64_t(s.cumsum_chunk_count) * 4; break;
    case TK_BUF_CUMSUM_ROW_STARTS: src = f->cs_row_starts; n = uint64_t(s.cumsum_row_count) * 4; break;
    case TK_BUF_CUMSUM_ROW_ENDS: src = f->cs_row_ends; n = uint64_t(s.cumsum_row_count) * 4; break;
    default: return fail(TK_ERR_INVALID, "unknown buffer");
    }
    if (bytes > n) return fail(TK_ERR_INVALID, "read larger than the buffer");
    if (bytes == 0) return TK_OK;
    TK_HIP(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, f->ctx->stream));
    TK_HIP(hipStreamSynchronize(f->ctx->stream));
    return TK_OK;
}

tk_status tk_frame_stats_read(tk_frame *f, tk_frame_stats *out) {
    if (!f || !out) return fail(TK_ERR_INVALID, "null argument");
    const tk_scene &s = f->meta;
    uint32_t status[4] = {};
    TK_HIP(hipMemcpyAsync(status, f->status, 16, hipMemcpyDeviceToHost, f->ctx->stream));
    TK_HIP(hipStreamSynchronize(f->ctx->stream));
    std::memset(out, 0, sizeof(*out));
    out->paths = s.path_count;
    out->lines = s.line_count;
    out->draws = s.draw_count;
    out->batches = s.batch_count;
    out->tiles = s.tiles_width * s.tiles_height;
    out->backdrop_len = s.backdrop_len;
    out->segment_capacity = s.segment_capacity;
    out->tile_draw_refs = s.tile_draw_index_count;
    out->ptcl_capacity = f->ptcl_capacity;
    out->ptcl_bound = s.ptcl_capacity;
    out->scan_chunk_count = s.scan_chunk_count;
    out->cumsum_chunk_count = s.cumsum_chunk_count;
    out->cumsum_row_count = s.cumsum_row_count;
    out->ptcl_records = status[1];
    out->overflow = status[0] | (status[3] << 1) | (status[1] > f->ptcl_capacity ? 4u : 0u) | (status[2] << 4);
    uint32_t emitted = 0;
    if (s.path_count) {
        // Segments emitted = sum of per-path bumps.
        std::vector<uint32_t> bumps(s.path_count);
        TK_HIP(hipMemcpy(bumps.data(), f->bumps, uint64_t(s.path_count) * 4, hipMemcpyDeviceToHost));
        for (uint32_t b : bumps) emitted += b;
    }
    out->segments_emitted = emitted;
    if (emitted > s.segment_capacity) out->overflow |= 8u;
    out->device_bytes = f->device_bytes;
    return TK_OK;
}

tk_status tk_render_to_image(tk_context *ctx, const tk_scene *scene, uint32_t *rgba, size_t pixel_count) {
    tk_frame *f = nullptr;
    TK_TRY(tk_frame_create(ctx, scene, &f));
    tk_status st = tk_frame_render(f);
    if (st == TK_OK) {
        tk_frame_stats stats;
        st = tk_frame_stats_read(f, &stats);
        if (st == TK_OK && stats.overflow) st = fail(TK_ERR_OVERFLOW, "a pipeline capacity was exceeded");
    }
    if (st == TK_OK) st = tk_frame_read_image(f, rgba, pixel_count);
    tk_frame_destroy(f);
    return st;
}

// --- stage entry points -----------------------------------------------------

tk_status tk_scan_clear(tk_context *ctx, const tk_scan_config *cfg, void *backdrops, void *ranges, void *counts,
                        void *cursors, void *bumps, void *chunk_totals, void *chunk_offsets) {
    if (!ctx || !cfg) return fail(TK_ERR_INVALID, "null argument");
    if (cfg->incremental) return fail(TK_ERR_INVALID, "incremental scan is not supported");
    tk_args_scan_clear a{*cfg,
                         static_cast<int32_t *>(backdrops),
                         static_cast<tk_segment_range *>(ranges),
                         static_cast<uint32_t *>(counts),
                         static_cast<uint32_t *>(cursors),
                         static_cast<uint32_t *>(bumps),
                         static_cast<uint32_t *>(chunk_totals),
                         static_cast<uint32_t *>(chunk_offsets)};
    return launch(ctx, ctx->f_scan_clear, div_up(cfg->clear_len, 256), 1, 256, a);
}

tk_status tk_scan_count(tk_context *ctx, const tk_scan_config *cfg, const tk_line *lines, const tk_path_record *paths,
                        uint32_t path_record_count, int32_t *backdrops, uint32_t *counts) {
    if (!ctx || !cfg) return fail(TK_ERR_INVALID, "null argument");
    if (cfg->incremental) return fail(TK_ERR_INVALID, "incremental scan is not supported");
    if (cfg->line_count && (!lines || !backdrops || !counts)) return fail(TK_ERR_INVALID, "null buffer");
    return run_scan_count(ctx, *cfg, lines, paths, path_record_count, backdrops, counts);
}

tk_status tk_scan_prefix(tk_context *ctx, const tk_scan_config *cfg, const tk_scan_chunk *chunks,
                         const tk_scan_chunk_range *chunk_ranges, const tk_path_record *paths,
                         const uint32_t *counts, tk_segment_range *ranges, uint32_t *cursors, uint32_t *bumps,
                         uint32_t *chunk_totals, uint32_t *chunk_offsets) {
    if (!ctx || !cfg) return fail(TK_ERR_INVALID, "null argument");
    if (cfg->incremental) return fail(TK_ERR_INVALID, "incremental scan is not supported");
    return run_scan_prefix(ctx, *cfg, chunks, chunk_ranges, paths, counts, ranges, cursors, bumps, chunk_totals,
                           chunk_offsets);
}

tk_status tk_scan_emit(tk_context *ctx, const tk_scan_config *cfg, const tk_line *lines, const tk_path_record *paths,
                       uint32_t path_record_count, uint32_t *cursors, tk_line_segment *segments) {
    if (!ctx || !cfg) return fail(TK_ERR_INVALID, "null argument");
    if (cfg->incremental) return fail(TK_ERR_INVALID, "incremental scan is not supported");
    return run_scan_emit(ctx, *cfg, lines, paths, path_record_count, cursors, segments);
}

tk_status tk_cumsum(tk_context *ctx, const tk_cumsum_config *cfg, const uint32_t *offs, const uint32_t *lens,
                    const uint32_t *row_starts, const uint32_t *row_ends, int32_t *backdrops, uint32_t *totals,
                    uint32_t *carries) {
    if (!ctx || !cfg) return fail(TK_ERR_INVALID, "null argument");
    return run_cumsum(ctx, *cfg, offs, lens, row_starts, row_ends, backdrops, totals, carries);
}

tk_status tk_coarse(tk_context *ctx, const tk_coarse_config *cfg, const tk_coarse_inputs *in,
                    const tk_coarse_outputs *out) {
    if (!ctx || !cfg || !in || !out) return fail(TK_ERR_INVALID, "null argument");
    if (uint64_t(cfg->tiles_width) * cfg->tiles_height != cfg->tile_count)
        return fail(TK_ERR_INVALID, "tile_count must equal tiles_width * tiles_height");
    TK_TRY(ensure_status(ctx, coarse_scan_blocks(*cfg)));
    return run_coarse(ctx, *cfg, *in, *out, ctx->status, &ctx->tickets[1], ctx->next_epoch(), nullptr);
}

// The lookback ticket sits in the first 256 bytes of the scratch array, the
// status words (one u64 per scan workgroup) after it.  Both are cleared by a fill
// kernel at every call, so the epoch can stay fixed and nothing a previous call
// left behind is read.
static_assert(TK_NO_ALLOC_MAX_TABLE ==
                  uint64_t(TK_NO_ALLOC_SCRATCH_BYTES - TK_NO_ALLOC_TICKET_BYTES) / 8u * TK_COARSE_SCAN_ITEMS,
              "the no-alloc table limit follows from the scratch size");
tk_status tk_coarse_no_alloc(tk_context *ctx, const tk_coarse_config *cfg, const tk_coarse_inputs *in,
                             const tk_coarse_outputs *out) {
    if (!ctx || !cfg || !in || !out) return fail(TK_ERR_INVALID, "null argument");
    if (uint64_t(cfg->tiles_width) * cfg->tiles_height != cfg->tile_count)
        return fail(TK_ERR_INVALID, "tile_count must equal tiles_width * tiles_height");
    if (!ctx->no_alloc_scratch) return fail(TK_ERR_HIP, "the context has no no-alloc scratch");
    const uint64_t entries = uint64_t(cfg->batch_count) * cfg->tile_count;
    const uint32_t blocks = coarse_scan_blocks(*cfg);
    const uint64_t bytes = TK_NO_ALLOC_TICKET_BYTES + uint64_t(blocks) * sizeof(unsigned long long);
    if (bytes > TK_NO_ALLOC_SCRATCH_BYTES) {
        set_error("batch_count * tile_count = %llu exceeds the %llu entries tk_coarse_no_alloc supports",
                  (unsigned long long)entries, (unsigned long long)TK_NO_ALLOC_MAX_TABLE);
        return TK_ERR_INVALID;
    }
    uint8_t *scratch = ctx->no_alloc_scratch;
    TK_TRY(fill_u32(ctx, scratch, bytes / 4, 0u));
    return run_coarse(ctx, *cfg, *in, *out, reinterpret_cast<unsigned long long *>(scratch + TK_NO_ALLOC_TICKET_BYTES),
                      reinterpret_cast<uint32_t *>(scratch), 1u, nullptr);
}

tk_status tk_fine(tk_context *ctx, const tk_fine_config *cfg, const tk_fine_inputs *in, uint32_t *rgba) {
    if (!ctx || !cfg || !in || !rgba) return fail(TK_ERR_INVALID, "null argument");
    if (uint64_t(cfg->tiles_width) * cfg->tiles_height != cfg->tile_count)
        return fail(TK_ERR_INVALID, "tile_count must equal tiles_width * tiles_height");
    TK_TRY(ensure_draw_brush(ctx, 2 * size_t(cfg->draw_count)));
    return run_fine(ctx, *cfg, *in, rgba, ctx->draw_brush);
}

tk_status tk_fill_coverage_words(tk_context *ctx, uint32_t count, const uint32_t *requests,
                                 const tk_line_segment *segments, uint32_t *alpha_out, float *coverage_out) {
    if (!ctx || (count && (!requests || !alpha_out))) return fail(TK_ERR_INVALID, "null argument");
    tk_args_fill_coverage a{count, requests, segments, alpha_out, coverage_out};
    return launch(ctx, ctx->f_fill_coverage, div_up(count, 256), 1, 256, a);
}

tk_status tk_pixel_math_words(tk_context *ctx, uint32_t count, const uint32_t *requests, uint32_t *out) {
    if (!ctx || (count && (!requests || !out))) return fail(TK_ERR_INVALID, "null argument");
    tk_args_pixel_math a{count, requests, out};
    return launch(ctx, ctx->f_pixel_math, div_up(count, 256), 1, 256, a);
}

tk_status tk_brush_words(tk_context *ctx, const uint32_t *paint, uint32_t paint_words, uint32_t paint_brush_base,
                         uint32_t count, const uint32_t *requests, uint32_t *out) {
    if (!ctx || (count && (!requests || !out || (paint_words && !paint)))) return fail(TK_ERR_INVALID, "null argument");
    if (!count) return TK_OK;
    tk_args_brush_words a{count, requests, paint, paint_words, paint_brush_base, out};
    return launch(ctx, ctx->f_brush_words, div_up(count, 256), 1, 256, a);
}

}  // extern "C"
