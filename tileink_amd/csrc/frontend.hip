// GPU front end for polygon scenes (tk_frame_set_polygons), included by
// kernels.hip so that its kernels share the one gfx950 code object.
//
// Given new vertices for every path of a frame whose paths are closed polygons,
// these kernels rebuild on the device everything the host recorder derives from
// geometry when it records the same polygons (tk_canvas_push_polygons) and
// encodes the scene (tk_canvas_encode):
//   - the lines, with degenerate edges dropped (src/shared/path_flatten.rs:59-82);
//   - each path record: pixel and tile bounds, backdrop slab, segment capacity
//     from line_scanned_tile_count (src/canvas/paths.rs:65-156);
//   - the path draws' pixel bounds;
//   - the scan and cumsum chunk plans (src/shared/gpu_plan/path_plans.rs:156-228,463-504);
//   - the painter-ordered tile bins minus the tiles early occlusion hides
//     (src/shared/gpu_plan/tile_bins/build.rs:4-83, transient.rs:7-146), with the
//     per-draw cover rectangles the recorder computes once (draw_covers);
//   - the PTCL capacity bound.
// The work runs in two phases with one host round trip between them: counts
// and prefix sums (whose totals size the frame's buffers), then the fills.
// Every result is what the recorder produces for the same vertices, bit for bit:
// the same FP64 bounds, the same FP32 tile counts (scan_line.h) and the same
// integer rules.

#include "flatten.h"
#include "scan_line.h"

namespace {

__device__ __forceinline__ uint32_t fe_scene_of(const uint32_t *map, uint32_t i) { return map ? map[i] : 0u; }

__device__ __forceinline__ int32_t fe_sat_i32(double v) {
    if (!(v == v)) return 0;
    if (v <= -2147483648.0) return INT32_MIN;
    if (v >= 2147483647.0) return INT32_MAX;
    return (int32_t)v;
}

__device__ __forceinline__ double fe_wave_min(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double fe_wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ unsigned long long fe_wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += (unsigned long long)__shfl_xor((long long)v, off, 64);
    return v;
}

// Edge i of a closed polygon of n vertices starting at vertex v0: from vertex i
// to vertex i + 1 (the last closes back to the first).  Degenerate when its
// squared length in FP64 is at most 1e-24 (is_non_degenerate).
struct FeEdge {
    float ax, ay, bx, by;
    bool keep;
};
// The vertex coordinates in FP64 (exact for FP32 input); the line stores them
// rounded to FP32, as push_line does.
__device__ __forceinline__ FeEdge fe_edge(const tk_args_frontend &a, uint32_t v0, uint32_t n, uint32_t i) {
    FeEdge e{0.0f, 0.0f, 0.0f, 0.0f, false};
    if (i >= n) return e;
    const uint32_t j = i + 1u == n ? 0u : i + 1u;
    double ax, ay, bx, by;
    if (a.points64) {
        ax = a.points64[2u * (v0 + i)];
        ay = a.points64[2u * (v0 + i) + 1u];
        bx = a.points64[2u * (v0 + j)];
        by = a.points64[2u * (v0 + j) + 1u];
    } else {
        ax = a.points[2u * (v0 + i)];
        ay = a.points[2u * (v0 + i) + 1u];
        bx = a.points[2u * (v0 + j)];
        by = a.points[2u * (v0 + j) + 1u];
    }
    const double dx = ax - bx, dy = ay - by;
    e.keep = dx * dx + dy * dy > 1.0e-12 * 1.0e-12;
    e.ax = (float)ax;
    e.ay = (float)ay;
    e.bx = (float)bx;
    e.by = (float)by;
    return e;
}

}  // namespace

// Phase 1, one wave per path: bounds, line count, tile box, backdrop slab,
// segment capacity and chunk counts.
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_paths(tk_args_frontend a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t p = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (p >= a.path_count) return;
    const tk_fe_scene sc = a.scenes[fe_scene_of(a.scene_of_path, p)];
    const uint32_t v0 = a.starts[p], v1 = a.starts[p + 1u];
    uint32_t n = v1 > v0 ? v1 - v0 : 0u;
    double x0 = __builtin_inf(), y0 = __builtin_inf(), x1 = -__builtin_inf(), y1 = -__builtin_inf();
    bool finite = true;
    for (uint32_t i = lane; i < n; i += 64u) {
        const double x = a.points64 ? a.points64[2u * (v0 + i)] : (double)a.points[2u * (v0 + i)];
        const double y = a.points64 ? a.points64[2u * (v0 + i) + 1u] : (double)a.points[2u * (v0 + i) + 1u];
        finite = finite && isfinite(x) && isfinite(y);
        x0 = fmin(x0, x);
        y0 = fmin(y0, y);
        x1 = fmax(x1, x);
        y1 = fmax(y1, y);
    }
    if (__ballot(!finite)) {
        if (lane == 0u) atomicOr(a.status, 1u);
        n = 0u;  // the host refuses the update; keep every count bounded meanwhile
    }
    x0 = fe_wave_min(x0);
    y0 = fe_wave_min(y0);
    x1 = fe_wave_max(x1);
    y1 = fe_wave_max(y1);
    // path_bounds of a polygon: every vertex once the path has an edge (a line
    // verb); of flattened curves: the curves' exact bounds, given.
    if (a.bounds64 && n) {
        x0 = a.bounds64[4u * p];
        y0 = a.bounds64[4u * p + 1u];
        x1 = a.bounds64[4u * p + 2u];
        y1 = a.bounds64[4u * p + 3u];
    }
    int32_t pb[4] = {0, 0, 0, 0};
    if (a.bounds64 ? (n && x0 <= x1) : n >= 2u) {
        pb[0] = fe_sat_i32(floor(x0));
        pb[1] = fe_sat_i32(floor(y0));
        pb[2] = fe_sat_i32(ceil(x1));
        pb[3] = fe_sat_i32(ceil(y1));
    }
    uint32_t tb[4];
    tk::tile_box_of(pb[0], pb[1], pb[2], pb[3], sc.tiles_width, sc.tiles_height, tb);
    uint32_t lines = 0u;
    unsigned long long capacity = 0ull;
    for (uint32_t base = 0; base < n; base += 64u) {
        const FeEdge e = fe_edge(a, v0, n, base + lane);
        lines += (uint32_t)__popcll(__ballot(e.keep));
        if (e.keep) capacity += tk::scanned_tile_count(e.ax, e.ay, e.bx, e.by, tb[0], tb[1], tb[2], tb[3]);
    }
    capacity = fe_wave_sum_u64(capacity);
    if (lane != 0u) return;
    const uint32_t stride = tb[2] > tb[0] ? tb[2] - tb[0] : 0u, height = tb[3] > tb[1] ? tb[3] - tb[1] : 0u;
    const uint32_t data_len = stride * height;
    const uint32_t rows = stride && height ? height : 0u;
    const uint32_t P = a.path_count;
    a.path_vals[TK_FE_LINES * P + p] = lines;
    a.path_vals[TK_FE_BACKDROPS * P + p] = data_len;
    a.path_vals[TK_FE_SEGMENTS * P + p] = capacity > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)capacity;
    a.path_vals[TK_FE_SCAN_CHUNKS * P + p] = (data_len + TK_SCAN_CHUNK_SIZE - 1u) / TK_SCAN_CHUNK_SIZE;
    a.path_vals[TK_FE_CUMSUM_ROWS * P + p] = rows;
    a.path_vals[TK_FE_CUMSUM_CHUNKS * P + p] = rows * ((stride + TK_CUMSUM_CHUNK_SIZE - 1u) / TK_CUMSUM_CHUNK_SIZE);
    for (int k = 0; k < 4; ++k) a.path_pb[4u * p + k] = pb[k];
    tk_path_record r{};
    r.path_id = p;
    r.line_count = lines;
    r.flags = 0u;
    r.data_len = data_len;
    r.tile_x0 = tb[0];
    r.tile_y0 = tb[1];
    r.tile_x1 = tb[2];
    r.tile_y1 = tb[3];
    r.segment_capacity = a.path_vals[TK_FE_SEGMENTS * P + p];
    r.segment_count = 0u;
    r.transform = tk_affine{1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f};
    a.paths[p] = r;  // line_start, data_offset and segment_start follow the scan
}

// Phase 1, one thread per draw: a path draw takes its path's pixel bounds.
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_draw_bounds(tk_args_frontend a) {
    const uint32_t d = blockIdx.x * 256u + threadIdx.x;
    if (d >= a.draw_count) return;
    const uint32_t path = a.draws[d].path_id;
    if (path >= a.path_count) return;
    const tk_pixel_bounds pb{a.path_pb[4u * path], a.path_pb[4u * path + 1u], a.path_pb[4u * path + 2u],
                             a.path_pb[4u * path + 3u]};
    a.draws[d].pixel_bounds = pb;
    a.draws[d].local_pixel_bounds = pb;
}

// Phase 1, one thread per batch: the tile box its clip stack allows.
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_batch_clips(tk_args_frontend a) {
    const uint32_t b = blockIdx.x * 256u + threadIdx.x;
    if (b >= a.batch_count) return;
    const tk_fe_scene sc = a.scenes[fe_scene_of(a.scene_of_batch, b)];
    uint32_t clip[4] = {0u, 0u, sc.tiles_width, sc.tiles_height};
    const tk_batch_record br = a.batches[b];
    for (uint32_t l = br.layer_start; l < br.layer_end && l < a.layer_count; ++l) {
        const uint32_t cd = a.layers[l].draw;
        if (cd >= a.draw_count) continue;
        const tk_pixel_bounds pb = a.draws[cd].pixel_bounds;
        uint32_t t[4];
        tk::tile_box_of(pb.x0, pb.y0, pb.x1, pb.y1, sc.tiles_width, sc.tiles_height, t);
        clip[0] = max(clip[0], t[0]);
        clip[1] = max(clip[1], t[1]);
        clip[2] = min(clip[2], t[2]);
        clip[3] = min(clip[3], t[3]);
    }
    for (int k = 0; k < 4; ++k) a.batch_clip[4u * b + k] = clip[k];
}

// Phase 1, one thread per draw of a batch: its tile box, its share of the PTCL
// bound, its rows, and the visible runs of each row (the box minus its cover).
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_draw_count(tk_args_frontend a) {
    const uint32_t d = blockIdx.x * 256u + threadIdx.x;
    if (d >= a.draw_count) return;
    const uint32_t b = a.batch_ids[d];
    uint32_t tb[4] = {0u, 0u, 0u, 0u};
    if (b < a.batch_count) {
        const tk_fe_scene sc = a.scenes[fe_scene_of(a.scene_of_draw, d)];
        const tk_pixel_bounds pb = a.draws[d].pixel_bounds;
        tk::tile_box_of(pb.x0, pb.y0, pb.x1, pb.y1, sc.tiles_width, sc.tiles_height, tb);
        if (tb[2] > tb[0] && tb[3] > tb[1]) {
            const uint32_t *clip = a.batch_clip + 4u * b;
            const uint32_t cx0 = max(tb[0], clip[0]), cy0 = max(tb[1], clip[1]);
            const uint32_t cx1 = min(tb[2], clip[2]), cy1 = min(tb[3], clip[3]);
            if (cx1 > cx0 && cy1 > cy0)
                atomicAdd(&a.batch_area[b], (unsigned long long)(cx1 - cx0) * (unsigned long long)(cy1 - cy0));
            const uint32_t *cv = a.covers + 4u * d;
            const uint32_t hx0 = max(tb[0], cv[0]), hy0 = max(tb[1], cv[1]);
            const uint32_t hx1 = min(tb[2], cv[2]), hy1 = min(tb[3], cv[3]);
            const bool hides = hx1 > hx0 && hy1 > hy0;
            const uint64_t base = (uint64_t)sc.column_start + sc.row_start;
            for (uint32_t y = tb[1]; y < tb[3]; ++y) {
                atomicAdd(&a.row_counts[sc.row_start + y], 1u);
                int32_t *row = a.diff + base + (uint64_t)y * (sc.tiles_width + 1u);
                if (hides && y >= hy0 && y < hy1) {
                    if (hx0 > tb[0]) {
                        atomicAdd(&row[tb[0]], 1);
                        atomicAdd(&row[hx0], -1);
                    }
                    if (tb[2] > hx1) {
                        atomicAdd(&row[hx1], 1);
                        atomicAdd(&row[tb[2]], -1);
                    }
                } else {
                    atomicAdd(&row[tb[0]], 1);
                    atomicAdd(&row[tb[2]], -1);
                }
            }
        } else {
            tb[0] = tb[1] = tb[2] = tb[3] = 0u;
        }
    }
    for (int k = 0; k < 4; ++k) a.draw_box[4u * d + k] = tb[k];
}

// Phase 1, one thread per batch: per batch, its particles plus (2 depth + 1)
// per tile that can hold one (the recorder's PTCL bound).
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_batch_caps(tk_args_frontend a) {
    const uint32_t b = blockIdx.x * 256u + threadIdx.x;
    if (b >= a.batch_count) return;
    const uint32_t *clip = a.batch_clip + 4u * b;
    const unsigned long long clip_area = (clip[2] > clip[0] && clip[3] > clip[1])
                                             ? (unsigned long long)(clip[2] - clip[0]) * (clip[3] - clip[1])
                                             : 0ull;
    const unsigned long long area = a.batch_area[b];
    const tk_batch_record br = a.batches[b];
    const unsigned long long depth = br.layer_end > br.layer_start ? br.layer_end - br.layer_start : 0u;
    atomicAdd(&a.totals[TK_FE_PTCL], area + (2ull * depth + 1ull) * (clip_area < area ? clip_area : area));
}

// Phase 1, one thread per tile row: the row's tile list lengths from its runs.
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_tile_counts(tk_args_frontend a) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= a.row_count) return;
    const tk_fe_scene sc = a.scenes[fe_scene_of(a.scene_of_row, r)];
    const uint32_t y = r - sc.row_start;
    const int32_t *row = a.diff + (uint64_t)sc.column_start + sc.row_start + (uint64_t)y * (sc.tiles_width + 1u);
    uint32_t *counts = a.tile_counts + sc.column_start + (uint64_t)y * sc.tiles_width;
    int32_t run = 0;
    for (uint32_t x = 0; x < sc.tiles_width; ++x) {
        run += row[x];
        counts[x] = (uint32_t)run;
    }
}

// In-place exclusive scans of several u32 columns (blocks of 1024 values).
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_scan_blocks(tk_args_fe_scan a) {
    __shared__ unsigned long long s_sum[4];
    const uint32_t col = blockIdx.y, blk = blockIdx.x;
    const uint32_t *v = a.values + (uint64_t)col * a.count;
    unsigned long long sum = 0ull;
    for (uint32_t k = 0; k < 4u; ++k) {
        const uint64_t i = (uint64_t)blk * 1024u + k * 256u + threadIdx.x;
        if (i < a.count) sum += v[i];
    }
    sum = fe_wave_sum_u64(sum);
    if ((threadIdx.x & 63u) == 0u) s_sum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0u) a.block_sums[(uint64_t)col * a.blocks + blk] = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
}

extern "C" __global__ __launch_bounds__(256) void tk_k_fe_scan_top(tk_args_fe_scan a) {
    __shared__ unsigned long long s_wave[4];
    const uint32_t col = blockIdx.x;
    unsigned long long *sums = a.block_sums + (uint64_t)col * a.blocks;
    unsigned long long carry = 0ull;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t base = 0; base < a.blocks; base += 256u) {
        const uint32_t i = base + threadIdx.x;
        const unsigned long long v = i < a.blocks ? sums[i] : 0ull;
        unsigned long long x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned long long y = (unsigned long long)__shfl_up((long long)x, off, 64);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63u) s_wave[wave] = x;
        __syncthreads();
        unsigned long long before = 0ull, total = 0ull;
        for (uint32_t w = 0; w < 4u; ++w) {
            if (w < wave) before += s_wave[w];
            total += s_wave[w];
        }
        if (i < a.blocks) sums[i] = carry + before + x - v;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0u) {
        a.totals[col] = carry;
        if (carry > 0xFFFFFFFFull) atomicOr(a.overflow, 1u);
    }
}

extern "C" __global__ __launch_bounds__(256) void tk_k_fe_scan_apply(tk_args_fe_scan a) {
    __shared__ uint32_t s_wave[4];
    const uint32_t col = blockIdx.y, blk = blockIdx.x;
    uint32_t *v = a.values + (uint64_t)col * a.count;
    const uint64_t first = (uint64_t)blk * 1024u + threadIdx.x * 4u;
    uint32_t x[4], sum = 0u;
    for (uint32_t k = 0; k < 4u; ++k) {
        x[k] = first + k < a.count ? v[first + k] : 0u;
        sum += x[k];
    }
    uint32_t total;
    uint32_t run = (uint32_t)a.block_sums[(uint64_t)col * a.blocks + blk] + block_exclusive_scan_256(sum, s_wave, total);
    for (uint32_t k = 0; k < 4u; ++k) {
        if (first + k < a.count) v[first + k] = run;
        run += x[k];
    }
}

// Phase 2, one wave per path: the lines, the path record's starts, and the
// path's scan and cumsum chunks.
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_path_fill(tk_args_frontend a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t p = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (p >= a.path_count) return;
    const uint32_t P = a.path_count;
    const uint32_t line_start = a.path_vals[TK_FE_LINES * P + p];
    const uint32_t data_offset = a.path_vals[TK_FE_BACKDROPS * P + p];
    const uint32_t segment_start = a.path_vals[TK_FE_SEGMENTS * P + p];
    const uint32_t chunk_start = a.path_vals[TK_FE_SCAN_CHUNKS * P + p];
    const uint32_t row_start = a.path_vals[TK_FE_CUMSUM_ROWS * P + p];
    const uint32_t cs_start = a.path_vals[TK_FE_CUMSUM_CHUNKS * P + p];
    tk_path_record &rec = a.paths[p];
    const uint32_t data_len = rec.data_len;
    const uint32_t stride = rec.tile_x1 > rec.tile_x0 ? rec.tile_x1 - rec.tile_x0 : 0u;
    const uint32_t height = rec.tile_y1 > rec.tile_y0 ? rec.tile_y1 - rec.tile_y0 : 0u;
    const uint32_t n_lines = rec.line_count;
    if (lane == 0u) {
        rec.line_start = line_start;
        rec.data_offset = data_offset;
        rec.segment_start = segment_start;
    }
    // lines, in edge order
    const uint32_t v0 = a.starts[p], v1 = a.starts[p + 1u];
    const uint32_t n = n_lines ? (v1 > v0 ? v1 - v0 : 0u) : 0u;
    uint32_t written = 0u;
    for (uint32_t base = 0; base < n; base += 64u) {
        const FeEdge e = fe_edge(a, v0, n, base + lane);
        const uint64_t keep = __ballot(e.keep);
        const uint64_t slot = (uint64_t)line_start + written + (uint32_t)__popcll(keep & lanes_below(lane));
        if (e.keep && slot < a.line_capacity) {
            a.lines[slot] = tk_line{p, 0.0f, {e.ax, e.ay}, {e.bx, e.by}};
            if (a.line_vertex) a.line_vertex[slot] = v0 + base + lane;
        }
        written += (uint32_t)__popcll(keep);
    }
    // scan chunks: 256 backdrop tiles each
    const uint32_t n_chunks = (data_len + TK_SCAN_CHUNK_SIZE - 1u) / TK_SCAN_CHUNK_SIZE;
    for (uint32_t k = lane; k < n_chunks; k += 64u) {
        const uint32_t local = k * TK_SCAN_CHUNK_SIZE;
        const uint32_t len = min(data_len - local, TK_SCAN_CHUNK_SIZE);
        if ((uint64_t)chunk_start + k < a.chunk_capacity)
            a.scan_chunks[chunk_start + k] = tk_scan_chunk{p, data_offset + local, segment_start, len};
    }
    if (lane == 0u) a.chunk_ranges[p] = tk_scan_chunk_range{chunk_start, chunk_start + n_chunks};
    // cumsum chunks: per tile row of the slab, 256 tiles each
    if (!stride || !height) return;
    const uint32_t per_row = (stride + TK_CUMSUM_CHUNK_SIZE - 1u) / TK_CUMSUM_CHUNK_SIZE;
    for (uint32_t r = lane; r < height; r += 64u) {
        const uint32_t first = cs_start + r * per_row;
        if ((uint64_t)row_start + r < a.cs_row_capacity) {
            a.cs_row_starts[row_start + r] = first;
            a.cs_row_ends[row_start + r] = first + per_row;
        }
        for (uint32_t j = 0; j < per_row; ++j) {
            if ((uint64_t)first + j >= a.cs_chunk_capacity) break;
            a.cs_offsets[first + j] = data_offset + r * stride + j * TK_CUMSUM_CHUNK_SIZE;
            a.cs_lens[first + j] = min(stride - j * TK_CUMSUM_CHUNK_SIZE, TK_CUMSUM_CHUNK_SIZE);
        }
    }
}

// Phase 2, one thread per column: its tile list record.
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_tile_records(tk_args_frontend a) {
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= a.column_count) return;
    const uint32_t start = a.tile_counts[c];
    const uint32_t next = c + 1u < a.column_count ? a.tile_counts[c + 1u] : (uint32_t)a.totals[TK_FE_TILE_REFS];
    const uint32_t len = next - start;
    a.tile_draws[c] = tk_tile_draw_record{len ? start : kNone, len};
}

// Phase 2, one workgroup per tile row: the row's draws in ascending order (the
// painter order of the batched draws), each with its x range and the x range its
// cover hides in this row.
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_rows(tk_args_frontend a) {
    __shared__ uint32_t s_wave[4];
    const uint32_t r = blockIdx.x;
    if (r >= a.row_count) return;
    const tk_fe_scene sc = a.scenes[fe_scene_of(a.scene_of_row, r)];
    const uint32_t y = r - sc.row_start;
    uint64_t out = a.row_counts[r];
    for (uint32_t base = sc.draw_start; base < sc.draw_end; base += 256u) {
        const uint32_t d = base + threadIdx.x;
        uint4 e = make_uint4(d, 0u, 0u, 0u);
        bool take = false;
        if (d < sc.draw_end) {
            const uint4 tb = reinterpret_cast<const uint4 *>(a.draw_box)[d];
            take = tb.w > tb.y && y >= tb.y && y < tb.w;
            if (take) {
                const uint32_t *cv = a.covers + 4u * d;
                const uint32_t hx0 = max(tb.x, cv[0]), hy0 = max(tb.y, cv[1]);
                const uint32_t hx1 = min(tb.z, cv[2]), hy1 = min(tb.w, cv[3]);
                const bool hides = hx1 > hx0 && hy1 > hy0 && y >= hy0 && y < hy1;
                e = make_uint4(d, tb.x, tb.z, hides ? (hx0 | (hx1 << 16)) : 0u);
            }
        }
        uint32_t total;
        const uint32_t at = block_exclusive_scan_256(take ? 1u : 0u, s_wave, total);
        if (take && out + at < a.row_capacity) a.row_entries[out + at] = e;
        out += total;
    }
}

// Phase 2, one workgroup per tile row, one lane per tile: each lane walks the
// row's draws in order and appends those visible in its tile to its list.
extern "C" __global__ __launch_bounds__(256) void tk_k_fe_tiles(tk_args_frontend a) {
    const uint32_t r = blockIdx.x;
    if (r >= a.row_count) return;
    const tk_fe_scene sc = a.scenes[fe_scene_of(a.scene_of_row, r)];
    const uint32_t y = r - sc.row_start;
    const uint32_t first = a.row_counts[r];
    const uint32_t last = r + 1u < a.row_count ? a.row_counts[r + 1u] : (uint32_t)a.totals[TK_FE_ROW_ENTRIES];
    for (uint32_t xb = 0; xb < sc.tiles_width; xb += 256u) {
        const uint32_t x = xb + threadIdx.x;
        const bool mine = x < sc.tiles_width;
        const uint64_t column = (uint64_t)sc.column_start + (uint64_t)y * sc.tiles_width + x;
        uint64_t dst = mine ? a.tile_counts[column] : 0u;
        for (uint32_t k = first; k < last && (uint64_t)k < a.row_capacity; ++k) {
            const uint4 e = a.row_entries[k];
            const uint32_t hx0 = e.w & 0xFFFFu, hx1 = e.w >> 16;
            if (mine && x >= e.y && x < e.z && !(x >= hx0 && x < hx1)) {
                if (dst < a.index_capacity) a.tile_draw_indices[dst] = e.x;
                ++dst;
            }
        }
    }
}

// GPU flattening: one thread per path, whose closed contour of cubics goes
// through the shared flattener (flatten.h) exactly as push_path records it: the
// start point, then each cubic's points after its start; the path's bounds are
// the cubics' exact bounds (add_cubic_bounds).  Pass 1 counts and bounds,
// pass 2 (after a scan of the counts) writes the points.
__device__ __forceinline__ bool fl_contour(const tk_args_flatten &a, uint32_t p, uint32_t &first, uint32_t &cubics) {
    first = a.starts[p];
    const uint32_t end = a.starts[p + 1u];
    const uint32_t n = end > first ? end - first : 0u;
    cubics = n >= 4u ? (n - 1u) / 3u : 0u;
    return n >= 4u && (n - 1u) % 3u == 0u;
}

extern "C" __global__ __launch_bounds__(256) void tk_k_flatten_count(tk_args_flatten a) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= a.path_count) return;
    uint32_t first, cubics;
    uint32_t count = 0u;
    double x0 = __builtin_inf(), y0 = __builtin_inf(), x1 = -__builtin_inf(), y1 = -__builtin_inf();
    if (!fl_contour(a, p, first, cubics)) {
        atomicOr(a.status, 2u);
    } else {
        const double *c = a.ctrl + 2u * first;
        bool finite = true;
        for (uint32_t i = 0; i < 2u * (3u * cubics + 1u); ++i) finite = finite && isfinite(c[i]);
        if (!finite) {
            atomicOr(a.status, 1u);
        } else {
            count = 1u;
            auto add = [&](tk::FPt q) {
                x0 = fmin(x0, q.x);
                y0 = fmin(y0, q.y);
                x1 = fmax(x1, q.x);
                y1 = fmax(y1, q.y);
            };
            for (uint32_t k = 0; k < cubics; ++k) {
                const tk::FPt p0{c[6u * k], c[6u * k + 1u]}, p1{c[6u * k + 2u], c[6u * k + 3u]},
                    p2{c[6u * k + 4u], c[6u * k + 5u]}, p3{c[6u * k + 6u], c[6u * k + 7u]};
                tk::flatten_cubic(p0, p1, p2, p3, a.tolerance, [&](tk::FPt) { ++count; });
                add(p0);
                add(p3);
                double ts[2];
                int nt;
                tk::cubic_axis_extrema(p0.x, p1.x, p2.x, p3.x, ts, &nt);
                for (int i = 0; i < nt; ++i)
                    add(tk::FPt{tk::cubic_coord(p0.x, p1.x, p2.x, p3.x, ts[i]),
                                tk::cubic_coord(p0.y, p1.y, p2.y, p3.y, ts[i])});
                tk::cubic_axis_extrema(p0.y, p1.y, p2.y, p3.y, ts, &nt);
                for (int i = 0; i < nt; ++i)
                    add(tk::FPt{tk::cubic_coord(p0.x, p1.x, p2.x, p3.x, ts[i]),
                                tk::cubic_coord(p0.y, p1.y, p2.y, p3.y, ts[i])});
            }
        }
    }
    a.counts[p] = count;
    a.bounds[4u * p] = x0;
    a.bounds[4u * p + 1u] = y0;
    a.bounds[4u * p + 2u] = x1;
    a.bounds[4u * p + 3u] = y1;
}

extern "C" __global__ __launch_bounds__(256) void tk_k_flatten_emit(tk_args_flatten a) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= a.path_count) return;
    const uint32_t at = a.counts[p];
    const uint32_t next = p + 1u < a.path_count ? a.counts[p + 1u] : 0xFFFFFFFFu;
    if (next == at) return;  // a path without points (malformed or non-finite)
    uint32_t first, cubics;
    if (!fl_contour(a, p, first, cubics)) return;
    const double *c = a.ctrl + 2u * first;
    uint64_t w = at;
    auto put = [&](tk::FPt q) {
        if (w < a.point_capacity) {
            a.points[2u * w] = q.x;
            a.points[2u * w + 1u] = q.y;
        }
        ++w;
    };
    put(tk::FPt{c[0], c[1]});
    for (uint32_t k = 0; k < cubics; ++k) {
        const tk::FPt p0{c[6u * k], c[6u * k + 1u]}, p1{c[6u * k + 2u], c[6u * k + 3u]},
            p2{c[6u * k + 4u], c[6u * k + 5u]}, p3{c[6u * k + 6u], c[6u * k + 7u]};
        tk::flatten_cubic(p0, p1, p2, p3, a.tolerance, put);
    }
}
