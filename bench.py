"""Benchmark: one frame of the compute-raster pipeline on one MI355X.

Workload: 4096x4096 RGBA8 canvas, 200,000 paths (60% fill, 40% stroke, 4-16
cubics each, 10% linear gradients, mixed fill rules) under 32 nested rounded-rect
clip layers, synthetic and seeded.  A step renders the whole frame (scan, backdrop
prefix, coarse, fine for every DrawBatch) from scene buffers already resident in
HBM.  Prints one JSON line.
"""
import argparse
import hashlib
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import tileink_amd as tk  # noqa: E402
from tileink_amd import _abi, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def fine_walked_work(frame, scene, stats):
    """Records and tile segments the fine stage reads: per (batch, tile) stream the
    records from start + opaque-prefix skip to END, and the segments of the FILL /
    BEGIN_CLIP records among them in INTERPRETER tiles."""
    Q, T = scene.batch_count, scene.tiles_width * scene.tiles_height
    K = int(stats["ptcl_records"])
    kinds = frame.buffer(_abi.BUF_TILE_KINDS, np.uint32, Q * T)
    ranges = frame.buffer(_abi.BUF_TILE_PTCL_RANGES, np.uint32, 2 * Q * T).reshape(-1, 2).astype(np.int64)
    ptcl = frame.buffer(_abi.BUF_PTCL, np.uint32, 6 * K).reshape(-1, 6)
    counts = ranges[:, 1] - ranges[:, 0]
    # The ranges are one exclusive scan in (batch, tile) order, so record k belongs
    # to the k-th position of the streams laid end to end.
    has = np.nonzero(counts > 0)[0]
    stream = np.repeat(has, counts[has])[:K]
    index = np.arange(stream.size, dtype=np.int64)
    kind = kinds[stream] & 7
    walked = (index >= ranges[stream, 0] + (kinds[stream] >> 3)) & (kind != _abi.TK_TILE_KIND_EMPTY)
    rec = ptcl[:stream.size]
    fills = walked & (kind == _abi.TK_TILE_KIND_INTERPRETER) & ((rec[:, 0] == 1) | (rec[:, 0] == 3))
    seg = rec[fills, 4].astype(np.int64) - rec[fills, 3].astype(np.int64)
    return int(walked.sum()), int(np.clip(seg, 0, None).sum())


def fine_algorithmic_bytes(walked_records, used_segments, scene):
    """Minimum HBM traffic of one fine launch: every walked PTCL record (24 B) and
    every segment those records use (20 B) read once, the (range, kind) word pair
    (12 B) of every batch x tile read once, the RGBA8 image (4 B/pixel) written once."""
    QT = scene.batch_count * scene.tiles_width * scene.tiles_height
    return 24 * walked_records + 20 * used_segments + 12 * QT + 4 * scene.width * scene.height


def source_digest():
    h = hashlib.sha256()
    for rel in ("tileink_amd/csrc/kernels.hip", "tileink_amd/csrc/raster_device.h"):
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(kernel):
    """HBM bytes per launch from the committed --pmc profile of this kernel source,
    or None when there is none for the current source."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic_4096px_200k_paths.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("source_digest") != source_digest():
        return None
    return rec.get("hbm_bytes_per_launch")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_threads():
    """The host cores this process may use: its CPU share (OMP_NUM_THREADS, set to the
    share on the GPU boxes) within its affinity mask."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(avail, share) if share > 0 else avail)


def cpu_baseline(args, arrays):
    """The oracle (the C restatement of every stage) on the host cores, timed on the
    same frame the GPU renders, threaded over lines, chunks and tiles; and, beside
    it, one thread on a bounded sample of the generator (1024x1024, 1/16 of the
    paths, the same clip schedule), since one thread takes minutes on the full frame."""
    import oracle
    threads = cpu_threads()
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    _, out = oracle.render_frame(arrays, outputs=False)
    dt = time.perf_counter() - t0
    side, paths = 1024, args.paths // 16
    sample = scenes.clipped_paths_scene(side, side, paths, clips=args.clips, workload=2).arrays()
    oracle.set_threads(1)
    t1 = time.perf_counter()
    _, out1 = oracle.render_frame(sample, outputs=False)
    dt1 = time.perf_counter() - t1
    W, H = arrays.width, arrays.height
    return {"value": W * H / dt / 1e9, "unit": "Gpix/s", "cores": threads, "kind": "port",
            "sample": f"oracle, whole frame of this workload ({W}x{H}, {arrays.path_count} paths), {threads} "
                      f"threads on {cpu_model()} ({os.cpu_count()} logical CPUs, {dt:.1f} s)",
            "ms_per_frame": round(dt * 1e3, 1),
            "stage_ms": {k: round(v, 1) for k, v in out["stage_ms"].items()},
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "single_thread": {"sample": f"{side}x{side}, {paths} paths, {args.clips} nested clips",
                              "ms_per_frame": round(dt1 * 1e3, 1), "gpix_per_s": side * side / dt1 / 1e9,
                              "stage_ms": {k: round(v, 1) for k, v in out1["stage_ms"].items()}}}


def incremental_render(frame, tiles_w, tiles_h, steps, side=32):
    """Damage-proportional rendering: the same frame re-rendered on a side x side
    block of tiles at the canvas centre (tk_frame_render_tiles), averaged over
    `steps` renders.  Reported beside the full-frame metric, not instead of it."""
    x0, y0 = (tiles_w - side) // 2, (tiles_h - side) // 2
    tiles = np.array([(y0 + y) * tiles_w + x0 + x for y in range(side) for x in range(side)], np.uint32)
    frame.render_tiles(tiles)  # warm-up
    total = {}
    for _ in range(steps):
        frame.render_tiles(tiles)
        for k, v in frame.stage_times().items():
            total[k] = total.get(k, 0.0) + v / steps
    frame.render()  # back to full tables for anything read afterwards
    return {"damaged_tiles": int(tiles.size), "ms": round(total["total"], 4),
            "stage_ms": {k: round(v, 4) for k, v in total.items()}}


def scene_loop(ctx, count=512, repeats=3):
    """The independent-scene workload: 512 scenes, canvas side log-uniform in
    64-2048 px, aspect 0.5-2, path counts lognormal(ln 1500, 1.2) clipped to
    10-20 000, each recorded by the clipped-paths generator (scenes.scene_set, in
    worker processes).  Each scene is its own frame in the per-scene loop, timed
    with kernel-by-kernel launches and with each frame replayed from its hipGraph;
    the loop issues every frame's launches back to back on the context's stream
    with no host synchronisation between scenes.  The same scenes are then
    rendered as one batch (tk_batch_*: one frame, one launch per stage for all of
    them), eager and replayed from a hipGraph."""
    t0 = time.perf_counter()
    sizes = scenes.scene_set_sizes(count)
    arrays = scenes.scene_set(sizes)
    encoded = [a.to_scene() for a in arrays]
    log(f"scene set: {count} scenes, {sum(n for _, _, n in sizes)} paths generated in {time.perf_counter() - t0:.1f} s")
    frames = [tk.Frame(ctx, sc) for sc in encoded]
    out = {"scenes": count, "paths": int(sum(n for _, _, n in sizes)),
           "lines": int(sum(a.line_count for a in arrays)),
           "pixels": int(sum(w * h for w, h, _ in sizes))}

    def timed(name, render_all):
        render_all()  # warm-up (captures graphs)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            render_all()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) / repeats * 1e3
        out[f"{name}_ms"] = round(ms, 3)
        out[f"{name}_scenes_per_s"] = round(count / (ms * 1e-3), 1)

    for mode in ("eager", "graph"):
        for f in frames:
            f.set_graph(mode == "graph")
        timed(mode, lambda: [f.render() for f in frames])
    out["graph_speedup"] = round(out["eager_ms"] / out["graph_ms"], 3)
    del frames
    batch = tk.Batch(ctx, encoded)
    t0 = time.perf_counter()
    out["batch_tuning"] = batch.frame.autotune(repeats=2)
    log(f"scene set: batch autotune {out['batch_tuning']} in {time.perf_counter() - t0:.1f} s")
    for mode in ("eager", "graph"):
        batch.frame.set_graph(mode == "graph")
        timed(f"batch_{mode}", batch.render)
    loop_ms = min(out["eager_ms"], out["graph_ms"])
    out["batch_speedup"] = round(loop_ms / min(out["batch_eager_ms"], out["batch_graph_ms"]), 3)
    del batch
    return out


def diff_batch(ctx, count=256, repeats=3, gradients=0.1, clips=True):
    """Forward + backward of the differentiable form over many scenes: 256 scenes of
    256x256, each with 64-512 closed paths of 4 cubics under one rounded-rectangle
    clip (inset 4-24 px, radius 4-16 px), 10 % of the paints gradients of 2-4 stops
    (half linear, half radial) and the rest solid, loss = squared error against a random target.  A step renders every scene in FP32
    (tk_frame_render_diff) and back-propagates dL/d(image) to line endpoints and
    colours (tk_frame_backward), from scenes already resident in HBM.  Timed once as
    a per-scene loop (one frame per scene) and once as one batch.  `gradients`
    (the share of gradient paints) and `clips` vary the workload for comparison."""
    import torch
    from tileink_amd import diff
    rng = np.random.default_rng(5)
    paths = rng.integers(64, 513, count)
    sc = [diff.random_scene(1000 + i, 256, 256, fills=int(n), cubics=4, radius_mu=np.log(16.0))
          for i, n in enumerate(paths)]
    for i, (s, n) in enumerate(zip(sc, paths)):
        inset = rng.uniform(4, 24, 4)
        if clips:
            s.clips = [(inset[0], inset[1], 256 - inset[2], 256 - inset[3], float(rng.uniform(4, 16)), 0, int(n))]
        diff.random_gradients(s, gradients, 5000 + i, radial=0.5)
    polys = [s.polygons() for s in sc]
    singles = [diff.Plan(ctx, [s], [v], [k]) for s, (v, k) in zip(sc, polys)]
    batch = diff.Plan(ctx, sc, [v for v, _ in polys], [k for _, k in polys])
    colors = [p.full_paint(s.paint_records()).to("cuda", torch.float32) for p, s in zip(singles, sc)]
    gen = torch.Generator(device="cuda").manual_seed(5)
    targets = [torch.rand((p.pixels, 4), device="cuda", generator=gen) for p in singles]
    target_b = torch.zeros((batch.pixels, 4), device="cuda")
    for (w, h, _), off, t in zip(batch.sizes, batch.offsets, targets):
        target_b[off:off + w * h] = t
    color_b = batch.full_paint(torch.cat([s.paint_records() for s in sc])).to("cuda", torch.float32)

    def step(plan, c, t):
        image = plan.forward(c)
        plan.backward(2.0 * (image - t))

    out = {"scenes": count, "size": "256x256", "paths": int(paths.sum()), "lines": int(batch.line_count),
           "clips_per_scene": int(bool(clips)), "gradient_paints": int(sum(len(s.gradients) for s in sc)),
           "radial_paints": int(sum(diff.gradient_parts(g)[3] == "radial" for s in sc for g in s.gradients.values()))}
    def set_sync(plans, sync):
        for p in plans:
            p.sync = sync

    # loop: each call waits for the device and reads the status flags;
    # loop_async: the same plans with stream-ordered calls only (no host wait
    # between scenes), the flags read once after the timed steps.
    for name, run, sync in (("loop", lambda: [step(p, c, t) for p, c, t in zip(singles, colors, targets)], True),
                            ("loop_async", lambda: [step(p, c, t) for p, c, t in zip(singles, colors, targets)], False),
                            ("batch", lambda: step(batch, color_b, target_b), True)):
        set_sync(singles, sync)
        run()  # warm-up: sizes the coverage buffers
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            run()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / repeats * 1e3
        if not sync:
            for p in singles:
                p.check_status()
        out[f"{name}_ms"] = round(ms, 3)
        out[f"{name}_scenes_per_s"] = round(count / (ms * 1e-3), 1)
    out["batch_speedup"] = round(out["loop_ms"] / out["batch_ms"], 3)
    out["batch_speedup_vs_async_loop"] = round(out["loop_async_ms"] / out["batch_ms"], 3)
    # The batch again with the forward's coverages kept for the backward instead of
    # recomputed there (tk_tuning.diff_keep_coverage).
    batch.frame.set_tuning(diff_keep_coverage=1)
    step(batch, color_b, target_b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        step(batch, color_b, target_b)
    torch.cuda.synchronize()
    out["batch_keep_coverage_ms"] = round((time.perf_counter() - t0) / repeats * 1e3, 3)
    batch.frame.set_tuning(diff_keep_coverage=0)
    del singles, batch
    return out


def fit(ctx, count=256, steps=200, lr=0.05, loop_steps=3):
    """Batched image fitting with the differentiable form: 256 scenes of 256x256 as
    in diff_batch; each scene's target is the render of a copy whose control points
    are perturbed by N(0, 4 px) and colours (and gradient stops) by N(0, 0.05);
    `steps` Adam steps on control points, colours and stops (FP32, on the GPU)
    against the squared error, the batch's parameters held as three tensors that
    every scene's tensors are slices of.  A renderer's first call records and encodes the
    scenes on the host (timed apart); every step evaluates the geometry and paint
    of all scenes at once (diff.batch_polygons, diff.batch_paint_records), gives the GPU front end
    only the moved vertices (diff.Renderer -> tk_frame_set_polygons), then runs
    forward + backward for all scenes at once.  Beside it, `loop_steps` steps of
    the per-scene loop (each scene its own Renderer, one frame per scene), its
    native calls ordered on streams with no host wait per scene.  Returns the loss
    curve and the per-step times."""
    import torch
    from tileink_amd import diff
    rng = np.random.default_rng(9)
    paths = rng.integers(64, 513, count)
    make = lambda i, n: diff.random_scene(3000 + i, 256, 256, fills=int(n), cubics=4, radius_mu=np.log(16.0),
                                          dtype=torch.float32, device="cuda")
    scenes_, targets = [], []
    for i, n in enumerate(paths):
        t, s = make(i, n), make(i, n)
        inset = rng.uniform(4, 24, 4)
        clip = [(inset[0], inset[1], 256 - inset[2], 256 - inset[3], float(rng.uniform(4, 16)), 0, int(n))]
        t.clips, s.clips = clip, list(clip)
        diff.random_gradients(t, 0.1, 7000 + i, radial=0.5)
        diff.random_gradients(s, 0.1, 7000 + i, radial=0.5)
        with torch.no_grad():
            for g in t.gradients.values():
                stops = g[2]
                stops += torch.as_tensor(rng.normal(0, 0.05, tuple(stops.shape)), dtype=torch.float32, device="cuda")
                stops.clamp_(0, 1)
            t.fills += torch.as_tensor(rng.normal(0, 4.0, tuple(t.fills.shape)), dtype=torch.float32, device="cuda")
            t.colors += torch.as_tensor(rng.normal(0, 0.05, tuple(t.colors.shape)), dtype=torch.float32, device="cuda")
            t.colors.clamp_(0, 1)
        scenes_.append(s)
        targets.append(t)
    with torch.no_grad():
        ref = [im.clone() for im in diff.render(ctx, targets, "cuda")]
    # The optimiser's parameters as three tensors (all control points, all colours,
    # all stop colours); each scene's tensors are slices of them laid end to end,
    # which the batched geometry and paint take whole (one leaf each for autograd
    # and Adam instead of one per scene and per gradient).
    fills = torch.cat([s.fills for s in scenes_]).requires_grad_(True)
    colors = torch.cat([s.colors for s in scenes_]).requires_grad_(True)
    stop_tensors = [g[2] for s in scenes_ for _, g in sorted(s.gradients.items())]
    stops = torch.cat(stop_tensors).requires_grad_(True) if stop_tensors else None
    f0 = c0 = g0 = 0
    for s in scenes_:
        n = s.fills.shape[0]
        s.fills = fills[f0:f0 + n]
        f0 += n
        n = s.colors.shape[0]
        s.colors = colors[c0:c0 + n]
        c0 += n
        for d, entry in sorted(s.gradients.items()):
            line, offsets, st, kind = diff.gradient_parts(entry)
            s.gradients[d] = (line, offsets, stops[g0:g0 + st.shape[0]], kind)
            g0 += st.shape[0]
    params = [p for p in (fills, colors, stops) if p is not None]


    def run(renderers, n_steps, params, clamped):
        opt = torch.optim.Adam(params, lr=lr)
        losses, step_ms = [], []
        refs = []  # each renderer's target in its flat pixel layout (zero padding)
        first = 0
        t0 = time.perf_counter()
        for r in renderers:
            r()  # the first call records and encodes on the host
            flat_ref = torch.zeros_like(r.flat.detach())
            for (w, h, _), off in zip(r.plan.sizes, r.plan.offsets):
                flat_ref[off:off + w * h] = ref[first].reshape(-1, 4)
                first += 1
            refs.append(flat_ref)
        torch.cuda.synchronize()
        record_ms = (time.perf_counter() - t0) * 1e3
        for _ in range(n_steps):
            t0 = time.perf_counter()
            opt.zero_grad()
            for r in renderers:
                r()
            loss = sum(((r.flat - t) ** 2).sum() for r, t in zip(renderers, refs))
            loss.backward()
            opt.step()
            with torch.no_grad():
                if clamped:  # colours and stop colours
                    torch._foreach_clamp_min_(clamped, 0.0)
                    torch._foreach_clamp_max_(clamped, 1.0)
            torch.cuda.synchronize()
            step_ms.append((time.perf_counter() - t0) * 1e3)
            losses.append(round(float(loss.detach()), 3))
        return losses, step_ms, record_ms

    start = [p.detach().clone() for p in params]
    batch = diff.Renderer(ctx, scenes_, "cuda", sync=False)
    losses, step_ms, record_ms = run([batch], steps, params, params[1:])
    batch.plan.check_status()
    # The per-scene loop from the same starting values, each scene with its own
    # parameter tensors (as a loop over independent scenes holds them), its native
    # calls ordered on streams with no host wait per scene.
    with torch.no_grad():
        for p, v in zip(params, start):
            p.copy_(v)
    loop_params, loop_clamped = [], []
    for s in scenes_:
        s.fills = s.fills.detach().clone().requires_grad_(True)
        s.colors = s.colors.detach().clone().requires_grad_(True)
        loop_params += [s.fills, s.colors]
        loop_clamped.append(s.colors)
        for d, entry in sorted(s.gradients.items()):
            line, offsets, st, kind = diff.gradient_parts(entry)
            st = st.detach().clone().requires_grad_(True)
            s.gradients[d] = (line, offsets, st, kind)
            loop_params.append(st)
            loop_clamped.append(st)
    loop = [diff.Renderer(ctx, s, "cuda", sync=False) for s in scenes_]
    loop_losses, loop_ms, loop_record_ms = run(loop, loop_steps, loop_params, loop_clamped)
    for r in loop:
        r.plan.check_status()
    med = float(np.median(step_ms[1:])) if len(step_ms) > 1 else step_ms[0]
    loop_med = float(np.median(loop_ms[1:])) if len(loop_ms) > 1 else loop_ms[0]
    return {"scenes": count, "size": "256x256", "paths": int(paths.sum()), "clips_per_scene": 1,
            "gradient_paints": int(sum(len(s.gradients) for s in scenes_)), "steps": steps,
            "loss": losses, "loss_ratio": round(losses[-1] / losses[0], 4),
            "record_ms": round(record_ms, 1), "loop_record_ms": round(loop_record_ms, 1),
            "first_step_ms": round(step_ms[0], 1), "step_ms_median": round(med, 2),
            "scenes_per_s": round(count / (med * 1e-3), 1),
            "loop_steps": loop_steps, "loop_step_ms_median": round(loop_med, 2),
            "loop_scenes_per_s": round(count / (loop_med * 1e-3), 1), "loop_loss": loop_losses,
            "speedup": round(loop_med / med, 3),
            "note": "each renderer's first call records and encodes on the host (record_ms, untimed in the "
                    "steps); every step moves the vertices through the GPU front end (tk_frame_set_polygons)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--paths", type=int, default=200_000)
    ap.add_argument("--clips", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scene-loop", action="store_true", help="skip the many-small-scenes loop")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable forward+backward batch")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch parameters")
    ap.add_argument("--no-incremental", action="store_true", help="skip the damaged-tile re-render")
    ap.add_argument("--diff-only", action="store_true",
                    help="run only the differentiable forward+backward batch and print its figures as JSON")
    ap.add_argument("--fit", type=int, default=0, metavar="STEPS",
                    help="run only batched Adam image fitting for STEPS steps and print its loss curve as JSON")
    ap.add_argument("--fit-steps", type=int, default=200,
                    help="Adam steps of the batched image fitting in the default run (0: skip it)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device three times with tk_k_copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the frame would shard by tile "
              f"stripes, one stripe of the canvas per GPU, and a batch of scenes by scene)", file=sys.stderr)
        sys.exit(2)

    if args.fit:
        ctx = tk.NativeContext(0)
        print(json.dumps({"fit": fit(ctx, steps=args.fit)}), flush=True)
        return
    if args.diff_only:
        ctx = tk.NativeContext(0)
        print(json.dumps({"diff_fwd_bwd": diff_batch(ctx)}), flush=True)
        return

    t0 = time.perf_counter()
    canvas = scenes.clipped_paths_scene(args.size, args.size, args.paths, clips=args.clips, workload=2)
    scene = canvas.encode(0)
    log(f"scene: {scene.path_count} paths, {scene.line_count} lines, {scene.batch_count} batches, "
        f"{scene.tile_draw_index_count} tile refs, generated in {time.perf_counter() - t0:.1f} s")

    ctx = tk.NativeContext(0)
    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        src, dst = tk.DeviceBuffer(ctx, nbytes), tk.DeviceBuffer(ctx, nbytes)
        for _ in range(3):
            _abi.check(_abi.lib().tk_copy_device(ctx.handle, dst.ptr, src.ptr, nbytes))
        log(f"calibration: copied {args.calibrate_copy} MiB three times with tk_k_copy")
        del src, dst
    frame = tk.Frame(ctx, scene)
    for _ in range(args.warmup):
        frame.render()
    ctx.synchronize()
    stats = frame.stats()
    if stats["overflow"]:
        raise SystemExit(f"capacity overflow flags {stats['overflow']}")
    log(f"warmup done: K={stats['ptcl_records']} PTCL records, S={stats['segments_emitted']} segments")
    tuning = frame.tuning()
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuning = frame.autotune(repeats=3)
        log(f"autotune: {tuning} in {time.perf_counter() - t0:.1f} s")

    # Per-stage HIP events on the pipeline's stream, over the timed steps.
    frame.enable_timing(True)
    stage_sum = {}
    try:
        import torch
        torch_sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only; the pipeline does not need it
        torch_sync = lambda: None  # noqa: E731
    torch_sync()
    ctx.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        frame.render()
        times = frame.stage_times()  # waits for this frame's last event
        for k, v in times.items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    ctx.synchronize()
    torch_sync()
    wall = time.perf_counter() - t_start
    stage_avg = {k: v / args.steps for k, v in stage_sum.items()}
    ms_per_step = stage_avg["total"]
    log("stages (ms): " + ", ".join(f"{k}={v:.3f}" for k, v in stage_avg.items()))

    W = H = args.size
    N = W * H
    gpix = N / (ms_per_step * 1e-3) / 1e9
    walked_records, used_segments = fine_walked_work(frame, scene, stats)
    fine_bytes = fine_algorithmic_bytes(walked_records, used_segments, scene)
    fine_gbs = fine_bytes / (stage_avg["fine"] * 1e-3) / 1e9
    incremental = incremental_render(frame, W // 16, H // 16, args.steps) if not args.no_incremental else None
    loop = scene_loop(ctx) if not args.no_scene_loop else None
    fwd_bwd = diff_batch(ctx) if not args.no_diff else None
    if fwd_bwd:
        log(f"differentiable fwd+bwd: {fwd_bwd}")
    fitting = None
    if args.fit_steps and not args.no_diff:
        fitting = fit(ctx, steps=args.fit_steps)
        curve = fitting.pop("loss")
        # the loss every 10 steps and at the last step
        fitting["loss_curve"] = {str(i): curve[i] for i in sorted(set(range(0, len(curve), 10)) | {len(curve) - 1})}
        log(f"fitting: {fitting}")
    result = {
        "metric": "ms/frame & Gpix/s at 4096², 200k paths; batched fwd+bwd scenes/s vs plain",
        "value": round(gpix, 4),
        "unit": "Gpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain forward render, {W}x{H} RGBA8, {scene.path_count} paths, {args.clips} nested "
                        f"rounded-rect clips, 16x16 tiles",
            "paths_per_s": round(scene.path_count / (ms_per_step * 1e-3), 1),
            "lines": scene.line_count, "batches": scene.batch_count, "tiles": (W // 16) * (H // 16),
            "ptcl_records": int(stats["ptcl_records"]), "segments": int(stats["segments_emitted"]),
            "wall_ms_per_step": round(wall / args.steps * 1e3, 4),
            "stage_ms": {k: round(v, 4) for k, v in stage_avg.items()},
            "incremental": incremental,
            "tuning": tuning,
            "scene_loop": loop,
            "diff_fwd_bwd": fwd_bwd,
            "fit": fitting,
        },
        "roofline": {"bound": "hbm", "kernel": "tk_k_fine", "achieved": round(fine_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(fine_gbs / HBM_PEAK_GBS, 5), "traffic": measured_traffic("tk_k_fine"),
                     "algorithmic_bytes": fine_bytes, "walked_records": walked_records,
                     "used_segments": used_segments, "fine_ms": round(stage_avg["fine"], 4)},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline: the oracle on the same frame ...")
        result["cpu_baseline"] = cpu_baseline(args, canvas.arrays())
        log(f"cpu baseline: {result['cpu_baseline']['ms_per_frame']} ms/frame")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
