#!/usr/bin/env python3
"""Benchmark of the exact-OIT forward frame on one MI355X.

Workload (the reference metric's configuration): 1,000,000 soup triangles at 1920x1080,
centroids in 64 log-normal-weighted Gaussian blobs (sigma 0.1 NDC), log-normal edges
(median 6 px), depth U[1,1000], fp32 per-vertex RGB + per-triangle alpha, seed 1001.
A step is one full forward frame (setup, binning, raster_low, raster_high, exact
fallback) with all inputs already resident in HBM.

Prints ONE JSON line: value = fragments blended per second; ms_per_step = ms/frame;
roofline = the dominant kernel's algorithmic bytes / its average duration (HIP events)
against MI355X's 8 TB/s HBM peak; cpu_baseline = the C oracle on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/frame and fragments blended/s @1080p 1M tris; fwd+bwd vs plain fwd ratio"
HBM_PEAK_GBS = 8000.0
DEFAULT_TRIS = 1_000_000
# Measured on this workload at the default tuning with rocprofv3 (profiles/c1_forward_kernels.md):
# HBM bytes of the block-raster stage (block_raster + block_raster_big) per frame, from
# FETCH_SIZE and WRITE_SIZE passes scaled by a 256 MiB copy, and the VALU issue of the same
# kernels (SQ_INSTS_VALU over kernel time x 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per
# wave64 instruction). No counters are read in-process.
RASTER_TRAFFIC_BYTES = 373.5e6
RASTER_VALU_BUSY = 1.02  # above 1: faster than one wave64 instruction per 4 cycles


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run a 256 MiB device copy (reference kernel for rocprof counters)")
    p.add_argument("--tris", type=int, default=DEFAULT_TRIS)
    p.add_argument("--no-backward", action="store_true",
                   help="skip the fwd+bwd / fwd ratio leg (200k triangles at 512x512)")
    p.add_argument("--bwd-steps", type=int, default=5)
    p.add_argument("--no-unsorted", action="store_true",
                   help="skip the unsorted-blend baseline (the same frame without depth order)")
    p.add_argument("--no-batch", action="store_true",
                   help="skip the batch leg (64 scenes in one hipGraph vs 64 sequential frames)")
    p.add_argument("--batch-scenes", type=int, default=64)
    p.add_argument("--only-backward", action="store_true",
                   help="profiling aid: run only the forward+backward leg and print its JSON")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the default tuning instead of autotuning on the workload first")
    return p.parse_args()


def batch_leg(num_scenes: int, steps: int = 3):
    """Batch throughput: the C3 scenes (config0 generator, log-normal 5k-500k triangles,
    resolutions {256^2, 512^2, 768x432, 1024^2}) rendered as one hipGraph launch vs the same
    frames as sequential lucid_render + lucid_finish calls. Each resolution class is
    autotuned once on its largest scene; the tuning is applied to the whole class."""
    import torch
    from lucid_amd import scenes
    from lucid_amd.batch import LucidBatch
    from lucid_amd.renderer import GpuScene, LucidRenderer
    scs = scenes.config3(num_scenes=num_scenes)
    rs = [LucidRenderer((s.width, s.height), max_quads=s.num_quads) for s in scs]
    gs = [GpuScene(s) for s in scs]
    outs = [r.alloc_outputs(rgba8=True, rgba32f=False, counts=False) for r in rs]
    tuned = {}
    for res in sorted({(s.width, s.height) for s in scs}):
        idx = max((i for i, s in enumerate(scs) if (s.width, s.height) == res),
                  key=lambda i: scs[i].num_quads)
        tuned[res] = rs[idx].autotune(gs[idx], outs[idx], iters=2)
    for s, r in zip(scs, rs):
        r.set_tuning(**tuned[(s.width, s.height)])
    frags = 0
    for r, g, o in zip(rs, gs, outs):
        r.render(g, o)
        frags += int(r.read_info()[60])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for r, g, o in zip(rs, gs, outs):
            r.render(g, o)
    torch.cuda.synchronize()
    t_seq = (time.perf_counter() - t0) / steps
    batch = LucidBatch(rs)
    batch.render(gs, outs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.render(gs, outs)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    return {"workload": f"{num_scenes} scenes (config0 generator, log-normal 5k-500k triangles, "
                        "resolutions 256^2/512^2/768x432/1024^2, seed 1003), RGBA8",
            "triangles": int(sum(s.num_quads for s in scs)), "fragments": frags,
            "sequential_ms": t_seq * 1e3, "batch_ms": t_batch * 1e3,
            "batch_fragments_per_s": frags / t_batch, "speedup": t_seq / t_batch,
            "tuning": {f"{w}x{h}": t for (w, h), t in tuned.items()}}


def unsorted_baseline(scene, tuning, steps, exact_ms):
    """The same frame with LUCID_OPT_UNSORTED: samples blended in block-list order, as
    hardware alpha blending would (the reference's SimpleRenderer), timed like the exact
    frame (back-to-back frames between HIP events). The ratio is what exact order costs."""
    import torch
    from lucid_amd import _lib
    from lucid_amd.renderer import GpuScene, LucidRenderer
    r = LucidRenderer((scene.width, scene.height), opts=_lib.OPT_UNSORTED,
                      max_quads=scene.num_quads)
    r.set_tuning(**tuning)
    g = GpuScene(scene)
    out = r.alloc_outputs(rgba8=True, counts=True)
    for _ in range(3):
        r.render(g, out)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        r.render_async(g, out)
    e1.record(stream)
    torch.cuda.synchronize()
    _lib.check(r.finish(), "unsorted frames")
    ms = e0.elapsed_time(e1) / steps
    return {"ms_per_frame": ms, "exact_over_unsorted": exact_ms / ms}


def backward_ratio(steps: int):
    """fwd+bwd vs plain fwd on the differentiable workload: 200k soup triangles at 512x512
    (the config0 generator, ~160 fragments/pixel), fp32 samples. A fwd+bwd step is a plain
    forward frame followed by the backward pass for a dL/dimage, as the autograd binding runs
    them by default: the backward recomputes every pixel's depth-sorted fragment list (it
    stores none from the forward) and reuses only the frame's bins and per-block triangle
    lists. Timed beside it: the full recompute (lucid_backward rebuilds bins and lists too),
    and the opt-in recording forward (the forward stores each pixel's blend order for the
    backward, which then skips its first pass)."""
    import torch
    from lucid_amd import _lib, scenes
    from lucid_amd.renderer import GpuScene, LucidRenderer
    sc = scenes.config4()
    r = LucidRenderer((sc.width, sc.height), opts=_lib.OPT_FP32_SAMPLES, max_quads=sc.num_quads)
    g = GpuScene(sc)
    out = r.alloc_outputs(rgba8=False, rgba32f=True, counts=False)
    grad = torch.randn((sc.height, sc.width, 4), dtype=torch.float32, device="cuda") * 1e-3
    r.render(g, out)
    frags = int(r.read_info()[60])
    r.backward(g, grad)
    r.render(g, out)
    r.backward(g, grad, reuse_frame=True)
    for _ in range(2):  # the first recording frame may size the backward buffers instead
        r.render(g, out, record_backward=True)
        r.backward(g, grad, reuse_frame=True)
    torch.cuda.synchronize()

    def timed(backward=None, record=False):
        t0 = time.perf_counter()
        for _ in range(steps):
            r.render(g, out, record_backward=record)
            if backward == "reuse":
                r.backward(g, grad, reuse_frame=True)
            elif backward == "full":
                r.backward(g, grad)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_fwd = timed()
    t_both = timed("reuse")
    t_full = timed("full")
    t_rec = timed("reuse", record=True)
    return {"workload": "200k soup triangles at 512x512 (config0 generator, seed 1004), fp32 "
                        "samples; host-timed synchronous steps; per-pixel sorted fragment "
                        "lists recomputed in the backward",
            "fragments": frags, "fwd_ms": t_fwd * 1e3, "fwd_bwd_ms": t_both * 1e3,
            "fwd_bwd_ratio": t_both / t_fwd,
            "fwd_bwd_full_recompute_ms": t_full * 1e3,
            "fwd_bwd_full_recompute_ratio": t_full / t_fwd,
            "fwd_bwd_recording_forward_ms": t_rec * 1e3,
            "fwd_bwd_recording_forward_ratio": t_rec / t_fwd}


def algorithmic_bytes(scene, r, info):
    """Per-kernel and per-frame algorithmic bytes (each input read once, each output
    written once) from the frame's own counters."""
    import numpy as np
    nb = r.bin_count
    INFO = 1152
    cnt = info[INFO:]
    qc, qo = cnt[0:nb].astype(np.int64), cnt[nb:2 * nb].astype(np.int64)
    tc, to = cnt[3 * nb:4 * nb].astype(np.int64), cnt[4 * nb:5 * nb].astype(np.int64)
    w, h = scene.width, scene.height
    # setup records: 84 B per visible triangle + 20 B per visible quad
    vis_q = int(info[1]) + int(info[2])
    t_vis = 2 * vis_q
    if scene.quad_indices.shape[0] and np.all(scene.quad_indices[:, 2] == scene.quad_indices[:, 3]):
        t_vis = vis_q  # degenerate quads (a, b, c, c): one triangle each
    b_in = scene.input_bytes()
    e_total = int(qc.sum() + tc.sum())
    P = w * h
    frame = b_in + 2 * (84 * t_vis + 20 * vis_q) + 2 * 4 * e_total + 8 * P
    # the rasteriser's share: setup records and bin lists read once, image + counts written once
    raster = 84 * t_vis + 20 * vis_q + 4 * e_total + 8 * P
    return {"frame": int(frame), "block_raster": int(raster), "t_vis": int(t_vis),
            "q_vis": vis_q, "e_quad": int(qc.sum()), "e_tri": int(tc.sum()), "pixels": P,
            "b_in": int(b_in)}


def cpu_baseline(scene, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    if threads:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    import oracle
    lib = oracle.lib()
    nthreads = lib.oracle_num_threads()
    t0 = time.perf_counter()
    res = oracle.render(scene, want_records=False)
    dt = time.perf_counter() - t0
    frags = res["num_fragments"]
    return {"value": frags / dt, "unit": "fragments/s", "cores": int(nthreads), "kind": "port",
            "sample": f"one full frame of the same workload in the C oracle (exact mode, OpenMP "
                      f"over bins, {nthreads} threads): {dt:.2f} s for {frags} fragments; "
                      "a restatement of the reference's shaders, not the reference itself",
            "ms_per_frame": dt * 1e3}, res


def parity(r, out, ref):
    """The benchmarked GPU frame against the oracle's frame of the same workload: LucidInfo
    counters and the per-bin count/offset arrays, per-bin lists (as multisets) and per-pixel
    fragment counts bit-exact; largest RGBA8 difference (the bar is 1)."""
    import numpy as np
    INFO = 1152
    nb = r.bin_count
    gi, ri = r.read_info(), ref["info"]
    counters_equal = bool(gi[0] == ri[0] and np.array_equal(gi[1:3], ri[1:3]) and
                          np.array_equal(gi[5:10], ri[5:10]) and
                          np.array_equal(gi[32:36], ri[32:36]) and
                          np.array_equal(gi[INFO:INFO + 6 * nb], ri[INFO:INFO + 6 * nb]))
    gq, gt = r.read_bin_lists()

    def per_bin_sorted(lst, counts, offs):
        counts, offs = counts.astype(np.int64), offs.astype(np.int64)
        idx = np.concatenate([np.arange(o, o + c) for o, c in zip(offs, counts)]) \
            if counts.sum() else np.zeros(0, np.int64)
        bins = np.repeat(np.arange(nb), counts)
        vals = lst[idx]
        order = np.lexsort((vals, bins))
        return vals[order]

    c = ri[INFO:]
    lists_equal = True
    for k_cnt, k_off, g_list, r_list in ((0, 1, gq, ref["bin_quads"]), (3, 4, gt, ref["bin_tris"])):
        cnt, off = c[k_cnt * nb:(k_cnt + 1) * nb], c[k_off * nb:(k_off + 1) * nb]
        if int(cnt.sum()) != len(g_list):
            lists_equal = False
            continue
        lists_equal &= bool(np.array_equal(per_bin_sorted(g_list, cnt, off),
                                           per_bin_sorted(r_list, cnt, off)))
    gc = out["frag_counts"].cpu().numpy().astype(np.uint32)
    d = np.abs(out["image"].cpu().numpy().astype(np.int32) - ref["image"].astype(np.int32))
    return {"oracle": "C restatement (oracle/lucid_oracle.c), exact mode",
            "counters_equal": counters_equal, "bin_lists_equal": lists_equal,
            "frag_counts_equal": bool(np.array_equal(gc, ref["frag_counts"])),
            "fragments_equal": int(gi[60]) == int(ref["num_fragments"]),
            "max_rgba8_diff": int(d.max()), "pixels_off_by_one": int((d.max(-1) > 0).sum()),
            "max_fragments_per_pixel": int(ref["frag_counts"].max())}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from lucid_amd import _lib, scenes
    from lucid_amd.renderer import GpuScene, LucidRenderer

    if args.only_backward:
        print(json.dumps({"backward": backward_ratio(args.bwd_steps)}), flush=True)
        return
    print(f"[bench] generating workload ({args.tris} tris)", file=sys.stderr, flush=True)
    scene = scenes.config1(num_tris=args.tris)
    r = LucidRenderer((scene.width, scene.height), opts=_lib.OPT_TIMERS,
                      max_quads=scene.num_quads)
    g = GpuScene(scene)
    out = r.alloc_outputs(rgba8=True, counts=True)
    torch.cuda.synchronize()
    tuning = r.get_tuning()
    if not args.no_autotune:
        print("[bench] autotune", file=sys.stderr, flush=True)
        tuning = r.autotune(g, out, iters=8)
    print("[bench] warmup", file=sys.stderr, flush=True)
    for _ in range(max(1, args.warmup)):
        r.render(g, out)
    info = r.read_info()
    frags = int(info[60])

    # Phase A: back-to-back frames, HIP events around the whole timed region.
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        r.render_async(g, out)
    e1.record(stream)
    torch.cuda.synchronize()
    rc = r.finish()
    if rc != 0:
        _lib.check(rc, "timed region")
    ms = e0.elapsed_time(e1) / args.steps

    # Phase B: the same frames one at a time, per-stage HIP events (library timers).
    stage_runs = []
    frame_ms = []
    for _ in range(args.steps):
        r.render_async(g, out)
        st = r.stage_times()
        stage_runs.append(st)
        frame_ms.append(sum(st.values()))
    _lib.check(r.finish(), "stage timing")
    stages = {k: float(np.mean([s[k] for s in stage_runs])) for k in stage_runs[0]}
    dominant = "block_raster"

    ab = algorithmic_bytes(scene, r, info)
    dom_bytes = ab[dominant]
    dom_s = stages[dominant] * 1e-3
    achieved = dom_bytes / dom_s / 1e9
    frame_gbs = ab["frame"] / (ms * 1e-3) / 1e9

    if args.copy_calibration:
        x = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()

    unsorted = None
    if not args.no_unsorted:
        print("[bench] unsorted baseline", file=sys.stderr, flush=True)
        unsorted = unsorted_baseline(scene, tuning, args.steps, ms)

    bwd = None
    if not args.no_backward:
        print("[bench] fwd+bwd ratio (200k tris, 512x512)", file=sys.stderr, flush=True)
        bwd = backward_ratio(args.bwd_steps)

    batch = None
    if not args.no_batch:
        print(f"[bench] batch ({args.batch_scenes} scenes)", file=sys.stderr, flush=True)
        batch = batch_leg(args.batch_scenes)

    cpu = None
    par = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        cpu, ref = cpu_baseline(scene, args.cpu_threads)
        # the last frame rendered into `out` is the benchmarked frame (same inputs, tuning)
        print("[bench] parity of the benchmarked frame against the oracle", file=sys.stderr,
              flush=True)
        r.render(g, out)
        par = parity(r, out, ref)
        del ref

    line = {
        "metric": METRIC,
        "value": frags / (ms * 1e-3),
        "unit": "fragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "1M soup triangles at 1920x1080, 64 Gaussian blobs (sigma 0.1 NDC, "
                        "log-normal populations), log-normal edges (median 6 px), depth "
                        "U[1,1000], fp32 colours, exact-OIT forward frame, seed 1001",
            "triangles": scene.num_quads, "width": scene.width, "height": scene.height,
            "fragments": frags, "mean_fragments_per_pixel": frags / (scene.width * scene.height),
            "median_frame_ms_stage_events": float(np.median(frame_ms)),
            "stages_ms": stages,
            "fallback_pixels": int(info[62]), "oversized_blocks": int(info[63]),
            "tuning": tuning,
            "high_bins": int(info[9]), "low_bins": int(info[7]),
            "backward": bwd,
            "batch": batch,
            "unsorted_baseline": unsorted,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            # HBM bytes of the block-raster stage (both kernels) per frame, measured on this
            # workload with rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by a 256 MiB copy
            # (profiles/c1_forward_kernels.md); not re-measured here (no counters in-process)
            "traffic": RASTER_TRAFFIC_BYTES if args.tris == DEFAULT_TRIS else None,
            # the bound that holds: the stage issues VALU instructions most of its time
            "valu_busy": RASTER_VALU_BUSY if args.tris == DEFAULT_TRIS else None,
            "traffic_source": "rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated",
            "algorithmic_bytes": dom_bytes,
            "frame_algorithmic_bytes": ab["frame"],
            "frame_achieved": frame_gbs,
            "frame_frac": frame_gbs / HBM_PEAK_GBS,
        },
        "cpu_baseline": cpu,
        "parity": par,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
