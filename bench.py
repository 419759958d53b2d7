"""Benchmark of the torchmorph operator path on one MI355X.

Headline (``value``): exact EDT throughput in Gvox/s on the plain EDT workload
-- bool masks B=8, C=1, 256x256, foreground p=0.9 (seeded), unit sampling,
float32 distances + int64 indices written in place -- bit-exact against the
CPU oracle.  One step = one replay of a HIP graph holding one call of
euclidean_distance_transform per buffer set: 48 batches over 48 distinct
buffer sets (528 MB, above the 256 MB Infinity Cache), so every call reads and
writes HBM.  ``value`` counts the voxels of every call of the timed steps.
The per-call time of eager Python calls is reported beside it
(``eager_call``), never as ``value``.

Printed: one JSON line (contract of the build driver), with ``roofline`` for
the dominant kernel (algorithmic bytes per launch / event-timed launch
duration) and ``cpu_baseline`` (the CPU oracle, one core, on a bounded
sample of the same workload).  ``extras`` holds the other workloads, each
timed with HIP events on its own stream and parity-checked:
  morph3d   -- grey dilation (flat cubes 3/7/15, non-flat ball r=3 with
               heights 0..3) and binary erosion x4 on 4x2x128^3;
  ragged    -- 512 2-D masks (sides log-uniform 32..1024) + 64 3-D volumes
               (48..256), smoothed-noise masks, anisotropic sampling; EDT,
               taxicab and chessboard chamfer in one persistent launch;
  grad      -- grey dilation backward on 16x1x512x512, and point-cloud
               Sinkhorn B=256, n=m=1024, d=3, eps=0.05*median, to 1e-5;
  warmstart -- held-out ragged pairs (256..4096 points, eps 0.1/0.01 x
               median) solved to 1e-5 cold and warm-started.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
C1 = dict(B=8, C=1, H=256, W=256, p=0.9)
BYTES_PER_VOX = 1 + 4 + 16  # bool mask in, f32 distance + 2 x int64 index out
HEADLINE_KEY = "edt2d 256 256 64 1"
# HBM bytes per call of each profiled headline launch variant: FETCH_SIZE x 2.00 +
# WRITE_SIZE x 1.00 (factors of a plain 256 MB copy), summed over the call's kernels
# (profiles/edt2d_b8_256x256_hbm_counters.md)
TRAFFIC_PER_CALL = {0: 13_260_000, 8: 13_397_504, 16: 11_724_954}
HEADLINE_KERNELS = {v: "edt2d_cols_kernel + edt2d_rows_kernel (one call)" for v in range(8)}
HEADLINE_KERNELS.update({v: "edt2d_cols_kernel + edt2d_rows4_kernel (one call)"
                         for v in (8, 9, 10, 11, 12, 13, 14, 15)})
HEADLINE_KERNELS.update({v: "edt2d_colsum_kernel + edt2d_rowsum_kernel (one call)"
                         for v in (16, 17)})


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="graph replays timed")
    ap.add_argument("--warmup", type=int, default=5, help="graph replays untimed")
    ap.add_argument("--sets", type=int, default=48, help="distinct buffer sets (>=256 MB total)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-extras", action="store_true", help="headline workload only")
    ap.add_argument("--only", choices=["morph3d", "ragged", "grad", "hist", "warmstart"],
                    default=None,
                    help="run one extra workload alone and print its JSON (for profiling)")
    ap.add_argument("--tuning", default=os.path.join(ROOT, "torchmorph_amd", "tuning", "mi355x.json"),
                    help="autotune table to preload ('' = tune live)")
    ap.add_argument("--no-graph", action="store_true",
                    help="eager launches instead of graph replays (for counter collection)")
    ap.add_argument("--streams", type=int, default=4,
                    help="streams the headline graph spreads its independent calls over "
                         "(MI355X has 4 hardware queues per process; 1 = one chain)")
    ap.add_argument("--counters", choices=["morph3d", "rows"], default=None,
                    help="only run the named workload's kernels a few times, untimed "
                         "(the program the hardware-counter passes profile)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="only run a plain device-to-device copy of MB megabytes 8 times "
                         "(reference for the HBM byte counters) and exit")
    return ap.parse_args()


def _oracle_edt_worker(args):
    """Pool worker: EDT of the given batches for `seconds`; returns voxels done."""
    masks_np, seconds = args
    sys.path.insert(0, ROOT)
    import oracle

    t0 = time.perf_counter()
    n = vox = 0
    while time.perf_counter() - t0 < seconds:
        oracle.edt(masks_np[n % len(masks_np)])
        vox += masks_np[0].size
        n += 1
    return vox, time.perf_counter() - t0


def cpu_baseline(masks_np, seconds):
    """The C oracle on one core, and on a pool over the host cores this job
    may use (16 at most); the pooled rate is `value`."""
    import multiprocessing as mp

    import oracle

    oracle.build()
    vox1, dt1 = _oracle_edt_worker((masks_np, seconds / 2))
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    ctx = mp.get_context("spawn")  # no fork of a process that holds the GPU
    with ctx.Pool(cores) as pool:
        t0 = time.perf_counter()
        res = pool.map(_oracle_edt_worker, [(masks_np, seconds / 2)] * cores)
        wall = time.perf_counter() - t0
        pool.close()
        pool.join()  # workers gone before the run goes on
    vox = sum(v for v, _ in res)
    return dict(value=vox / wall / 1e6, unit="Mvox/s", cores=cores, kind="port",
                single_core_value=vox1 / dt1 / 1e6,
                sample=f"batches of 8x1x256x256 through oracle/tm_oracle.c: {cores} processes "
                       f"x {seconds / 2:.0f} s (pool start-up included), and one core alone")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import numpy as np
    import torch

    import torchmorph_amd as tm

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.tuning:
        # launch variants chosen by an offline sweep of every variant on MI355X
        # (the live tuner's timings of near-equal variants are noisy)
        from torchmorph_amd import autotune

        autotune.load_json(args.tuning)
    if args.calibrate_copy:
        nbytes = args.calibrate_copy * 1_000_000
        src = torch.empty(nbytes, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        for _ in range(8):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 8}), flush=True)
        return
    if args.counters == "morph3d":
        print(json.dumps({"counters_morph3d": counters_morph3d()}), flush=True)
        return
    if args.counters == "rows":
        print(json.dumps({"counters_rows": counters_rows()}), flush=True)
        return
    if args.only:
        fn = {"morph3d": extra_morph3d, "ragged": extra_ragged, "grad": extra_grad_and_sinkhorn,
              "hist": extra_hist, "warmstart": extra_warmstart}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return
    B, C, H, W, p = C1["B"], C1["C"], C1["H"], C1["W"], C1["p"]
    nvox = B * C * H * W
    S = args.sets
    masks, dists, idxs = [], [], []
    for s in range(S):
        g = torch.Generator().manual_seed(1000 + s)
        masks.append((torch.rand((B, C, H, W), generator=g) < p).to(dev))
        dists.append(torch.empty((B, C, H, W), dtype=torch.float32, device=dev))
        idxs.append(torch.empty((2, B, C, H, W), dtype=torch.int64, device=dev))
    footprint_mb = S * nvox * BYTES_PER_VOX / 1e6

    def step(k):
        s = k % S
        tm.euclidean_distance_transform(masks[s], distances=dists[s], indices=idxs[s])

    # parity of the benchmarked configuration
    step(0)
    torch.cuda.synchronize()
    import oracle

    od, oi = oracle.edt(masks[0].cpu().numpy())
    parity = (torch.equal(dists[0].cpu(), torch.from_numpy(od.astype(np.float32)))
              and np.array_equal(idxs[0].cpu().numpy(), oi))

    # eager calls through the Python API (host path included), HIP events
    # around 2 passes over the buffer sets: reported beside the headline, never
    # as `value`
    for k in range(S):
        step(k)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t_e = time.perf_counter()
    e0.record(stream)
    for k in range(2 * S):
        step(k)
    e1.record(stream)
    torch.cuda.synchronize()
    eager_wall_us = (time.perf_counter() - t_e) * 1e6 / (2 * S)
    eager_event_us = e0.elapsed_time(e1) * 1e3 / (2 * S)

    # The per-call reference: a HIP graph of the S calls as one chain on one
    # stream (buffer sets > the 256 MB Infinity Cache), timed with HIP events;
    # the calls do not overlap, so the per-call time is the two kernels'
    # duration and the roofline below is a per-kernel figure.
    def capture(nstreams):
        streams = [torch.cuda.Stream() for _ in range(nstreams)]
        for st_ in streams:
            st_.wait_stream(torch.cuda.current_stream())
        for k in range(S):  # warm-up, and each stream's workspace, outside capture
            with torch.cuda.stream(streams[k * nstreams // S]):
                step(k)
        for st_ in streams:
            torch.cuda.current_stream().wait_stream(st_)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cur = torch.cuda.current_stream()
            for st_ in streams:
                st_.wait_stream(cur)
            # consecutive calls per branch (measured faster than round robin)
            for k in range(S):
                with torch.cuda.stream(streams[k * nstreams // S]):
                    step(k)
            for st_ in streams:
                cur.wait_stream(st_)
        return g, streams

    def replay_ms(g, reps):
        e0_ = torch.cuda.Event(enable_timing=True)
        e1_ = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0_.record(stream)
        for _ in range(reps):
            g.replay()
        e1_.record(stream)
        torch.cuda.synchronize()
        return e0_.elapsed_time(e1_)

    if args.no_graph:  # eager calls only (hardware-counter passes see every dispatch)
        graph1 = None
        per_call_s = eager_event_us / 1e6
    else:
        graph1, _ = capture(1)
        for _ in range(max(1, args.warmup)):
            graph1.replay()
        per_call_s = replay_ms(graph1, max(args.steps, 10)) / 1e3 / (max(args.steps, 10) * S)
    achieved = nvox * BYTES_PER_VOX / per_call_s / 1e9

    # One step = one replay of a HIP graph of the S independent calls spread
    # over `streams` forked streams (parallel graph branches): one call's
    # latency-bound kernels overlap another's, as independent requests do on
    # MI355X's hardware queues.
    NS = max(1, args.streams)
    graph = None
    if not args.no_graph:
        graph, _ = capture(NS) if NS > 1 else (graph1, None)
        for _ in range(max(1, args.warmup)):
            graph.replay()
    else:
        for k in range(S):
            step(k)
    torch.cuda.synchronize()
    # the graph's calls wrote the last buffer set: recheck its parity
    od, oi = oracle.edt(masks[S - 1].cpu().numpy())
    parity = parity and (torch.equal(dists[S - 1].cpu(), torch.from_numpy(od.astype(np.float32)))
                         and np.array_equal(idxs[S - 1].cpu().numpy(), oi))

    K = args.steps
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    if args.no_graph:
        for k in range(K * S):
            step(k)
    else:
        for _ in range(K):
            graph.replay()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    calls = K * S
    value = calls * nvox / wall / 1e9
    step_share_s = gpu_ms / 1e3 / calls  # GPU time per call inside the concurrent step
    aggregate = nvox * BYTES_PER_VOX / step_share_s / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline([m.cpu().numpy() for m in masks[:4]], args.cpu_seconds)
        # SciPy, the library the reference's tests compare against, per image
        cpu["scipy"] = _scipy_rates("edt_indices", [m[b, 0].cpu().numpy() for m in masks[:2]
                                                    for b in range(B)], args.cpu_seconds / 2)

    extras = None
    if not args.no_extras:
        extras = {}
        for name, fn in (("morph3d", extra_morph3d), ("ragged", extra_ragged),
                         ("grad", extra_grad_and_sinkhorn), ("hist", extra_hist),
                         ("warmstart", extra_warmstart)):
            t_x = time.perf_counter()
            extras[name] = fn()
            extras[name]["wall_s"] = round(time.perf_counter() - t_x, 1)
            print(f"# extra {name} done in {extras[name]['wall_s']} s", file=sys.stderr,
                  flush=True)

    from torchmorph_amd import autotune

    tuned = autotune.export()
    out = {
        "metric": "EDT Gvox/s (bit-exact); Sinkhorn solves/s at 1e-5 marginal err, warm vs cold",
        "value": round(value, 3),
        "unit": "Gvox/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / K, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+f64",
        "data": "synthetic",
        "config": {
            "workload": "plain exact EDT: bool masks B=8 C=1 256x256 p_fg=0.9, unit sampling, "
                        "f32 distances + int64 indices in place",
            "batch": B * C, "shape": [H, W], "calls_per_step": S,
            "voxels_per_step": S * nvox, "buffer_sets": S,
            "rotating_footprint_MB": round(footprint_mb, 1),
            "streams": 1 if args.no_graph else NS,
            "timing": (f"eager: {S} Python calls per step" if args.no_graph else
                       f"one step = one replay of a HIP graph of {S} independent calls, one per "
                       f"distinct buffer set, spread over {NS} parallel streams"),
        },
        "parity": "bit-exact vs oracle" if parity else "MISMATCH",
        "eager_call": {"us_per_call_wall": round(eager_wall_us, 2),
                       "us_per_call_events": round(eager_event_us, 2),
                       "note": "euclidean_distance_transform called from Python, host path "
                               "included; not the headline"},
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            # the counter figure belongs to the profiled launch variant
            "traffic": TRAFFIC_PER_CALL.get(tuned.get(HEADLINE_KEY)),
            "kernel": HEADLINE_KERNELS.get(tuned.get(HEADLINE_KEY), "edt2d (two launches)"),
            "bytes_per_launch": nvox * BYTES_PER_VOX,
            "event_us_per_call": round(per_call_s * 1e6, 3),
            "measured_on": f"a HIP graph of the {S} calls as one chain (no overlap): per-call "
                           f"time = the two kernels back to back",
            "concurrent": {"achieved": round(aggregate, 1),
                           "frac": round(aggregate / HBM_PEAK_GBS, 4),
                           "us_per_call": round(step_share_s * 1e6, 3),
                           "note": f"algorithmic bytes over GPU time per call in the headline "
                                   f"step ({NS} streams): an aggregate rate, the calls overlap"},
            "autotune": tuned,
        },
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------------ extras --

def _timed(fn, reps, stream):
    """Mean milliseconds of fn() over reps, HIP events on `stream`."""
    import torch

    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _timed_sets(fn, sets, reps, stream):
    """Mean milliseconds per call of fn(k), k cycling over `sets` distinct
    buffer sets (together larger than the 256 MB Infinity Cache), reps passes,
    HIP events on `stream`."""
    import torch

    for k in range(sets):
        fn(k)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        for k in range(sets):
            fn(k)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / (reps * sets)


def _gb(bytes_, ms):
    return round(bytes_ / ms / 1e6, 1)


def counters_morph3d():
    """Grey dilation size 3 and binary erosion x 4 (default cross) on the
    C2 volumes (4 x 2 x 128^3, 256-level values; Bernoulli(0.7) masks), three
    calls each: the program of the morphology HBM-counter passes."""
    import torch

    import torchmorph_amd as tm

    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.floor(torch.rand((4, 2, 128, 128, 128), generator=g, device="cuda") * 256) / 256
    m = torch.rand((4, 2, 128, 128, 128), generator=g, device="cuda") < 0.7
    y = torch.empty_like(x)
    for _ in range(3):
        tm.grey_dilation(x, size=3, output=y)
        tm.binary_erosion(m, iterations=4)
    torch.cuda.synchronize()
    return {"grey_dilation_size3": 3, "binary_erosion_x4": 3}


def counters_rows():
    """Anisotropic 2-D EDT (sampling 1.3, 0.7, distances only) of one and of
    64 smooth-noise images of 256 x 1024, two eager calls each, autotuner
    off: the program of the row-envelope counter passes (edt_pass0_kernel,
    then edt_rows_kernel<1, 32> over 256 and 16,384 rows of 1,024 voxels)."""
    import torch

    import torchmorph_amd as tm
    from torchmorph_amd import autotune

    autotune.enable(False)
    gen = torch.Generator(device="cuda").manual_seed(0)
    for b in (1, 64):
        m = _smooth_noise_mask((b * 256, 1024), gen).view(b, 1, 256, 1024)
        for _ in range(2):
            tm.euclidean_distance_transform(m, sampling=[1.3, 0.7])
    torch.cuda.synchronize()
    return {"images": [1, 64], "shape": [256, 1024], "calls_each": 2}


def extra_morph3d():
    """4x2x128^3 float32 quantised to 256 levels (134 MB per volume set); each
    workload rotates over several input / output sets (> 256 MB together);
    parity on item 0."""
    import numpy as np
    import torch
    import torch.nn.functional as F

    import oracle
    import torchmorph_amd as tm

    g = torch.Generator(device="cuda").manual_seed(7)
    SG, SB = 3, 8  # grey sets (3 x 268 MB in + out), binary sets (8 x 33.6 MB)
    xs = [torch.floor(torch.rand((4, 2, 128, 128, 128), generator=g, device="cuda") * 256) / 256
          for _ in range(SG)]
    ys = [torch.empty_like(xs[0]) for _ in range(SG)]
    masks = [torch.rand((4, 2, 128, 128, 128), generator=g, device="cuda") < 0.7
             for _ in range(SB)]
    mouts = [torch.empty_like(masks[0]) for _ in range(SB)]
    x, mask = xs[0], masks[0]
    n = x.numel()
    st = torch.cuda.current_stream()
    out = {"rotating_sets": {"grey": SG, "binary": SB,
                             "grey_MB": round(SG * 2 * 4 * n / 1e6, 1),
                             "binary_MB": round(SB * 2 * n / 1e6, 1)}}
    x0 = x[:1, :1].cpu().numpy()
    for size in (3, 7, 15):
        y = tm.grey_dilation(x, size=size)
        ms = _timed_sets(lambda k: tm.grey_dilation(xs[k], size=size, output=ys[k]), SG, 3, st)
        # flat cube, reflect border: the max over the in-image part of the window
        ref = F.max_pool3d(x.view(8, 1, 128, 128, 128), size, 1, size // 2).view_as(x)
        out[f"grey_dilation_cube{size}"] = {"ms": round(ms, 3), "Gvox_s": round(n / ms / 1e6, 2),
                                            "GB_s": _gb(8 * n, ms),
                                            "parity": bool(torch.equal(y, ref))}
    r = np.sqrt(((np.indices((7, 7, 7)) - 3) ** 2).sum(0))
    fp = r <= 3.0
    heights = np.where(fp, np.floor(3.0 - r), 0.0).astype(np.float32)
    s_t = torch.from_numpy(heights)
    f_t = torch.from_numpy(fp)
    y = tm.grey_dilation(x, structure=s_t, footprint=f_t)
    ms = _timed_sets(lambda k: tm.grey_dilation(xs[k], structure=s_t, footprint=f_t,
                                                output=ys[k]), SG, 2, st)
    want = oracle.grey(x0, heights, fp, (7, 7, 7), (0, 0, 0), dilation=True)
    out["grey_dilation_ball_r3"] = {"ms": round(ms, 3), "Gvox_s": round(n / ms / 1e6, 2),
                                    "GB_s": _gb(8 * n, ms),
                                    "parity": bool(np.array_equal(y[:1, :1].cpu().numpy(), want))}
    yb = tm.binary_erosion(mask, iterations=4)
    ms = _timed_sets(lambda k: tm.binary_erosion(masks[k], iterations=4, output=mouts[k]), SB,
                     5, st)
    cross = np.zeros((3, 3, 3), bool)
    cross[1, 1, :] = cross[1, :, 1] = cross[:, 1, 1] = True
    want = oracle.binary(mask[:1, :1].cpu().numpy(), cross, (0, 0, 0), iterations=4)
    out["binary_erosion_x4"] = {"ms": round(ms, 3), "Gvox_s": round(n / ms / 1e6, 2),
                                "GB_s": _gb(2 * n, ms),
                                "parity": bool(np.array_equal(yb[:1, :1].cpu().numpy(),
                                                              want.astype(bool)))}
    # composites, fused epilogue against the plain composition of the same kernels
    composed = {
        "morphological_gradient": lambda v: tm.grey_dilation(v, size=3) - tm.grey_erosion(v, size=3),
        "white_tophat": lambda v: v - tm.grey_opening(v, size=3),
        "morphological_laplace": lambda v: (tm.grey_dilation(v, size=3) + tm.grey_erosion(v, size=3)
                                            - 2 * v),
    }
    for name, plain in composed.items():
        fn = getattr(tm, name)
        y = fn(x, size=3)
        ms = _timed_sets(lambda k: fn(xs[k], size=3, output=ys[k]), SG, 3, st)
        ms_plain = _timed_sets(lambda k: plain(xs[k]), SG, 3, st)
        out[f"{name}_cube3"] = {"ms": round(ms, 3), "composed_ms": round(ms_plain, 3),
                                "parity_with_composed": bool(torch.equal(y, plain(x)))}
    out["bytes_per_vox"] = "grey 4 in + 4 out; binary 1 in + 1 out (algorithmic)"
    # CPU beside it: the C oracle (one 128^3 item, one thread), and SciPy
    # (the library the reference's tests compare against) per item, one core
    # and a pool over the host cores
    t0 = time.perf_counter()
    oracle.grey(x0, None, None, (3, 3, 3), (0, 0, 0), dilation=True)
    t_grey = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.binary(mask[:1, :1].cpu().numpy(), cross, (0, 0, 0), iterations=4)
    t_bin = time.perf_counter() - t0
    out["cpu_baseline"] = {"grey_dilation_cube3_Mvox_s": round(x0.size / t_grey / 1e6, 2),
                           "binary_erosion_x4_Mvox_s": round(x0.size / t_bin / 1e6, 2),
                           "cores": 1, "kind": "port",
                           "sample": "one 128^3 item through oracle/tm_oracle.c"}
    items = [xs[0][b, c].cpu().numpy() for b in range(2) for c in range(2)]
    out["scipy"] = _scipy_rates("grey_dilation_size3", items, 3.0)
    mitems = [masks[0][b, c].cpu().numpy() for b in range(2) for c in range(2)]
    out["scipy"].update(_scipy_rates("binary_erosion_x4", mitems, 3.0))
    return out


def _scipy_job(args):
    """Pool worker: SciPy `op` over the items for `seconds`; (voxels, seconds)."""
    op, items, seconds = args
    from scipy import ndimage as ndi

    fns = {"grey_dilation_size3": lambda v: ndi.grey_dilation(v, size=3),
           "binary_erosion_x4": lambda v: ndi.binary_erosion(v, iterations=4),
           "edt_indices": lambda v: ndi.distance_transform_edt(v, return_indices=True)}
    fn = fns[op]
    t0 = time.perf_counter()
    vox = k = 0
    while time.perf_counter() - t0 < seconds:
        fn(items[k % len(items)])
        vox += items[k % len(items)].size
        k += 1
    return vox, time.perf_counter() - t0


def _scipy_rates(op, items, seconds):
    """SciPy on one core and on a pool over the host cores (16 at most), in
    Mvox/s, or a note when SciPy is not importable."""
    try:
        import scipy
    except ImportError:
        return {f"{op}_scipy": "scipy not importable"}
    import multiprocessing as mp

    v1, t1 = _scipy_job((op, items, seconds / 2))
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    with mp.get_context("spawn").Pool(cores) as pool:
        t0 = time.perf_counter()
        res = pool.map(_scipy_job, [(op, items, seconds / 2)] * cores)
        wall = time.perf_counter() - t0
        pool.close()
        pool.join()
    return {f"{op}_scipy": {"one_core_Mvox_s": round(v1 / t1 / 1e6, 2),
                            "pooled_Mvox_s": round(sum(v for v, _ in res) / wall / 1e6, 2),
                            "cores": cores, "version": scipy.__version__}}


def _smooth_noise_mask(shape, gen, sigma=4.0):
    """N(0,1) noise, Gaussian filter (sigma 4, truncate 4, reflect), > 0."""
    import torch
    import torch.nn.functional as F

    x = torch.randn(shape, generator=gen, device="cuda")
    rad = int(4 * sigma + 0.5)
    w = torch.exp(-0.5 * (torch.arange(-rad, rad + 1, dtype=torch.float64) / sigma) ** 2)
    w = (w / w.sum()).tolist()
    for ax in range(x.dim()):
        L = x.shape[ax]
        y = x.movedim(ax, -1)
        lead = y.shape[:-1]
        y = F.pad(y.reshape(1, -1, L), (rad, rad), mode="reflect" if L > rad else "replicate")
        acc = torch.zeros(y.shape[:-1] + (L,), device="cuda")
        for j, wj in enumerate(w):
            acc += wj * y[..., j:j + L]
        x = acc.reshape(lead + (L,)).movedim(-1, ax)
    return x > 0


def extra_ragged():
    import math

    import numpy as np
    import torch

    import oracle
    import torchmorph_amd as tm
    from torchmorph_amd.ragged import RaggedDistanceTransform

    rng = np.random.default_rng(0)
    gen = torch.Generator(device="cuda").manual_seed(0)
    shapes = [tuple(int(round(math.exp(rng.uniform(math.log(32), math.log(1024)))))
                    for _ in range(2)) for _ in range(512)]
    shapes += [tuple(int(round(math.exp(rng.uniform(math.log(48), math.log(256)))))
                     for _ in range(3)) for _ in range(64)]
    masks = [_smooth_noise_mask(sh, gen) for sh in shapes]
    samp = [list(rng.uniform(0.5, 2.0, size=len(sh))) for sh in shapes]
    nvox = sum(math.prod(sh) for sh in shapes)
    ops = ("edt", "cdt_taxicab", "cdt_chessboard")
    st = torch.cuda.current_stream()
    res = {"items": len(shapes), "voxels": nvox, "ops": list(ops)}
    outs = {}
    for sched in ("per_pass", "persistent", "per_level"):
        plan = RaggedDistanceTransform(shapes, ops, samp, schedule=sched, device="cuda")
        flat = plan.pack(masks)
        plan.run(flat)
        ms = _timed(lambda: plan.run(flat, check=False), 5, st)
        plan.run(flat)  # checks the dependency-wait flag
        res[sched] = {"ms": round(ms, 3), "Gvox_op_s": round(3 * nvox / ms / 1e6, 2),
                      "GB_s": _gb(3 * 5 * nvox, ms)}
        outs[sched] = plan.distances.clone()
        del plan
    res["schedules_agree"] = bool(all(torch.equal(outs["per_pass"], v) for v in outs.values()))
    # the per-pass launches captured once as a HIP graph and replayed; the
    # row-pass chunk depth the autotuner picked for this shape class
    from torchmorph_amd import autotune
    gplan = RaggedDistanceTransform(shapes, ops, samp, schedule="per_pass", graph=True,
                                    device="cuda")
    gflat = gplan.pack(masks)
    gplan.run(gflat)
    ms = _timed(lambda: gplan.run(gflat, check=False), 5, st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        gplan.run(gflat, check=False)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / 5 * 1e3
    res["per_pass_graph"] = {"ms": round(ms, 3), "wall_ms": round(wall, 3),
                             "equal": bool(torch.equal(gplan.distances, outs["per_pass"])),
                             "tuned": {k: v for k, v in autotune.export().items()
                                       if k.startswith("ragged_rows")}}
    del gplan
    plan = RaggedDistanceTransform(shapes, ops, samp, schedule="per_pass", device="cuda")
    plan.run(plan.pack(masks))
    out = plan.results()
    # dense per-item calls for comparison (the reference's only option), with
    # the default launch variants (no tuning runs inside the timed loop)
    from torchmorph_amd import autotune
    autotune.enable(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dense = []
    for k, m in enumerate(masks):
        xk = m[None, None]
        dense.append((tm.euclidean_distance_transform(xk, sampling=samp[k]),
                      tm.chamfer_distance_transform(xk, metric="taxicab"),
                      tm.chamfer_distance_transform(xk, metric="chessboard")))
    torch.cuda.synchronize()
    res["dense_per_item_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    autotune.enable(True)
    same = all(torch.equal(out[o][k], dense[k][o][0, 0]) for k in range(len(masks))
               for o in range(3))
    ok = True
    for k in (0, 1, 2, 512, 513):
        if math.prod(shapes[k]) > 3_000_000:
            continue
        mk = masks[k].cpu().numpy()[None, None]
        od, _ = oracle.edt(mk, samp[k])
        ok &= np.array_equal(out[0][k].cpu().numpy(), od[0, 0].astype(np.float32))
        for o, metric in ((1, "taxicab"), (2, "chessboard")):
            cd, _ = oracle.cdt(mk, metric)
            ok &= np.array_equal(out[o][k].cpu().numpy(), cd[0, 0].astype(np.float32))
    res["parity"] = {"equal_to_dense_per_item": bool(same), "oracle_items_ok": bool(ok)}
    # CPU oracle beside it: the three ops on the first 2-D items until ~3 s
    t0 = time.perf_counter()
    done_vox = 0
    k = 0
    while time.perf_counter() - t0 < 3.0 and k < 512:
        mk = masks[k].cpu().numpy()[None, None]
        oracle.edt(mk, samp[k])
        oracle.cdt(mk, "taxicab")
        oracle.cdt(mk, "chessboard")
        done_vox += 3 * mk.size
        k += 1
    dt = time.perf_counter() - t0
    res["cpu_baseline"] = {"value": round(done_vox / dt / 1e6, 2), "unit": "Mvox-op/s", "cores": 1,
                           "kind": "port", "sample": f"{k} 2-D items x 3 ops, oracle/tm_oracle.c"}
    res["bytes_per_vox_op"] = "1 mask in + 4 f32 out (algorithmic)"
    # ragged morphology on the same item shapes (one launch per list and step)
    from torchmorph_amd.ragged import RaggedBinaryMorphology, RaggedGreyMorphology
    morph = {}
    for nd in (2, 3):
        idx = [k for k, sh in enumerate(shapes) if len(sh) == nd]
        sh_nd = [shapes[k] for k in idx]
        vox = sum(math.prod(sh) for sh in sh_nd)
        vals = [torch.floor(torch.rand(sh, generator=gen, device="cuda") * 256) / 256
                for sh in sh_nd]
        gplan = RaggedGreyMorphology(sh_nd, "dilation", size=3, device="cuda")
        gflat = gplan.pack(vals)
        gplan.run(gflat)
        ms = _timed(lambda: gplan.run(gflat), 5, st)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dense = [tm.grey_dilation(v[None, None], size=3) for v in vals]
        torch.cuda.synchronize()
        dense_ms = (time.perf_counter() - t0) * 1e3
        same = all(torch.equal(a, b[0, 0]) for a, b in zip(gplan.results(), dense))
        bplan = RaggedBinaryMorphology(sh_nd, "erosion", iterations=4, device="cuda")
        bmasks = [masks[k] for k in idx]
        bflat = bplan.pack(bmasks)
        bres = bplan(bflat)
        bms = _timed(lambda: bplan.run(bflat), 5, st)
        bsame = all(torch.equal(bres[j], tm.binary_erosion(bmasks[j][None, None],
                                                          iterations=4)[0, 0])
                    for j in range(0, len(idx), 7))
        morph[f"{nd}d"] = {"items": len(idx), "voxels": vox,
                           "grey_dilation_size3": {"ms": round(ms, 3),
                                                   "GB_s": _gb(8 * vox, ms),
                                                   "dense_per_item_ms": round(dense_ms, 2),
                                                   "equal_to_dense": bool(same)},
                           "binary_erosion_x4": {"ms": round(bms, 3),
                                                 "equal_to_dense": bool(bsame)}}
    res["morphology"] = morph
    return res


def extra_grad_and_sinkhorn():
    import numpy as np
    import torch

    import torchmorph_amd as tm
    from torchmorph_amd.warmstart import gmm_clouds

    st = torch.cuda.current_stream()
    out = {}
    g = torch.Generator(device="cuda").manual_seed(3)
    SETS = 6  # 6 x (x, grad_out, grad_in) of 16.8 MB each = 302 MB
    xs = [torch.rand((16, 1, 512, 512), generator=g, device="cuda").requires_grad_(True)
          for _ in range(SETS)]
    ys = [tm.grey_dilation(x, size=3, differentiable=True) for x in xs]
    gos = [torch.rand_like(ys[0]) for _ in range(SETS)]
    n = xs[0].numel()
    ms = _timed_sets(lambda k: torch.autograd.grad(ys[k], xs[k], gos[k], retain_graph=True),
                     SETS, 5, st)
    out["grey_dilation_backward"] = {"ms": round(ms, 3), "Gvox_s": round(n / ms / 1e6, 2),
                                     "GB_s": _gb(12 * n, ms), "rotating_sets": SETS,
                                     "bytes_per_vox": "x 4 + grad_out 4 + grad_in 4"}

    rng = np.random.default_rng(11)
    B, N = 256, 1024
    xs = torch.from_numpy(np.stack(gmm_clouds(rng, [N] * B))).cuda()
    ys = torch.from_numpy(np.stack(gmm_clouds(rng, [N] * B))).cuda()
    eps = torch.stack([0.05 * tm.optimal_transport.median_sqeuclid(xs[p], ys[p])
                       for p in range(B)]).cpu()
    tol = 8e-6  # float32 error estimate; FP64 re-measurement reported below
    res = tm.sinkhorn_points(xs, ys, eps, tol=tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        res = tm.sinkhorn_points(xs, ys, eps, tol=tol)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    # FP64 re-measurement of the row/column marginal errors on 8 pairs
    errs = []
    f = res.f.view(B, N).double()
    gg = res.g.view(B, N).double()
    for p in range(0, B, 32):
        xp, yp = xs[p].double(), ys[p].double()
        C = ((xp[:, None] - yp[None]) ** 2).sum(-1)
        P = torch.exp((f[p][:, None] + gg[p][None] - C) / float(eps[p])) / (N * N)
        errs.append(max(float((P.sum(1) - 1.0 / N).abs().sum()),
                        float((P.sum(0) - 1.0 / N).abs().sum())))
    # CPU beside it: the FP64 numpy oracle solving two of the pairs to the same tolerance
    import oracle

    t0 = time.perf_counter()
    for p in (0, 1):
        xp = xs[p].double().cpu().numpy()
        yp = ys[p].double().cpu().numpy()
        oracle.sinkhorn_points(xp, yp, np.full(N, 1.0 / N), np.full(N, 1.0 / N), float(eps[p]),
                               tol=1e-5)
    cpu_rate = 2 / (time.perf_counter() - t0)
    out["cpu_baseline"] = {"value": round(cpu_rate, 3), "unit": "solves/s", "cores": 1,
                           "kind": "port",
                           "sample": "2 pairs, FP64 numpy log-domain oracle (elementwise numpy, "
                                     "one thread)"}
    it = res.iterations.float()
    out["sinkhorn_points"] = {"pairs": B, "n": N, "m": N, "solves_s": round(B / dt, 1),
                              "ms": round(dt * 1e3, 2), "iters_mean": round(float(it.mean()), 1),
                              "iters_max": int(it.max()), "tol_f32": tol,
                              "fp64_marginal_err_max_of_8": float(max(errs)),
                              "all_converged": bool((res.err <= tol).all())}
    return out


def extra_hist():
    """Histogram Sinkhorn (the reference's SinkhornSolver) on grid costs
    d = 32x32 = 1024 and 64x64 = 4096, batches of 1 and 64 histograms, 100
    iterations, both forms; per iteration the two update kernels read the
    (d, d) matrix once each (8 d^2 bytes in float32), the bytes counted here.
    At d = 4096 the 64 MB matrix stays in the 256 MB Infinity Cache between
    half-iterations, so the rate can exceed the HBM figure.  Parity: float32
    GPU values against the CPU path at d = 1024 (rtol / atol 1e-4)."""
    import torch

    import torchmorph_amd as tm

    st = torch.cuda.current_stream()
    out = {}
    g = torch.Generator().manual_seed(42)
    for side in (32, 64):
        d = side * side
        cost = tm.build_cost_matrix((side, side), device="cuda")
        for nb in (1, 64):
            src = torch.rand((nb, d), generator=g).cuda()
            tgt = torch.rand((nb, d), generator=g).cuda()
            for log_space in (False, True):
                eps = 4.0 if log_space else 16.0
                solver = tm.SinkhornSolver(epsilon=eps, max_iter=100, log_space=log_space)
                ms = _timed(lambda: solver(src, tgt, cost), 3, st)
                us_it = ms * 1e3 / 100
                key = f"d{d}_n{nb}_{'log' if log_space else 'scaling'}"
                bytes_it = 8 * d * d + 16 * nb * d
                rec = {"ms_per_solve": round(ms, 3), "us_per_iteration": round(us_it, 2),
                       "GB_s": round(bytes_it / us_it / 1e3, 1),
                       "frac_of_hbm": round(bytes_it / us_it / 1e3 / HBM_PEAK_GBS, 4)}
                if side == 32 and nb == 1:
                    want = solver(src.cpu(), tgt.cpu(), cost.cpu())
                    got = solver(src, tgt, cost).cpu()
                    rec["parity_vs_cpu"] = bool(torch.allclose(got, want, rtol=1e-4, atol=1e-4))
                out[key] = rec
    out["bytes_per_iteration"] = "8 d^2 (matrix read by both updates) + 16 n d (scalings)"
    return out


def extra_warmstart():
    import os as _os

    import numpy as np
    import torch

    from torchmorph_amd import train_warmstart
    from torchmorph_amd.warmstart import WarmStartNet

    path = _os.path.join(ROOT, "torchmorph_amd", "weights", "warmstart_gmm3d.pt")
    if not _os.path.exists(path):
        return {"skipped": "no trained weights"}
    net = WarmStartNet.load(path)
    ev = train_warmstart.evaluate(net, count=256, seed=4242, tol=1e-5, chunk=256,
                                  log=lambda s: None)
    keys = ("pairs", "cold_iters", "warm_iters", "iter_ratio", "cold_solves_per_s",
            "warm_solves_per_s", "speedup", "cold_err_max", "warm_err_max", "warm_fallbacks")
    return {k: (float(f"{ev[k]:.5g}") if isinstance(ev[k], float) else ev[k]) for k in keys}


if __name__ == "__main__":
    main()
