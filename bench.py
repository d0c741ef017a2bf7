"""Benchmark: env-frames/s of the batch renderer on the 512-environment
apartment workload (128x128 RGB-D + semantic per env, ~200k triangles per
scene, 20 shared textured templates, trilinear sampling).

A step = one draw() of all environments after a per-step camera update (a
seeded random walk), with every input already resident in HBM; outputs stay
in HBM (colour RGBA8, linear depth FP32, semantic uint32 planes).  The walk is
run once per culling mode (plain = every triangle set up, frustum, hiz =
two-phase depth-pyramid occlusion with the previous frame as proposal,
learned = the same with the visibility MLP as proposal, trained first on
other walks); each mode's launch parameters come from the library's
autotuner (gbr_autotune) on the walk's first view; the headline value is the
learned mode, and the final frame of every mode is compared pixel by pixel
with the plain one.

Prints ONE JSON line (rank 0).  ``roofline`` is the headline mode's frame
(its kernels back to back on the renderer's stream): the frozen formula

  B_frame = sum_env [ sum_drawn (12 T + 16 V + 64) + sum_sensor W H 4 + 8 W H ]

(drawn = chunks actually set up; culling adds 120 B of work item, bounds and
transform per chunk, the learned proposal 64 B of features per scored chunk
and its bf16 weights, the depth pyramid 8 B read per pixel) over the frame's
mean GPU time from HIP events on that stream around the timed steps; the
per-stage split comes from a separate replay with events around each stage
(they add a few microseconds between kernels, so the timed run has none).
``cpu_baseline`` times the CPU oracle (the project's C restatement of the GL
pipeline, OpenMP over environments) on a bounded sample of the same workload
on this host.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "habitat-sim_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "frames/s @128x128 RGB-D, 512 envs; learned-cull vs plain speedup at 0 px diff"
HBM_PEAK_GBS = 8000.0


def camera_walk(spec, steps, seed=7):
    """Per step and env a view matrix: the env's pose turned by a seeded
    random yaw walk (10 degree steps) and moved 0.25 m forward, clamped to the
    apartment."""
    import numpy as np
    from habitat_sim_amd import math3d as m3

    rng = np.random.default_rng(seed)
    n = len(spec.envs)
    poses = [env.camera_pose.copy() for env in spec.envs]
    views = np.empty((steps, n, 4, 4), np.float32)
    for s in range(steps):
        acts = rng.integers(0, 3, n)
        for e in range(n):
            p = poses[e]
            if acts[e] == 1:
                p = m3.mul(p, m3.rotation_y(math.radians(10)))
            elif acts[e] == 2:
                p = m3.mul(p, m3.rotation_y(math.radians(-10)))
            else:
                fwd = -p[2, :3]
                pos = np.clip(p[3, :3] + 0.25 * fwd, 0.3, 11.7)
                p = p.copy()
                p[3, :3] = pos.astype(np.float32)
                p[3, 1] = 1.5
            poses[e] = p
            views[s, e] = m3.inverted_rigid(p)
    return views


def cpu_baseline(spec, seconds=12.0, max_seconds=30.0):
    """Oracle env-frames/s on a sample of the workload on this host."""
    import numpy as np
    from habitat_sim_amd import math3d as m3
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    sample = list(range(min(64, len(spec.envs))))
    o = O.OracleRenderer(spec.tile, spec.grid)
    o.add_file_data(spec.file)
    for e in sample:
        env = spec.envs[e]
        for inst in env.instances:
            o.add_node_hierarchy(e, inst.template, inst.bake, inst.semantic)
        o.update_camera(e, env.projection, m3.inverted_rigid(env.camera_pose))
    job = O.build_job(o, sample)
    w, h = spec.tile
    n = len(sample)
    out = dict(color=np.empty((n, h, w, 4), np.uint8), depth=np.empty((n, h, w), np.float32),
               lindepth=np.empty((n, h, w), np.float32), semantic=np.empty((n, h, w), np.uint32))
    frames = 0
    t0 = time.perf_counter()
    while True:
        O.run_job(job, out, threads)
        frames += n
        el = time.perf_counter() - t0
        if el >= seconds or el >= max_seconds:
            break
    return {"value": frames / el, "unit": "env-frames/s", "cores": threads, "kind": "port",
            "sample": f"{n} of the {len(spec.envs)} apartment envs (~{spec.triangles(0) // 1000}k "
                      f"tris each), {frames // n} passes, {el:.1f} s, all planes"}


def copy_probe(nbytes: int) -> None:
    """Device-to-device copies of a known size: reads nbytes, writes nbytes."""
    from habitat_sim_amd import hip
    a, b = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
    for _ in range(3):
        hip.memcpy_d2d(b.value, a.value, nbytes)
    hip.hip().hipDeviceSynchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}))
    a.close()
    b.close()


MODES = (("plain", 0), ("frustum", 1), ("hiz", 2), ("learned", 3))
HEADLINE = "learned"


def stage_bytes(stage, spec, st, n_items, n_nodes, learned=False):
    """Algorithmic bytes a stage is responsible for in one frame (the frozen
    frame formula split by stage; intermediate buffers count nothing)."""
    w, h = spec.tile
    envs = len(spec.envs)
    t2, v2 = st["triangles_phase2"], st["vertices_phase2"]
    t1, v1 = st["triangles_setup"] - t2, st["vertices_setup"] - v2
    if stage == "setup1":
        return 12 * t1 + 16 * v1 + 64 * st["items_phase1"]
    if stage == "setup2":
        return 12 * t2 + 16 * v2 + 64 * st["items_phase2"]
    if stage == "cull":  # work item, bounds, transform; learned: feature rows + bf16 weights
        rows = st.get("items_phase1", 0) + st.get("items_candidates", 0)
        return n_items * (32 + 24 + 64) + (64 * rows + 2 * 20865 if learned else 0)
    if stage == "hiz":  # depth pyramid: read the phase-1 depth, write 4/3 of a float per 16 px
        return envs * (w * h * 8 + w * h * 4 // 12)
    if stage == "resolve":
        return envs * (3 * w * h * 4 + 8 * w * h)
    if stage == "nodes":
        return n_nodes * 128
    return 0


def run_mode(r, mode, spec, views, proj, warmup, steps):
    """Times `steps` camera-update + draw() steps after `warmup` untimed
    ones, with stage timing off (its events sit between kernels); then
    replays the same views with stage timing on for the per-stage split,
    and once more one frame at a time for the p50/p99 frame latency."""
    import numpy as np
    from habitat_sim_amd import hip
    r.set_tuning(cull_mode=mode, timing=0)
    stream = r.stream()
    # launch parameters for this mode from the autotuner, on the walk's
    # first views (every candidate renders the same pixels)
    r.update_cameras(0, proj, views[0])
    tuned = r.autotune(warmup=2, frames=3)
    for k in range(warmup):
        r.update_cameras(0, proj, views[k])
        r.draw()
    r.synchronize()
    start, stop = hip.Event(), hip.Event()
    hip.hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for k in range(steps):
        r.update_cameras(0, proj, views[warmup + k])
        r.draw()
    stop.record(stream)
    r.synchronize()
    hip.hip().hipDeviceSynchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_ms(stop) / steps
    stats = r.frame_stats()
    visibility = r.frame_visibility()
    planes = {"color": r.color_image(), "semantic": r.semantic_image(),
              "linear_depth": r.linear_depth_image()}
    # stage split (events between the stages; not part of the timed run)
    r.set_tuning(timing=1)
    for k in range(min(steps, 8)):
        r.update_cameras(0, proj, views[warmup + k])
        r.draw()
    r.synchronize()
    stages = r.frame_timings()
    r.set_tuning(timing=0)
    # frame latency: camera update + draw() to the frame's completion, one
    # frame at a time (not part of the timed run)
    lat = []
    for k in range(min(steps, 30)):
        t1 = time.perf_counter()
        r.update_cameras(0, proj, views[warmup + k])
        r.draw()
        r.synchronize()
        lat.append((time.perf_counter() - t1) * 1e3)
    latency = {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
               "frames": len(lat)}
    return {"wall": wall, "gpu_ms": gpu_ms, "stages": stages, "stats": stats, "planes": planes,
            "latency_ms": latency, "visibility": visibility,
            "tuning": {k: tuned["best"][k] for k in ("bin_size", "bin_capacity_factor",
                                                     "lane_pixels", "queue_grain", "raster_mode")},
            "autotune": {"candidates": tuned["candidates"], "default_ms": tuned["initial_ms"],
                         "tuned_ms": tuned["best_ms"]}}


def rollout(args):
    """Trajectory rollout: `--rollout` envs of the apartment workload with
    128x128 linear depth + semantic sensors, `--rollout-steps` steps of the
    seeded random walk (0.25 m forward / 10 degrees left / right, clamped to
    the apartment), learned proposal with the previous-frame depth pyramid;
    visibility rows are collected every frame and the proposal MLP is
    fine-tuned on them every 50 steps.  Prints one JSON line: env-frames/s
    over the whole rollout (row collection and fine-tuning included, the
    recall readback excluded), p50/p99 frame latency (camera update to the
    frame's completion) and the proposal's recall (visible work items that
    phase 1 set up)."""
    import numpy as np
    from habitat_sim_amd import RendererConfiguration, RendererStandalone, scenes, visibility
    from habitat_sim_amd import _native as N

    t_setup = time.perf_counter()
    spec = scenes.apartments(n_envs=args.rollout, tile=128)
    gx, gy = spec.grid
    cfg = RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy))
    r = RendererStandalone(cfg, planes=N.PLANE_LINEAR_DEPTH | N.PLANE_SEMANTIC, top_down=True)
    scenes.populate(r, spec)
    steps = args.rollout_steps
    views = camera_walk(spec, steps, seed=11)
    proj = np.stack([env.projection for env in spec.envs]).astype(np.float32)
    print(f"bench: rollout of {len(spec.envs)} envs set up in {time.perf_counter() - t_setup:.0f} s",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    initial = visibility.train(r, spec, frames=args.train_frames, steps=args.train_steps, seed=3)
    train_s = time.perf_counter() - t0
    r.set_tuning(cull_mode=N.CULL_LEARNED)
    r.update_cameras(0, proj, views[0])
    r.draw()
    r.synchronize()
    r.vis_clear_rows()
    lat, recall, fits = [], [], []
    excluded = 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        t1 = time.perf_counter()
        r.update_cameras(0, proj, views[k])
        r.vis_collect()
        r.synchronize()
        lat.append((time.perf_counter() - t1) * 1e3)
        t2 = time.perf_counter()
        recall.append(r.frame_visibility()["recall"])
        excluded += time.perf_counter() - t2
        if (k + 1) % 50 == 0:
            fits.append(r.vis_fit(steps=20, batch=4096, learning_rate=1e-3, seed=100 + k))
            r.vis_clear_rows()
            print(f"bench: rollout step {k + 1}, recall {recall[-1]:.3f}, "
                  f"loss {fits[-1]['loss_last']:.3f}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0 - excluded
    n = len(spec.envs)
    out = {
        "metric": "rollout env-frames/s, learned cull with online fine-tuning",
        "value": n * steps / wall, "unit": "env-frames/s", "n_gpus": 1, "steps": steps,
        "higher_is_better": True, "data": "synthetic", "dtype": "f32",
        "config": {"workload": f"apartments-{n}env-rollout-{steps}steps-128x128-depth-semantic",
                   "envs": n, "resolution": [128, 128],
                   "triangles_per_frame": int(sum(spec.triangles(e) for e in range(n))),
                   "sensors": ["linear_depth_f32", "semantic_u32"], "cull": "learned",
                   "finetune_every": 50},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "proposal_recall": {"mean": float(np.mean(recall)), "min": float(np.min(recall)),
                            "first": recall[0], "last": recall[-1]},
        "initial_training": {"seconds": train_s, "loss_last": initial["loss_last"]},
        "fine_tune_loss_last": [f["loss_last"] for f in fits],
    }
    print(json.dumps(out), flush=True)
    r.close()


def diff_pixels(a, b):
    import numpy as np
    d = np.zeros(a["semantic"].shape, bool)
    for k in a:
        x = a[k] != b[k]
        d |= x.any(-1) if x.ndim == 3 else x
    return int(d.sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--envs", type=int, default=512)
    ap.add_argument("--res", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--train-frames", type=int, default=64,
                    help="frames of training rows for the visibility model")
    ap.add_argument("--train-steps", type=int, default=600)
    ap.add_argument("--modes", default=",".join(m for m, _ in MODES),
                    help="culling modes to time, comma-separated (profiling runs time "
                         "fewer; the speedup needs plain and the headline mode)")
    ap.add_argument("--rollout", type=int, default=0, metavar="ENVS",
                    help="instead of the benchmark, a trajectory rollout of ENVS envs with "
                         "online fine-tuning (its own JSON line)")
    ap.add_argument("--rollout-steps", type=int, default=500)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (the reference kernel "
                         "for scaling hardware byte counters) and exit")
    args = ap.parse_args()
    modes = [(m, c) for m, c in MODES if m in args.modes.split(",")]
    if not modes:
        ap.error(f"--modes: none of {[m for m, _ in MODES]}")
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "environments would shard across GPUs with no exchange between them", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__
    __graft_entry__.ensure_built()
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
        return
    if args.rollout:
        rollout(args)
        return
    import numpy as np
    from habitat_sim_amd import RendererConfiguration, RendererStandalone, hip, scenes
    from habitat_sim_amd import _native as N

    spec = scenes.apartments(n_envs=args.envs, tile=args.res)
    gx, gy = spec.grid
    cfg = RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy))
    planes = N.PLANE_COLOR | N.PLANE_LINEAR_DEPTH | N.PLANE_SEMANTIC
    r = RendererStandalone(cfg, planes=planes, top_down=True)
    scenes.populate(r, spec)
    views = camera_walk(spec, args.warmup + args.steps)
    proj = np.stack([env.projection for env in spec.envs]).astype(np.float32)
    n = len(spec.envs)

    # the visibility model trains on its own walks (random starts, other
    # seeds) before the learned mode is timed
    training = None
    if any(m == "learned" for m, _ in modes):
        from habitat_sim_amd import visibility
        t0 = time.perf_counter()
        training = visibility.train(r, spec, frames=args.train_frames, steps=args.train_steps,
                                    seed=3)
        training["seconds"] = time.perf_counter() - t0
        training["frames"] = args.train_frames
        training["poses"] = args.train_frames * len(spec.envs)
        training["steps"] = args.train_steps
        print(f"bench: visibility model trained in {training['seconds']:.1f} s", file=sys.stderr,
              flush=True)
    runs = {}
    for name, mode in modes:
        runs[name] = run_mode(r, mode, spec, views, proj, args.warmup, args.steps)
        print(f"bench: {name} {runs[name]['gpu_ms']:.3f} ms/frame", file=sys.stderr, flush=True)
    base = runs.get("plain")
    headline = HEADLINE if HEADLINE in runs else modes[-1][0]
    head = runs[headline]
    n_items = head["stats"]["work_items"]
    n_nodes = sum(2 + len(env.instances) for env in spec.envs)  # roots + children (upper bound)

    cull = {}
    for name, _ in modes:
        x = runs[name]
        cull[name] = {
            "env_frames_per_s": n * args.steps / x["wall"],
            "gpu_ms_per_frame": x["gpu_ms"],
            "speedup_vs_plain": base["wall"] / x["wall"] if base else None,
            "diff_pixels_vs_plain": diff_pixels(x["planes"], base["planes"]) if base else None,
            "items_phase1": x["stats"]["items_phase1"],
            "items_phase2": x["stats"]["items_phase2"],
            "triangles_setup": x["stats"]["triangles_setup"],
            "bin_entries": x["stats"]["bin_entries"],
            "stage_ms": {k: round(v, 4) for k, v in x["stages"].items() if v > 0},
            "latency_ms": x["latency_ms"],
            "proposal_recall": x["visibility"]["recall"],
            "visible_items": x["visibility"]["visible_items"],
            "tuning": x["tuning"],
            "autotune": x["autotune"],
        }

    # roofline of the headline mode's frame (its kernels back to back on one
    # stream): the frame formula's bytes over the frame's GPU time
    stages = head["stages"]
    dom = max(stages, key=lambda k: stages[k])
    learned = headline == "learned"
    bytes_frame = sum(stage_bytes(k, spec, head["stats"], n_items, n_nodes, learned)
                      for k in stages)
    frame_ms = head["gpu_ms"]
    achieved = bytes_frame / (frame_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_apartments512.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("frame_bytes", {}).get(headline)

    cpu = None if args.no_cpu_baseline else cpu_baseline(spec, args.cpu_seconds)
    value = n * args.steps / head["wall"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "env-frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"apartments-{n}env-{args.res}x{args.res}-rgbd-semantic",
                   "envs": n, "resolution": [args.res, args.res],
                   "triangles_per_frame": int(head["stats"]["triangles"]),
                   "sensors": ["rgba8", "linear_depth_f32", "semantic_u32"],
                   "cull": headline, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "frame", "algorithmic_bytes_per_launch": bytes_frame,
                     "launch_ms": frame_ms,
                     "dominant_stage": {"name": dom, "ms": stages[dom],
                                        "share": stages[dom] / max(sum(stages.values()), 1e-9)},
                     "alu": {"pixels_tested_per_s": head["stats"]["raster_pixels"] / (frame_ms * 1e-3),
                             "fragments_per_s": head["stats"]["fragments"] / (frame_ms * 1e-3),
                             "triangles_setup_per_s":
                                 head["stats"]["triangles_setup"] / (frame_ms * 1e-3)}},
        "cpu_baseline": cpu,
        "batch_steps_per_s": value / n,
        "learned_cull_speedup": cull["learned"]["speedup_vs_plain"] if "learned" in cull else None,
        "vis_training": training,
        "cull": cull,
        "frame_stats": head["stats"],
    }
    if cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))
    r.close()


if __name__ == "__main__":
    main()
