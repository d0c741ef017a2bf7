"""Benchmark: env-frames/s of the batch renderer on the 512-environment
apartment workload (128x128 RGB-D + semantic per env, ~200k triangles per
scene, 20 shared textured templates, trilinear sampling).

A step = one draw() of all environments after a per-step camera update (a
seeded random walk), with every input already resident in HBM; outputs stay
in HBM (colour RGBA8, linear depth FP32, semantic uint32 planes).  The walk is
run once per culling mode (plain = every triangle set up, frustum, hiz =
two-phase depth-pyramid occlusion with the previous frame as proposal,
learned = the same with the visibility MLP's proposal together with the
previous frame's visible chunks, the proposal for a camera walk, where the
MLP scores the chunks that come into view, and learned_mlp = the MLP's
proposal alone; the MLP is trained first on other walks); each mode's launch parameters come from the library's autotuner
(gbr_autotune) on the walk's first view; the headline value is the learned
mode, and the final frame of every mode is compared pixel by pixel with the
plain one.  The learned-proposal workload (teleports, no history) times the
MLP's proposal alone against Hi-Z.

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
    """Per step and env a view matrix (steps, n, 4, 4): from the env's pose a
    seeded random walk of 0.25 m forward / 10 degrees left / 10 degrees right
    steps, clamped to the apartment."""
    import numpy as np
    from habitat_sim_amd import visibility

    return visibility.camera_walk(spec, steps, seed)


def host_cpu():
    """The host's CPU model, logical CPU count and the CPU share this
    process may use (its affinity mask, capped by OMP_NUM_THREADS when the
    host sets it: a GPU box grants each GPU a fixed share of its cores)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    share = affinity
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity, "share": share}


def llvmpipe_probe():
    """Whether the reference's CPU path (its batch renderer on Mesa's
    llvmpipe through EGL) could run on this host: an EGL library, Mesa's
    EGL vendor library and the habitat_sim module, looked up by name only
    (nothing is loaded)."""
    import ctypes.util
    import glob
    import importlib.util
    dirs = ["/usr/lib/x86_64-linux-gnu", "/usr/lib64", "/usr/lib", "/usr/local/lib"]
    mesa = sorted({os.path.basename(p) for d in dirs for p in glob.glob(os.path.join(d, "libEGL_mesa.so*"))})
    egl = ctypes.util.find_library("EGL")
    swrast = sorted({p for d in dirs for p in glob.glob(os.path.join(d, "dri", "swrast_dri.so"))})
    found = {"libEGL": egl, "libEGL_mesa": mesa or None, "swrast_dri": swrast or None,
             "habitat_sim": importlib.util.find_spec("habitat_sim") is not None}
    found["usable"] = bool(egl and mesa and found["habitat_sim"])
    found["used"] = "none: " + ("the reference's renderer on llvmpipe would be timed here"
                                if found["usable"] else
                                "no EGL + llvmpipe + habitat_sim on this host, so the CPU "
                                "baseline is the oracle (the project's C restatement of the "
                                "GL pipeline)")
    return found


def cpu_baseline(spec, seconds=12.0, max_seconds=30.0):
    """Oracle env-frames/s on a sample of the workload on this host, on every
    CPU of this process's share (OpenMP over environments), with the same
    planes the GPU writes (colour, window and linear depth, semantic)."""
    import numpy as np
    from habitat_sim_amd import math3d as m3
    from oracle import oracle as O

    host = host_cpu()
    threads = host["share"]
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
                      f"tris each), {frames // n} passes, {el:.1f} s, all planes",
            "host": host, "llvmpipe_probe": llvmpipe_probe()}


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


# (name, gbr cull mode, proposal_history)
MODES = (("plain", 0, 0), ("frustum", 1, 0), ("hiz", 2, 0), ("learned", 3, 1), ("learned_mlp", 3, 0))
HEADLINE = "learned"


def frame_bytes(spec, st, mode, sensors=3):
    """Algorithmic bytes of one frame by the frozen formula

      B_frame = sum_env [ sum_drawn (12 T + 16 V + 64) + sum_sensor W H 4 + 8 W H ]

    drawn = the work items (chunks) actually set up, T / V their triangles
    and vertex-range entries; `sensors` planes of 4 B per pixel; 8 B of
    texture footprint per pixel.  Culled modes add 24 B of chunk bounds per
    chunk tested; the two-phase modes the depth pyramid, 4/3 W H 4 B written
    and as much read per env; the learned mode 64 B of features per scored
    chunk and the model's bf16 weights."""
    w, h = spec.tile
    envs = len(spec.envs)
    items = st["items_phase1"] + st["items_phase2"]
    b = 12 * st["triangles_setup"] + 16 * st["vertices_setup"] + 64 * items
    b += envs * (sensors * w * h * 4 + 8 * w * h)
    if mode != "plain":
        b += 24 * st["work_items"]
    if mode == "hiz" or mode.startswith("learned"):
        b += envs * 2 * (4 * w * h * 4 // 3)
    if mode.startswith("learned"):
        from habitat_sim_amd import _native as N
        b += 64 * st["items_scored"] + 2 * N.VIS_PARAMS
    return int(b)


def build_sha():
    """Hash of the sources libgbr.so is built from: a hardware-counter
    profile counts only for the build it was taken on."""
    import hashlib
    hsh = hashlib.sha256()
    csrc = os.path.join(ROOT, "habitat-sim_amd", "csrc")
    for name in sorted(os.listdir(csrc)):
        with open(os.path.join(csrc, name), "rb") as f:
            hsh.update(name.encode() + b"\0" + f.read())
    with open(os.path.join(ROOT, "include", "gbr.h"), "rb") as f:
        hsh.update(f.read())
    return hsh.hexdigest()[:16]


def profile_of_this_build(name):
    """A counter summary under profiles/ if it was taken on this build."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, "missing"
    with open(path) as f:
        prof = json.load(f)
    if prof.get("build_sha") != build_sha():
        return None, f"profiles/{name} was taken on another build ({prof.get('build_sha')})"
    return prof, f"profiles/{name}"


def run_mode(r, mode, spec, views, proj, warmup, steps, history=0):
    """Times `steps` camera-update + draw() steps after `warmup` untimed
    ones, with stage timing off (its events sit between kernels); then
    replays the same views with stage timing on for the per-stage split,
    and once more one frame at a time for the p50/p99 frame latency."""
    import numpy as np
    from habitat_sim_amd import hip
    r.set_tuning(cull_mode=mode, proposal_history=history, timing=0)
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
                                                     "lane_pixels", "queue_grain", "raster_mode",
                                                     "heavy_entries", "proposal_threshold")},
            "autotune": {"candidates": tuned["candidates"], "default_ms": tuned["initial_ms"],
                         "tuned_ms": tuned["best_ms"]}}


def rollout(args):
    """Trajectory rollout: `--rollout-envs` envs of the apartment workload
    with 128x128 linear depth + semantic sensors, `--rollout-steps` steps of
    the seeded random walk (0.25 m forward / 10 degrees left / right, clamped
    to the apartment), learned proposal together with the previous frame's
    visible chunks; visibility rows are collected every frame and the
    proposal MLP is fine-tuned on them every 50 steps.  Returns env-frames/s
    over the whole rollout (row collection and fine-tuning included, the
    recall readback excluded), p50/p99 frame latency (camera update to the
    frame's completion), the proposal's recall (visible work items that
    phase 1 set up) and the pixels in which the last frame differs from a
    plain render of the same views."""
    import numpy as np
    from habitat_sim_amd import RendererConfiguration, RendererStandalone, scenes, visibility
    from habitat_sim_amd import _native as N

    t_setup = time.perf_counter()
    spec = scenes.apartments(n_envs=args.rollout_envs, tile=128)
    gx, gy = spec.grid
    cfg = RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy))
    r = RendererStandalone(cfg, planes=N.PLANE_LINEAR_DEPTH | N.PLANE_SEMANTIC, top_down=True)
    scenes.populate(r, spec)
    steps = args.rollout_steps
    views = camera_walk(spec, steps, seed=11)
    proj = np.stack([env.projection for env in spec.envs]).astype(np.float32)
    setup_s = time.perf_counter() - t_setup
    print(f"bench: rollout of {len(spec.envs)} envs set up in {setup_s:.1f} s",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    initial = visibility.train(r, spec, frames=args.train_frames, steps=args.train_steps, seed=3)
    train_s = time.perf_counter() - t0
    r.set_tuning(cull_mode=N.CULL_LEARNED, proposal_history=1)
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
    wall = time.perf_counter() - t0 - excluded
    print(f"bench: rollout {steps} steps in {wall:.1f} s, recall {np.mean(recall):.3f}",
          file=sys.stderr, flush=True)
    last = {"linear_depth": r.linear_depth_image(), "semantic": r.semantic_image()}
    r.set_tuning(cull_mode=N.CULL_NONE)
    r.draw()
    plain = {"linear_depth": r.linear_depth_image(), "semantic": r.semantic_image()}
    n = len(spec.envs)
    out = {
        "workload": f"apartments-{n}env-rollout-{steps}steps-128x128-depth-semantic",
        "value": n * steps / wall, "unit": "env-frames/s", "envs": n, "steps": steps,
        "triangles_per_frame": int(sum(spec.triangles(e) for e in range(n))),
        "sensors": ["linear_depth_f32", "semantic_u32"], "cull": "learned+previous-frame",
        "finetune_every": 50, "setup_s": setup_s,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "proposal_recall": {"mean": float(np.mean(recall)), "min": float(np.min(recall)),
                            "first": recall[0], "last": recall[-1]},
        "initial_training": {"seconds": train_s, "poses": args.train_frames * n,
                             "loss_last": initial["loss_last"]},
        "fine_tune_loss_last": [f["loss_last"] for f in fits],
        "diff_pixels_vs_plain_last_frame": diff_pixels(last, plain),
    }
    r.close()
    return out


def learned_proposal(args):
    """Learned-proposal workload: the uneven batch (`--lp-envs` envs of
    10k-1M triangles log-uniform, 64/128/256 px viewports, 1-3 sensors) cut
    into work items of at most 1024 triangles; the visibility model (32 ->
    256 -> 256 -> 1, camera-pose and chunk-box features) trained on
    `--lp-train-frames` x envs uniform random poses, then per culling mode
    `--lp-frames` frames of fresh uniform random poses (every frame a
    teleport, so the previous frame says nothing).  Reports ms per frame,
    chunks set up and recall per mode, learned vs hiz (the two-phase method
    without the model), and the pixels in which each mode's last frame
    differs from the plain render."""
    import numpy as np
    from habitat_sim_amd import RendererConfiguration, RendererStandalone, hip, scenes, visibility
    from habitat_sim_amd import _native as N

    t_setup = time.perf_counter()
    spec = scenes.heterogeneous(n_envs=args.lp_envs)
    gx, gy = spec.grid
    r = RendererStandalone(RendererConfiguration().set_tile_size_count(spec.tile, (gx, gy)))
    scenes.populate(r, spec)
    n = len(spec.envs)
    proj = np.stack([env.projection for env in spec.envs]).astype(np.float32)
    rng = np.random.default_rng(2024)
    frames = [visibility.inverted_rigid(visibility.random_poses(rng, n))
              for _ in range(args.lp_frames + 2)]
    setup_s = time.perf_counter() - t_setup
    t0 = time.perf_counter()
    fit = visibility.train(r, spec, frames=args.lp_train_frames, steps=args.lp_train_steps,
                           seed=3, walk_length=1)
    train_s = time.perf_counter() - t0
    print(f"bench: learned proposal: {n} uneven envs set up in {setup_s:.1f} s, model trained on "
          f"{args.lp_train_frames * n} random poses in {train_s:.1f} s", file=sys.stderr, flush=True)
    stream = r.stream()
    res, planes = {}, {}
    for name, mode, _ in MODES[:4]:
        r.set_tuning(cull_mode=mode, proposal_history=0, timing=0)
        r.update_cameras(0, proj, frames[0])
        r.autotune(warmup=1, frames=2)
        for k in range(2):
            r.update_cameras(0, proj, frames[k])
            r.draw()
        r.synchronize()
        start, stop = hip.Event(), hip.Event()
        start.record(stream)
        t1 = time.perf_counter()
        for k in range(args.lp_frames):
            r.update_cameras(0, proj, frames[2 + k])
            r.draw()
        stop.record(stream)
        r.synchronize()
        wall = time.perf_counter() - t1
        st = r.frame_stats()
        vis = r.frame_visibility()
        planes[name] = {k: getattr(r, f"{k}_image")() for k in ("color", "depth", "semantic")}
        res[name] = {"gpu_ms_per_frame": start.elapsed_ms(stop) / args.lp_frames,
                     "env_frames_per_s": n * args.lp_frames / wall,
                     "items_phase1": st["items_phase1"], "items_phase2": st["items_phase2"],
                     "triangles_setup": st["triangles_setup"], "proposal_recall": vis["recall"],
                     "diff_pixels_vs_plain": diff_pixels_sets(planes[name], planes["plain"], spec)}
        print(f"bench: learned proposal workload, {name}: {res[name]['gpu_ms_per_frame']:.3f} ms",
              file=sys.stderr, flush=True)
    r.close()
    return {"workload": f"uneven-{n}env-teleport-64to256px-1to3sensors-1024tri-chunks",
            "envs": n, "frames": args.lp_frames,
            "triangles_per_frame": int(sum(spec.triangles(e) for e in range(n))),
            "work_items": int(st["work_items"]),
            "training": {"poses": args.lp_train_frames * n, "seconds": train_s,
                         "steps": args.lp_train_steps, "loss_last": fit["loss_last"]},
            "modes": res,
            "learned_vs_hiz": res["hiz"]["gpu_ms_per_frame"] / res["learned"]["gpu_ms_per_frame"],
            "learned_vs_plain": res["plain"]["gpu_ms_per_frame"] / res["learned"]["gpu_ms_per_frame"]}


def diff_pixels_sets(a, b, spec):
    """Pixels that differ between two framebuffer readbacks of an uneven
    batch, counting each scene's planes only where its sensor set has them
    (a plane outside the set holds unspecified values)."""
    import numpy as np
    from habitat_sim_amd import _native as N
    gx, gy = spec.grid
    tw, th = spec.tile
    bits = {"color": N.PLANE_COLOR, "depth": N.PLANE_DEPTH, "semantic": N.PLANE_SEMANTIC}
    sets = np.zeros(gx * gy, np.uint32)
    sets[:len(spec.envs)] = [e.planes or 0xF for e in spec.envs]
    total = 0
    d = np.zeros((gy, gx, th, tw), bool)
    for k, bit in bits.items():
        x = a[k] != b[k]
        if x.ndim == 3:
            x = x.any(-1)
        x = x.reshape(gy, th, gx, tw).transpose(0, 2, 1, 3)
        d |= x & ((sets & bit) != 0).reshape(gy, gx)[:, :, None, None]
    total = int(d.sum())
    return total


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
    ap.add_argument("--modes", default=",".join(m for m, _, _ in MODES),
                    help="culling modes to time, comma-separated (profiling runs time "
                         "fewer; the speedup needs plain and the headline mode)")
    ap.add_argument("--rollout-envs", type=int, default=4096, metavar="ENVS",
                    help="trajectory rollout of ENVS envs with online fine-tuning (0: skip)")
    ap.add_argument("--rollout-steps", type=int, default=500)
    ap.add_argument("--lp-envs", type=int, default=1024,
                    help="learned-proposal workload on the uneven batch of this many envs (0: skip)")
    ap.add_argument("--lp-train-frames", type=int, default=1024,
                    help="frames of uniform random poses the proposal model trains on")
    ap.add_argument("--lp-train-steps", type=int, default=3000)
    ap.add_argument("--lp-frames", type=int, default=10)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (the reference kernel "
                         "for scaling hardware byte counters) and exit")
    args = ap.parse_args()
    modes = [(m, c, h) for m, c, h in MODES if m in args.modes.split(",")]
    if not modes:
        ap.error(f"--modes: none of {[m for m, _, _ in MODES]}")
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "environments would shard across GPUs with no exchange between them", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__
    __graft_entry__.ensure_built()
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
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
    if any(m.startswith("learned") for m, _, _ in modes):
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
    for name, mode, history in modes:
        runs[name] = run_mode(r, mode, spec, views, proj, args.warmup, args.steps, history)
        print(f"bench: {name} {runs[name]['gpu_ms']:.3f} ms/frame", file=sys.stderr, flush=True)
    base = runs.get("plain")
    headline = HEADLINE if HEADLINE in runs else modes[-1][0]
    head = runs[headline]

    cull = {}
    for name, _, _ in modes:
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
    bytes_frame = frame_bytes(spec, head["stats"], headline)
    frame_ms = head["gpu_ms"]
    achieved = bytes_frame / (frame_ms * 1e-3) / 1e9
    prof, prof_source = profile_of_this_build("counters_apartments512.json")
    traffic = prof["frame_bytes"].get(headline) if prof else None
    valu = prof.get("valu_busy", {}).get(headline) if prof else None

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
                   "cull": headline + (" (MLP proposal + previous frame's visible chunks)"
                                       if headline == "learned" else ""),
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": prof_source,
                     "kernel": "frame", "algorithmic_bytes_per_launch": bytes_frame,
                     "launch_ms": frame_ms, "valu_busy": valu,
                     "dominant_stage": {"name": dom, "ms": stages[dom],
                                        "share": stages[dom] / max(sum(stages.values()), 1e-9)},
                     "alu": {"pixels_tested_per_s": head["stats"]["raster_pixels"] / (frame_ms * 1e-3),
                             "fragments_per_s": head["stats"]["fragments"] / (frame_ms * 1e-3),
                             "triangles_setup_per_s":
                                 head["stats"]["triangles_setup"] / (frame_ms * 1e-3)}},
        "cpu_baseline": cpu,
        "batch_steps_per_s": value / n,
        "learned_cull_speedup": cull["learned"]["speedup_vs_plain"] if "learned" in cull else None,
        "learned_vs_hiz": ({"learned_ms": cull["learned"]["gpu_ms_per_frame"],
                            "learned_mlp_ms": (cull["learned_mlp"]["gpu_ms_per_frame"]
                                               if "learned_mlp" in cull else None),
                            "hiz_ms": cull["hiz"]["gpu_ms_per_frame"],
                            "speedup": runs["hiz"]["wall"] / runs["learned"]["wall"],
                            "teleports": "learned_proposal.learned_vs_hiz"}
                           if "learned" in cull and "hiz" in cull else None),
        "vis_training": training,
        "cull": cull,
        "frame_stats": head["stats"],
    }
    if cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    r.close()
    line["learned_proposal"] = learned_proposal(args) if args.lp_envs else None
    line["rollout"] = rollout(args) if args.rollout_envs else None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
