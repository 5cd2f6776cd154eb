#!/usr/bin/env python3
"""Throughput benchmark of the MI355X path tracer.

Workload: the 2,097,152-triangle displaced sphere (GGX conductor) with a
20,480-triangle glass sphere, a diffuse ground and a 1024x512 sky + sun
environment map, rendered at 1920x1080 with max depth 16.  One step is one
wavefront pass of the unidirectional MIS path tracer over the whole frame
(1 sample per pixel, 2,073,600 paths), accumulated progressively; the
default 256 steps are the workload's 256 spp.

Prints one JSON line: Mpaths/s over the timed steps, the roofline of the
dominant kernel (closest-hit traversal) and the CPU oracle's throughput on a
bounded sample of the same workload, timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s + relMSE@equal-time (guided vs plain); p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, RAY_BYTES, HIT_BYTES = 64, 48, 32, 16


def traffic_from_profiles():
    """HBM bytes per trace_closest launch measured with rocprofv3 PMC runs
    (profiles/trace_closest_c1_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "trace_closest_c1_traffic.json")
    try:
        with open(path) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(desc, target_s, threads):
    import oracle

    t0 = time.perf_counter()
    osc = oracle.from_desc(desc)
    build_s = time.perf_counter() - t0
    # probe: one block per thread, then scale the block count to ~target_s
    _, st = osc.render(1, seed=99, max_depth=desc.max_depth, rr_depth=desc.rr_depth, threads=threads,
                       max_blocks=threads, want_image=False)
    wave_s = max(st["seconds"], 1e-3)
    total_blocks = len(oracle.spiral_blocks(desc.camera.width, desc.camera.height))
    blocks = int(min(total_blocks, max(threads, threads * target_s / wave_s)))
    _, st = osc.render(1, seed=7, max_depth=desc.max_depth, rr_depth=desc.rr_depth, threads=threads,
                       max_blocks=blocks, want_image=False)
    return {
        "value": st["paths"] / st["seconds"] / 1e6,
        "unit": "Mpaths/s",
        "cores": st["threads"],
        "kind": "port",
        "sample": "first %d of %d 32x32 blocks of the spiral order (frame centre), 1 spp, %d paths, %.1f s; "
                  "reference octree (robust slabs) built in %.1f s, not timed" % (
                      blocks, total_blocks, int(st["paths"]), st["seconds"], build_s),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=256)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X, so --gpus %d is not run "
              "(the frame would shard by sample index across GPUs)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import drawlab_amd
    from drawlab_amd import scenes

    desc = scenes.displaced_sphere()
    W, H = desc.camera.width, desc.camera.height
    be = drawlab_amd.Backend(0)
    be.upload_scene(desc)
    accel = be.build_accel(drawlab_amd.BUILD_GPU)

    # Untimed pass with in-kernel counters: traversal work per closest-hit ray.
    _, cst = be.render(spp=1, seed=1 << 20, count_traversal=True, want_image=False)
    nodes_per_ray = cst["nodes_closest"] / cst["closest_rays"]
    tris_per_ray = cst["tris_closest"] / cst["closest_rays"]
    bytes_per_ray = RAY_BYTES + HIT_BYTES + NODE_BYTES * nodes_per_ray + TRI_BYTES * tris_per_ray

    for i in range(args.warmup):
        be.render(spp=1, first_sample=i, accumulate=i > 0, want_image=False)
    be.sync()

    agg = {}
    t0 = time.perf_counter()
    for k in range(args.steps):
        _, st = be.render(spp=1, first_sample=args.warmup + k, accumulate=True, want_image=False)
        for key, v in st.items():
            agg[key] = agg.get(key, 0.0) + v
    be.sync()
    t1 = time.perf_counter()
    seconds = t1 - t0
    paths = agg["paths"]
    value = paths / seconds / 1e6
    achieved = agg["closest_rays"] * bytes_per_ray / (agg["ms_closest"] * 1e-3) / 1e9
    traffic = traffic_from_profiles()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(desc, args.cpu_seconds, min(16, os.cpu_count() or 1))

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * seconds / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene generated in code)",
        "config": {
            "workload": "displaced_sphere_2M",
            "triangles": int(desc.num_triangles),
            "width": W,
            "height": H,
            "spp": args.steps,
            "max_depth": desc.max_depth,
            "env_map": "1024x512",
            "bvh": "gpu-lbvh",
            "bvh_build_ms": round(accel["build_ms"], 2),
        },
        "mrays_per_s": round((agg["closest_rays"] + agg["shadow_rays"]) / seconds / 1e6, 2),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_trace_queue (closest hit)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_ray": round(bytes_per_ray, 1),
            "node_visits_per_ray": round(nodes_per_ray, 2),
            "tri_tests_per_ray": round(tris_per_ray, 2),
            "launches": int(agg["launches_closest"]),
            "avg_launch_ms": round(agg["ms_closest"] / max(1.0, agg["launches_closest"]), 4),
        },
        "stage_ms": {k: round(agg[k], 1) for k in ("ms_closest", "ms_shadow", "ms_shade", "ms_film", "ms_other")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    be.close()


if __name__ == "__main__":
    main()
