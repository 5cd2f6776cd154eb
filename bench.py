#!/usr/bin/env python3
"""Throughput benchmark of the MI355X path tracer.

Workload: the 2,097,152-triangle displaced sphere (GGX conductor) with a
20,480-triangle glass sphere, a diffuse ground and a 1024x512 sky + sun
environment map, rendered at 1920x1080 with max depth 16, 256 spp.  One step
is one wavefront pass of the unidirectional MIS path tracer over the whole
frame at 256 samples per pixel (530,841,600 paths, the largest pass
dl_hip_render takes, ~120 GB of path state), accumulated progressively; the
default single step is the workload's 256 spp (--spp-per-step and --steps
split it into smaller passes; the other scenes default to 64 spp per step
and 4 steps).

Prints one JSON line: Mpaths/s over the timed steps, the roofline of the
dominant kernel (closest-hit traversal, k_trace_queue) and the CPU oracle's
throughput on a bounded sample of the same workload, timed on this host's
cores.

`--workload service`: the render-service workload instead -- jobs drawn over
the Cornell box, MIS plates, the 2M-triangle displaced sphere and procedural
10k-500k-triangle scenes, resolutions {256^2, 512^2, 1024^2, 1920x1080},
spp = floor(16 * 64^U); Poisson arrivals at 0.9x the one-at-a-time service
rate.  Prints p50/p99 job latency of the multi-lane service and of
one-at-a-time rendering, and both saturated throughputs (jobs/s).

`--workload trace`: the traversal micro-bench on the same geometry -- three
ray sets (2,073,600 primary rays, 2^22 diffuse-bounce rays and 2^22 shadow
rays toward the sky and sun, both from the points the primary rays hit)
through the batch traversal kernels, each with its own HBM roofline
fraction in algorithmic bytes.

`--workload guided`: the learned proposal against plain rendering at equal
wall time on the same scene at 480x270 -- a 65,536-spp plain reference, then
ladders of 16..4096 spp of both methods (the guided network trained from
scratch inside every timed render), each point averaged over three seeds.
Prints relMSE against wall time for both, the time each takes to reach
relMSE 1e-3 (when the ladder gets there) and the guided relMSE at the
plain render's time.

Profile helpers (they read rocprofv3 CSV output and write the summaries that
`roofline.traffic` and the kernel-time tables come from):
  bench.py --summarize-stats DIR OUT.md       kernel time table
  bench.py --summarize-pmc FETCH_DIR WRITE_DIR OUT.json
  bench.py --summarize-counters DIR [DIR ...] OUT.md   occupancy, VALU, L2 table
"""
import argparse
import csv
import glob
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s + relMSE@equal-time (guided vs plain); p50/p99 job latency"
WORKLOAD = "displaced_sphere_2M"
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, RAY_BYTES, HIT_BYTES = 64, 48, 32, 16  # 4-wide node, triangle, ray in, hit out
OCC_BYTES = 1  # shadow query result
# Path state each stage reads and writes per path or per shading vertex
# (structure-of-arrays records of the integrator, 16-byte vectors unless
# stated): camera ray generation writes ray origin and direction, the
# pcg32 state (8), sample index (4), film position (8), scene byte (1) and
# a queue entry (4); the first-hit stage reads the queue
# entry, hit and direction and writes radiance and a material-queue entry;
# a shading vertex reads its queue entry, hit, direction, throughput,
# aux, pcg32 state and sample index (a path's first vertex: not
# throughput and aux, which are implied) and writes the shadow ray's
# direction and its NEE term, the continuation ray (whose origin the shadow
# ray shares), throughput,
# pcg32 state, aux and two queue entries (radiance only when an emitter
# sample that casts no shadow ray adds to it); its post-processing
# reads aux, radiance, NEE term, occlusion byte, throughput, hit, direction,
# pcg32 state and sample index and writes radiance, aux, throughput, pcg32
# state and a queue entry; the film gather reads each sample's position and
# radiance once and writes a pixel per 64 samples.
PATH_STATE_BYTES = {"raygen": 2 * 16 + 8 + 4 + 8 + 1 + 4,                     # 57
                    "first_hit": 4 + 2 * 16 + 16 + 4,                          # 56
                    "shade": (4 + 4 * 16 + 8 + 4) + (6 * 16 + 8 + 2 * 4),      # 80 read + 112 written
                    "post": (4 + 6 * 16 + 1 + 8 + 4) + (3 * 16 + 8 + 4),       # 113 read + 60 written
                    "film": 8 + 16 + 16 / 64.0,
                    "first_vertex_skip": 2 * 16}                               # per path: an undercount
TRACE_KERNEL = "k_trace_queue<false>"
COPY_KERNEL = "k_copy"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_trace_queue_%s_traffic.json" % WORKLOAD)


def traffic_from_profiles():
    """HBM bytes per closest-hit traversal launch of this workload, measured
    by two rocprofv3 PMC runs (FETCH_SIZE, WRITE_SIZE) scaled by a copy
    kernel of known size; None when the summary is absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def host_cpus():
    """CPUs this process may run on (its affinity mask: the machine's share
    given to this job), the machine's logical CPU count and the CPU model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return avail, os.cpu_count() or avail, model


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
    avail, logical, model = host_cpus()
    return {
        "value": st["paths"] / st["seconds"] / 1e6,
        "unit": "Mpaths/s",
        "cores": st["threads"],
        "cpus_available": avail,
        "hardware_concurrency": logical,
        "cpu_model": model,
        "kind": "port",
        "sample": "first %d of %d 32x32 blocks of the spiral order (frame centre), 1 spp, %d paths, %.1f s; "
                  "reference octree (robust slabs) built in %.1f s, not timed" % (
                      blocks, total_blocks, int(st["paths"]), st["seconds"], build_s),
    }


def path_bytes(agg, bytes_closest, bytes_shadow):
    """Algorithmic bytes per path: its traversal queries (closest-hit and
    shadow, at their measured bytes per query) plus the path state its
    stages read and write (PATH_STATE_BYTES; every continuation ray comes
    from one shading vertex)."""
    paths = agg["paths"]
    vertices = agg["closest_rays"] - paths
    trav = (agg["closest_rays"] * (bytes_closest + 4) + agg["shadow_rays"] * (bytes_shadow + 4)) / paths
    # the throughput and aux a first shading vertex does not read are taken
    # off once per path, also for paths that never reach one (so the sum is
    # a lower bound)
    state = (paths * (PATH_STATE_BYTES["raygen"] + PATH_STATE_BYTES["first_hit"] + PATH_STATE_BYTES["film"] -
                      PATH_STATE_BYTES["first_vertex_skip"]) +
             vertices * (PATH_STATE_BYTES["shade"] + PATH_STATE_BYTES["post"])) / paths
    return {"traversal": round(trav, 1), "path_state": round(state, 1), "total": trav + state,
            "closest_rays_per_path": round(agg["closest_rays"] / paths, 3),
            "shadow_rays_per_path": round(agg["shadow_rays"] / paths, 3)}


COUNTERS_FILE = os.path.join(ROOT, "profiles", "bench_%s_counters.md" % WORKLOAD)


def counters_from_profiles(kernel):
    """Occupancy, VALU busy and L2 hit rate of `kernel` from the committed
    hardware-counter summary (None when absent)."""
    try:
        with open(COUNTERS_FILE) as f:
            rows = [l for l in f if l.startswith("| ")]
    except OSError:
        return None
    head = [c.strip() for c in rows[0].strip().strip("|").split("|")]
    for r in rows[1:]:
        cells = [c.strip() for c in r.strip().strip("|").split("|")]
        if cells[0] == kernel:
            d = dict(zip(head, cells))
            try:
                return {"occupancy_pct": float(d["OccupancyPercent"]), "valu_busy_pct": float(d["VALUBusy"]),
                        "l2_hit_pct": float(d["L2 hit %"])}
            except (KeyError, ValueError):
                return None
    return None


def limiter_from_profiles(kernel, traffic, ms_per_launch):
    """What limits the kernel below the HBM roofline, from its counters:
    actual HBM traffic near the peak means bandwidth; otherwise a busy VALU
    with misses to HBM means decode/ALU work behind dependent fetches, an
    idle one dependent-fetch latency."""
    c = counters_from_profiles(kernel)
    if c is None:
        return None
    actual = traffic / (ms_per_launch * 1e-3) / 1e9 if traffic else None
    if actual is not None and actual > 0.8 * HBM_PEAK_GBS:
        what = "HBM bandwidth"
    elif c["valu_busy_pct"] >= 50:
        what = "VALU (plane decode, slab and triangle tests) interleaved with dependent node fetches"
    else:
        what = "latency of dependent fetches"
    c["what"] = what
    if actual is not None:
        c["actual_hbm_gbs"] = round(actual, 1)
    return c


def short_name(name):
    n = name.replace("void ", "").replace("dl::(anonymous namespace)::", "").replace("dl::", "")
    if n.startswith("rocprim::"):
        return n.split("<")[0] + "<...>" + ("" if "(" not in n else "")
    return n.split("(")[0]


def summarize_stats(d, out_md, command, title=""):
    path = sorted(glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True))[0]
    rows = list(csv.DictReader(open(path)))
    title = title or "%s bench (1920x1080, 256 spp per pass, max depth 16)" % WORKLOAD
    lines = ["# Kernel time: %s" % title, "",
             "Command (repository root, MI355X):", "", "```", command, "```", "",
             "| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        lines.append("| %s | %s | %.2f | %.1f | %.1f | %.1f | %.2f |" % (
            short_name(r["Name"]), r["Calls"], float(r["TotalDurationNs"]) / 1e6, float(r["AverageNs"]) / 1e3,
            float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    with open(out_md, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines[:20]))


def _pmc_means(d, counter):
    path = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))[0]
    acc = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        acc.setdefault(short_name(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def summarize_pmc(fetch_dir, write_dir, out_json, copy_bytes, commands):
    f = _pmc_means(fetch_dir, "FETCH_SIZE")
    w = _pmc_means(write_dir, "WRITE_SIZE")
    cf = copy_bytes / (f[COPY_KERNEL][0] * 1024.0)
    cw = copy_bytes / (w[COPY_KERNEL][0] * 1024.0)
    fetch = f[TRACE_KERNEL][0] * 1024.0 * cf
    write = w[TRACE_KERNEL][0] * 1024.0 * cw
    out = {
        "kernel": TRACE_KERNEL,
        "workload": WORKLOAD + " (256 spp per pass)",
        "dispatches": f[TRACE_KERNEL][1],
        "fetch_kib_per_dispatch": f[TRACE_KERNEL][0],
        "write_kib_per_dispatch": w[TRACE_KERNEL][0],
        "copy_kernel_bytes": copy_bytes,
        "copy_fetch_kib": f[COPY_KERNEL][0],
        "copy_write_kib": w[COPY_KERNEL][0],
        "fetch_scale": cf,
        "write_scale": cw,
        "hbm_bytes_per_launch": fetch + write,
        "commands": commands,
        "other_kernels_fetch_kib": {k: v[0] for k, v in f.items()},
        "other_kernels_write_kib": {k: v[0] for k, v in w.items()},
    }
    with open(out_json, "w") as fo:
        json.dump(out, fo, indent=1)
    print(json.dumps({k: out[k] for k in list(out)[:12]}, indent=1))


def summarize_counters(dirs, out_md, command, kernels):
    """Per-kernel table of rocprofv3 counters (one column per counter, the
    dispatch-time-weighted mean over the kernel's dispatches) from several
    PMC runs' counter_collection.csv files."""
    acc = {}
    names = []
    for d in dirs:
        path = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))[0]
        for r in csv.DictReader(open(path)):
            k = short_name(r["Kernel_Name"])
            if k not in kernels:
                continue
            c = r["Counter_Name"]
            if c not in names:
                names.append(c)
            w = max(1.0, float(r["End_Timestamp"]) - float(r["Start_Timestamp"]))
            s = acc.setdefault((k, c), [0.0, 0.0, 0])
            s[0] += w * float(r["Counter_Value"])
            s[1] += w
            s[2] += 1
    lines = ["# Hardware counters: %s bench (1920x1080, 256 spp per pass)" % WORKLOAD, "",
             "Commands (repository root, MI355X):", "", "```", command, "```", "",
             "Dispatch-time-weighted means over each kernel's dispatches.  VALUBusy: % of GPU time the vector ALUs "
             "issue; VALUUtilization: % of active lanes per VALU instruction; OccupancyPercent: resident waves as "
             "% of the maximum; L2 hit: TCC_HIT / (TCC_HIT + TCC_MISS); MemUnitStalled: % of GPU time the memory "
             "unit is stalled.", "",
             "| kernel | dispatches | " + " | ".join(names) + " | L2 hit % |",
             "|---|---:|" + "---:|" * (len(names) + 1)]
    for k in kernels:
        row = []
        n = 0
        vals = {}
        for c in names:
            s = acc.get((k, c))
            v = s[0] / s[1] if s else float("nan")
            vals[c] = v
            n = max(n, s[2] if s else 0)
            row.append("%.3g" % v)
        h, m = vals.get("TCC_HIT_sum"), vals.get("TCC_MISS_sum")
        l2 = 100.0 * h / (h + m) if h is not None and m is not None and h + m > 0 else float("nan")
        lines.append("| %s | %d | %s | %.1f |" % (k, n, " | ".join(row), l2))
    with open(out_md, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def _run_jobs(svc, sids, pool, jobs, arrivals):
    """Submit jobs at their arrival times (seconds from now); wait for all.
    Returns the per-job latencies, the paths traced and the makespan."""
    import dataclasses
    t0 = time.perf_counter()
    ids = []
    for (si, w, h, spp), ta in zip(jobs, arrivals):
        d = time.perf_counter() - t0
        if ta > d:
            time.sleep(ta - d)
        desc = pool[si]
        cam = dataclasses.replace(desc.camera, width=w, height=h)
        ids.append(svc.submit(sids[si], spp, seed=len(ids), max_depth=desc.max_depth, rr_depth=desc.rr_depth,
                              camera=cam))
        if len(ids) % 64 == 0:
            print("  %d jobs submitted at %.1f s" % (len(ids), time.perf_counter() - t0), file=sys.stderr, flush=True)
    infos = []
    for j in ids:
        infos.append(svc.wait(j))
        if len(infos) % 64 == 0:
            print("  %d jobs done at %.1f s" % (len(infos), time.perf_counter() - t0), file=sys.stderr, flush=True)
    for j in ids:
        svc.result(j)
    for (si, w, h, spp), i in zip(jobs, infos):
        if i["peak_bytes"] > i["reserved_bytes"]:
            print("reservation exceeded: scene %d %dx%d spp %d lane %d reserved %d peak %d" % (
                si, w, h, spp, i["lane"], i["reserved_bytes"], i["peak_bytes"]), file=sys.stderr, flush=True)
    assert all(i["state"] == 2 for i in infos), [i["state"] for i in infos]
    lat = np.array([i["t_finish"] - i["t_submit"] for i in infos])
    span = max(i["t_finish"] for i in infos) - min(i["t_submit"] for i in infos)
    return lat, sum(i["paths"] for i in infos), span


def _run_standalone(pool, jobs, arrivals):
    """The jobs rendered one after another the way a single render loads
    its scene: upload, BVH build, render, on one context (no resident
    scenes, no shared passes).  Returns per-job latencies and the makespan."""
    import dataclasses
    import drawlab_amd
    be = drawlab_amd.Backend(0)
    try:
        t0 = time.perf_counter()
        done = []
        for k, ((si, w, h, spp), ta) in enumerate(zip(jobs, arrivals)):
            d = time.perf_counter() - t0
            if ta > d:
                time.sleep(ta - d)
            desc = pool[si]
            be.upload_scene(dataclasses.replace(desc, camera=dataclasses.replace(desc.camera, width=w, height=h)))
            be.build_accel(drawlab_amd.BUILD_GPU)
            be.render(spp=spp, seed=k, max_depth=desc.max_depth, rr_depth=desc.rr_depth)
            done.append(time.perf_counter() - t0)
    finally:
        be.close()
    lat = np.array([f - max(a, 0.0) for f, a in zip(done, arrivals)])
    return lat, done[-1] - min(arrivals)


def service_bench(args):
    import drawlab_amd
    from drawlab_amd import scenes

    pool = scenes.service_scene_pool()
    jobs = scenes.service_jobs(args.jobs)
    print("service workload: %d jobs, %.2f Gpaths" % (len(jobs), sum(w * h * s for _, w, h, s in jobs) / 1e9),
          file=sys.stderr, flush=True)
    policy = drawlab_amd.POLICY_SJF if args.policy == "sjf" else drawlab_amd.POLICY_FIFO
    results = {}

    def run(lanes, pol, arrivals):
        # one lane renders one job at a time (no shared passes)
        svc = drawlab_amd.Service(lanes=lanes, policy=pol, max_batch=1 if lanes == 1 else 0,
                                  single_scene_passes=args.single_scene_passes)
        try:
            sids = [svc.add_scene(d) for d in pool]
            lat, paths, span = _run_jobs(svc, sids, pool, jobs, arrivals)
            st = svc.stats()
        finally:
            svc.stop()
        return lat, paths, span, st

    zero = [0.0] * len(jobs)
    # saturated (all jobs at t = 0): one-at-a-time, then the service
    lat1, paths, span1, _ = run(1, drawlab_amd.POLICY_FIFO, zero)
    print("one-at-a-time saturated: %.2f s, %.2f jobs/s" % (span1, len(jobs) / span1), file=sys.stderr, flush=True)
    latn, _, spann, stn = run(args.lanes, policy, zero)
    print("service saturated: %.2f s, %.2f jobs/s" % (spann, len(jobs) / spann), file=sys.stderr, flush=True)
    # Poisson arrivals at 0.9x the one-at-a-time rate
    cap1 = len(jobs) / span1
    rate = 0.9 * cap1
    rng = np.random.default_rng(7)
    unit = np.cumsum(rng.exponential(1.0, len(jobs)))  # arrivals at rate 1, scaled per rate
    arrivals = unit / rate
    lat1p, _, _, _ = run(1, drawlab_amd.POLICY_FIFO, arrivals)
    latnp, _, _, stp = run(args.lanes, policy, arrivals)
    # p99 latency against the arrival rate (same arrival pattern, rescaled):
    # the rate each mode sustains at the one-at-a-time p99 of the 0.9x run
    curve1 = [(0.9 * cap1, float(np.percentile(lat1p, 99)))]
    curven = [(0.9 * cap1, float(np.percentile(latnp, 99)))]
    for f in args.rate_sweep:
        r = f * cap1
        if f < 1.0:
            l1, _, _, _ = run(1, drawlab_amd.POLICY_FIFO, unit / r)
            curve1.append((r, float(np.percentile(l1, 99))))
        ln, _, _, _ = run(args.lanes, policy, unit / r)
        curven.append((r, float(np.percentile(ln, 99))))
        print("rate %.2f jobs/s: p99 one-at-a-time %s, service %.3f s" % (
            r, "%.3f s" % curve1[-1][1] if f < 1.0 else "-", curven[-1][1]), file=sys.stderr, flush=True)
    curve1.sort()
    curven.sort()
    target = float(np.percentile(lat1p, 99))

    def rate_at(curve, p99):
        # highest rate whose p99 stays within p99 (linear between points)
        best = None
        for (r0, l0), (r1, l1) in zip([(0.0, 0.0)] + curve[:-1], curve):
            if l1 <= p99:
                best = r1
            elif l0 <= p99:
                best = r0 + (r1 - r0) * (p99 - l0) / (l1 - l0)
                break
            else:
                break
        return best

    r1 = rate_at(curve1, target)
    rn = rate_at(curven, target)
    # the reference's flow: every job loads its scene (upload + BVH build)
    curves = []
    if args.skip_standalone:
        lats, spans = None, None
    else:
        lats, spans = _run_standalone(pool, jobs, zero)
        print("standalone saturated: %.2f s, %.2f jobs/s" % (spans, len(jobs) / spans), file=sys.stderr, flush=True)
        caps = len(jobs) / spans
        for f in (0.6, 0.8, 0.9):
            ls, _ = _run_standalone(pool, jobs, unit / (f * caps))
            curves.append((f * caps, float(np.percentile(ls, 99))))
            print("standalone at %.2f jobs/s: p99 %.3f s" % (f * caps, curves[-1][1]), file=sys.stderr, flush=True)
    target_s = curves[-1][1] if curves else None
    rs_ = rate_at(curves, target_s) if curves else None
    rn_s = rate_at(curven, target_s) if curves else None

    def q(a):
        return {"p50": round(float(np.percentile(a, 50)), 4), "p99": round(float(np.percentile(a, 99)), 4),
                "mean": round(float(a.mean()), 4)}

    out = {
        "metric": METRIC,
        "value": q(latnp)["p99"],
        "unit": "s (p99 job latency)",
        "n_gpus": 1,
        "steps": len(jobs),
        "warmup": 0,
        "ms_per_step": round(1e3 * spann / len(jobs), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scenes and job list)",
        "config": {"workload": "render_service", "jobs": len(jobs), "lanes": args.lanes, "policy": args.policy,
                   "arrival_rate_hz": round(rate, 4), "scenes": [d.name for d in pool],
                   "gpaths": round(paths / 1e9, 3)},
        "latency_s": {"service": q(latnp), "one_at_a_time": q(lat1p)},
        "p99_vs_rate": {"one_at_a_time": [[round(r, 3), round(l, 3)] for r, l in curve1],
                        "service": [[round(r, 3), round(l, 3)] for r, l in curven]},
        "equal_p99": {"p99_s": round(target, 3), "rate_one_at_a_time": r1, "rate_service": rn,
                      "ratio": (rn / r1) if (rn and r1) else None,
                      "note": "service rate is a lower bound when it stays within the p99 at every swept rate"},
        "saturated_jobs_per_s": {"service": round(len(jobs) / spann, 4),
                                 "one_at_a_time": round(len(jobs) / span1, 4),
                                 "standalone": round(len(jobs) / spans, 4) if spans else None},
        "standalone_equal_p99": {"p99_s": round(target_s, 3) if target_s else None, "rate_standalone": rs_,
                                 "rate_service": rn_s,
                                 "ratio": (rn_s / rs_) if (rn_s and rs_) else None,
                                 "p99_vs_rate": [[round(r, 3), round(l, 3)] for r, l in curves],
                                 "note": "standalone = per job: scene upload + BVH build + render, one context"},
        "saturated_mpaths_per_s": {"service": round(paths / spann / 1e6, 1),
                                   "one_at_a_time": round(paths / span1 / 1e6, 1)},
        "service_stats": {"peak_reserved_bytes": int(stp["peak_reserved_bytes"]),
                          "estimate_misses": int(stp["estimate_misses"] + stn["estimate_misses"]),
                          "scene_uploads": int(stp["scene_uploads"]),
                          "batched_jobs": int(stp["batched_jobs"]), "shared_passes": int(stp["shared_passes"]),
                          "cross_scene_passes": int(stp["cross_scene_passes"]),
                          "batched_jobs_saturated": int(stn["batched_jobs"]),
                          "shared_passes_saturated": int(stn["shared_passes"]),
                          "cross_scene_passes_saturated": int(stn["cross_scene_passes"]),
                          "single_scene_passes": bool(args.single_scene_passes)},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def _relmse(x, y):
    """mean over pixels and channels of (x - y)^2 / (y^2 + 1e-2)."""
    return float(np.mean((x.astype(np.float64) - y) ** 2 / (y.astype(np.float64) ** 2 + 1e-2)))


def _time_to(target, times, errs):
    """Wall time at which relMSE first reaches `target`, interpolated
    log-log between the two ladder renders that bracket it (None if the
    ladder never reaches it)."""
    for i, e in enumerate(errs):
        if e <= target:
            if i == 0:
                return times[0]
            t0, t1, e0, e1 = times[i - 1], times[i], errs[i - 1], errs[i]
            if e0 <= e1:
                return t1
            a = (np.log(e0) - np.log(target)) / (np.log(e0) - np.log(e1))
            return float(np.exp(np.log(t0) + a * (np.log(t1) - np.log(t0))))
    return None


def _err_at(t, times, errs):
    """relMSE at wall time t, interpolated log-log on the ladder."""
    if t <= times[0]:
        return errs[0]
    for i in range(1, len(times)):
        if t <= times[i]:
            a = (np.log(t) - np.log(times[i - 1])) / (np.log(times[i]) - np.log(times[i - 1]))
            return float(np.exp(np.log(errs[i - 1]) + a * (np.log(errs[i]) - np.log(errs[i - 1]))))
    return None


def guided_bench(args):
    """Guided against plain rendering at equal wall time.  The reference
    is the plain HIP integrator at --ref-spp samples per pixel (its own
    per-path parity with the CPU oracle is tested separately); each method
    renders a ladder of sample counts from scratch -- the guided one with a
    freshly initialised network trained online inside the timed render --
    and the wall time of every render (including the image download) is
    paired with its relMSE against the reference, both averaged over
    --seeds independent renders (relMSE of one render is dominated by a few
    rare paths and varies several-fold between seeds)."""
    import drawlab_amd
    from drawlab_amd import scenes

    W, H = (int(v) for v in args.guided_res.lower().split("x"))
    if args.guided_scene == "chimney_room":
        desc = scenes.chimney_room(W, H, max_depth=6, rr_depth=3)
    else:
        desc = scenes.displaced_sphere(width=W, height=H)
    be = drawlab_amd.Backend(0)
    be.upload_scene(desc)
    be.build_accel(drawlab_amd.BUILD_GPU)
    ref_seed, test_seed, net_seed = 0x5EED + 1, 7, 0x5EED
    t0 = time.perf_counter()
    ref, _ = be.render(spp=args.ref_spp, seed=ref_seed)
    t_ref = time.perf_counter() - t0
    print("reference: %dx%d x %d spp in %.1f s" % (W, H, args.ref_spp, t_ref), file=sys.stderr, flush=True)
    ladder = [16 << i for i in range(64) if (16 << i) <= args.max_spp]

    def one(spp, guided, seed):
        be.sync()
        t = time.perf_counter()
        if guided:
            be.guide_init(seed=net_seed + seed, lr=1e-2)
        img, st = be.render(spp=spp, seed=seed, guided=guided, guide_steps=args.guide_steps)
        be.sync()
        return time.perf_counter() - t, img, st

    one(16, False, 1)
    one(16, True, 1)
    seeds = [test_seed + k for k in range(args.seeds)]
    curves = {}
    for name, g in (("plain", False), ("guided", True)):
        ts, es, gms, trs = [], [], [], []
        for spp in ladder:
            runs = [one(spp, g, s) for s in seeds]
            ts.append(float(np.mean([r[0] for r in runs])))
            es.append(float(np.mean([_relmse(r[1], ref) for r in runs])))
            gms.append(float(np.mean([r[2]["ms_guide"] for r in runs])))
            trs.append(float(np.mean([r[2]["ms_train"] for r in runs])))
            print("%-6s %6d spp  %8.3f s  relMSE %.4e  (inference %.1f ms, training %.1f ms)" % (
                name, spp, ts[-1], es[-1], gms[-1], trs[-1]), file=sys.stderr, flush=True)
        curves[name] = {"spp": ladder, "seconds": [round(v, 4) for v in ts], "relmse": [float("%.5g" % v) for v in es],
                        "ms_inference": [round(v, 2) for v in gms], "ms_training": [round(v, 2) for v in trs]}
    target = args.target_relmse
    tp = _time_to(target, curves["plain"]["seconds"], curves["plain"]["relmse"])
    tg = _time_to(target, curves["guided"]["seconds"], curves["guided"]["relmse"])
    eq = _err_at(tp, curves["guided"]["seconds"], curves["guided"]["relmse"]) if tp is not None else None
    out = {
        "metric": METRIC,
        "value": eq,
        "unit": "relMSE of the guided render at the plain render's time to relMSE %g" % target,
        "n_gpus": 1,
        "steps": len(ladder),
        "warmup": 2,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (network bf16 with f32 accumulation)",
        "data": "synthetic (seeded procedural scene generated in code)",
        "config": {"workload": "guided_vs_plain_" + ("chimney_room" if args.guided_scene == "chimney_room"
                                                     else "displaced_sphere_2M"), "width": W, "height": H,
                   "reference_spp": args.ref_spp, "reference_seconds": round(t_ref, 2), "target_relmse": target,
                   "seeds_per_point": args.seeds,
                   "guide_steps_per_pass": args.guide_steps or "16 for the first 256 steps, then 4"},
        "time_to_target_s": {"plain": tp, "guided": tg,
                             "ratio_guided_over_plain": (tg / tp) if (tp and tg) else None},
        "curves": curves,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def primary_rays(desc, seed=0x5EED):
    """One camera ray per pixel of the frame through a uniformly jittered
    film position (VirtualCamera::sampleRay, src/tracer/camera/virtual.cpp:
    52-65); 8 planes of n floats."""
    import drawlab_amd

    rng = np.random.default_rng(seed)
    f32 = np.float32
    cd = drawlab_amd.camera_desc(desc.camera)
    W, H = desc.camera.width, desc.camera.height
    eye, U, V, Wv = (np.array(getattr(cd, k)[:], f32) for k in ("eye", "U", "V", "W"))
    ys, xs = np.divmod(np.arange(W * H), W)
    px = (xs + rng.random(W * H)).astype(f32)
    py = (ys + rng.random(W * H)).astype(f32)
    dx = (px * f32(1.0 / W) * f32(2) - f32(1)).astype(f32)
    dy = (((f32(H) - py) - f32(1)) * f32(1.0 / H) * f32(2) - f32(1)).astype(f32)
    d = dx[:, None] * U + dy[:, None] * V + Wv
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(f32)
    return _planes(np.broadcast_to(eye, d.shape), d, 1e-4, np.inf)


def secondary_rays(desc, primary, prim, u, v, n=1 << 22, seed=0x5EED + 1):
    """Rays leaving the surface points the primary rays hit (n drawn with
    replacement; barycentric hit points, face normals turned toward the
    camera):
      bounce  cosine-distributed directions about the normal, as diffuse
              bounces leave;
      shadow  directions drawn from the sky-and-sun map in proportion to
              max(rgb) times the pixel's solid angle (the render's emitter
              sampling; about three quarters go to the sun), directions
              below the surface redrawn, ending 1e16 away as the render's
              environment shadow rays do."""
    rng = np.random.default_rng(seed)
    Vs = np.concatenate([m.V for m in desc.meshes]).astype(np.float64)
    offs = np.cumsum([0] + [m.V.shape[0] for m in desc.meshes])[:-1]
    F = np.concatenate([m.F.astype(np.int64) + o for m, o in zip(desc.meshes, offs)])
    hits = np.flatnonzero(prim != 0xFFFFFFFF)
    pick = hits[rng.integers(0, hits.size, n)]
    f = prim[pick].astype(np.int64)
    P0, P1, P2 = Vs[F[f, 0]], Vs[F[f, 1]], Vs[F[f, 2]]
    b1, b2 = u[pick].astype(np.float64)[:, None], v[pick].astype(np.float64)[:, None]
    o = P0 * (1 - b1 - b2) + P1 * b1 + P2 * b2
    nrm = np.cross(P1 - P0, P2 - P0)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    dcam = primary[3:6, pick].T.astype(np.float64)
    nrm *= np.where((nrm * dcam).sum(1) > 0, -1.0, 1.0)[:, None]
    t = np.where(np.abs(nrm[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(nrm, t)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(nrm, t)
    r, phi = np.sqrt(rng.random(n)), 2 * np.pi * rng.random(n)
    z = np.sqrt(np.maximum(0.0, 1 - r * r))
    dd = t * (r * np.cos(phi))[:, None] + b * (r * np.sin(phi))[:, None] + nrm * z[:, None]
    bounce = _planes(o, dd, 1e-4, np.inf)

    env = desc.envmap.rgb.astype(np.float64)
    h, w = env.shape[:2]
    th = np.pi * np.arange(h + 1) / h
    band = (np.cos(th[:-1]) - np.cos(th[1:]))[:, None] * np.ones((1, w))
    pcdf = np.cumsum((env.max(axis=2) * band).reshape(-1))
    pcdf /= pcdf[-1]
    dirs = np.zeros((n, 3))
    todo = np.arange(n)
    for _ in range(64):
        if not todo.size:
            break
        k = np.minimum(np.searchsorted(pcdf, rng.random(todo.size)), len(pcdf) - 1)
        yy, xx = np.divmod(k, w)
        ph = (xx + rng.random(todo.size)) / w * 2 * np.pi - np.pi
        c0, c1 = np.cos(th[yy]), np.cos(th[yy + 1])
        ct = c0 + (c1 - c0) * rng.random(todo.size)
        st = np.sqrt(np.maximum(0.0, 1 - ct * ct))
        cand = np.stack([np.cos(ph) * st, ct, np.sin(ph) * st], 1)
        ok = (cand * nrm[todo]).sum(1) > 0
        dirs[todo[ok]] = cand[ok]
        todo = todo[~ok]
    dirs[todo] = nrm[todo]  # points that see no part of the map above them (none in this scene)
    shadow = _planes(o, dirs, 1e-4, 1e16)
    return bounce, shadow


def _planes(o, d, tmin, tmax):
    n = d.shape[0]
    r = np.empty((8, n), np.float32)
    r[0:3] = np.asarray(o, np.float32).T
    r[3:6] = np.asarray(d, np.float32).T
    r[6] = tmin
    r[7] = tmax
    return r


def trace_bench(args):
    """Traversal micro-bench: each ray set traced by the batch kernels
    (dl_hip_trace_closest / dl_hip_trace_shadow, persistent grids over
    device-resident rays), HIP-event time of the kernel on the context's
    stream, median of --steps runs after --warmup; node visits and triangle
    tests per ray from the counting variants (untimed), giving algorithmic
    bytes per ray B = 32 + 16 (shadow: 1) + 64 N_node + 48 N_tri."""
    import drawlab_amd
    from drawlab_amd import scenes

    desc = scenes.displaced_sphere()
    t0 = time.perf_counter()
    primary = primary_rays(desc)
    be = drawlab_amd.Backend(0)
    be.upload_scene(desc)
    accel = be.build_accel(drawlab_amd.BUILD_GPU)
    prim, _, u, v = be.trace_closest(primary)
    bounce, shadow = secondary_rays(desc, primary, prim, u, v)
    gen_s = time.perf_counter() - t0
    sets = {"primary": primary, "bounce": bounce, "shadow": shadow}
    res = {}
    total_rays, total_ms = 0, 0.0
    for name, rays in sets.items():
        n = rays.shape[1]
        any_hit = name == "shadow"
        d_rays = be.alloc(rays.nbytes)
        d_rays.upload(np.ascontiguousarray(rays))
        outs = [be.alloc(max(1, n)) if any_hit else be.alloc(4 * n) for _ in range(1 if any_hit else 4)]
        cnt = [be.alloc(4 * n) for _ in range(2)]
        if any_hit:
            be.trace_shadow_dev(d_rays, n, outs[0], *cnt)
        else:
            be.trace_closest_dev(d_rays, n, *outs, *cnt)
        be.sync()
        nodes = cnt[0].download(np.uint32, n).astype(np.float64).mean()
        tris = cnt[1].download(np.uint32, n).astype(np.float64).mean()
        ms = []
        for k in range(args.warmup + args.steps):
            if any_hit:
                be.trace_shadow_dev(d_rays, n, outs[0])
            else:
                be.trace_closest_dev(d_rays, n, *outs)
            be.sync()
            if k >= args.warmup:
                ms.append(be.last_kernel_ms())
        hits = outs[0].download(np.uint8 if any_hit else np.uint32, n)
        hit_frac = float((hits != 0).mean() if any_hit else (hits != 0xFFFFFFFF).mean())
        for b in outs + cnt + [d_rays]:
            b.free()
        med = float(np.median(ms))
        bpr = RAY_BYTES + (OCC_BYTES if any_hit else HIT_BYTES) + NODE_BYTES * nodes + TRI_BYTES * tris
        gbs = n * bpr / (med * 1e-3) / 1e9
        res[name] = {"rays": n, "kernel": "k_trace_shadow" if any_hit else "k_trace_closest<false>",
                     "ms": round(med, 4), "mrays_per_s": round(n / (med * 1e-3) / 1e6, 1),
                     "node_visits_per_ray": round(nodes, 3), "tri_tests_per_ray": round(tris, 3),
                     "bytes_per_ray": round(bpr, 1), "achieved_gbs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4), ("occluded" if any_hit else "hit") + "_fraction":
                         round(hit_frac, 4)}
        total_rays += n
        total_ms += med
        print("%-8s %9d rays  %8.3f ms  %8.1f Mrays/s  %6.1f B/ray  %7.1f GB/s (%.3f of HBM peak)" % (
            name, n, med, res[name]["mrays_per_s"], bpr, gbs, gbs / HBM_PEAK_GBS), file=sys.stderr, flush=True)
    value = total_rays / (total_ms * 1e-3) / 1e6
    dom = max(res, key=lambda k: res[k]["ms"])
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene and ray sets generated in code)",
        "config": {"workload": "traversal_" + WORKLOAD, "triangles": int(desc.num_triangles),
                   "bvh": "gpu-ploc, quantised 4-wide nodes", "bvh_build_ms": round(accel["build_ms"], 2),
                   "ray_generation_s": round(gen_s, 1)},
        "ray_sets": res,
        "roofline": {"bound": "hbm", "kernel": res[dom]["kernel"] + " (%s rays)" % dom,
                     "achieved": res[dom]["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": res[dom]["frac"], "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    be.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 1 on the displaced sphere, the 256-spp frame; else 4)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp-per-step", type=int, default=None,
                    help="samples per pixel of a step, one wavefront pass (default: 256 on the displaced "
                         "sphere, 2^29 paths at 1080p; else 64)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="after the timed steps, run the plain copy kernel twice over this many MiB")
    ap.add_argument("--summarize-stats", nargs=2, metavar=("DIR", "OUT_MD"))
    ap.add_argument("--summarize-pmc", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "OUT_JSON"))
    ap.add_argument("--summarize-counters", nargs="+", metavar="DIR_OR_OUT",
                    help="PMC run directories, then the output .md")
    ap.add_argument("--command", default="", help="command line recorded in a summary")
    ap.add_argument("--title", default="", help="title of a kernel-time summary")
    ap.add_argument("--copy-mib", type=int, default=1024, help="calibration copy size used by the PMC runs")
    ap.add_argument("--scene", choices=("displaced_sphere", "cornell_box", "mis_plates"), default="displaced_sphere",
                    help="plain workload: the scene (1920x1080 displaced sphere, 512x512 Cornell box, 1280x720 "
                         "MIS plates, at their configured depths)")
    ap.add_argument("--workload", choices=("displaced_sphere", "service", "guided", "trace"),
                    default="displaced_sphere")
    ap.add_argument("--guided-res", default="480x270", help="guided workload: image size WxH")
    ap.add_argument("--guided-scene", choices=("displaced_sphere", "chimney_room"), default="displaced_sphere",
                    help="guided workload: the 2M-triangle scene, or a room lit through a small ceiling patch")
    ap.add_argument("--ref-spp", type=int, default=65536, help="guided workload: reference samples per pixel")
    ap.add_argument("--max-spp", type=int, default=4096, help="guided workload: largest ladder render")
    ap.add_argument("--target-relmse", type=float, default=1e-3)
    ap.add_argument("--seeds", type=int, default=3, help="guided workload: renders averaged per ladder point")
    ap.add_argument("--guide-steps", type=int, default=0,
                    help="guided workload: Adam steps per pass (0: 16 for the first 256 steps, then 4)")
    ap.add_argument("--jobs", type=int, default=64, help="service workload: number of jobs")
    ap.add_argument("--lanes", type=int, default=4, help="service workload: worker lanes")
    ap.add_argument("--policy", choices=("fifo", "sjf"), default="fifo")
    ap.add_argument("--single-scene-passes", action="store_true",
                    help="service workload: only jobs of one scene share a pass")
    ap.add_argument("--skip-standalone", action="store_true",
                    help="service workload: no standalone (upload + build + render per job) baseline")
    ap.add_argument("--rate-sweep", type=float, nargs="*", default=[0.6, 0.8, 1.0, 1.2, 1.4],
                    help="service workload: arrival rates (x one-at-a-time capacity) of the p99 sweep")
    args = ap.parse_args()
    # the 1080p displaced sphere renders its 256 spp in one pass of 2^29
    # paths (about 120 GB of path state): larger passes amortise the sparse
    # last bounces (128 spp per pass measured 4% faster than 64 there, 256
    # another 3-5%; the smaller scenes run faster at 64)
    big = args.workload == "displaced_sphere" and args.scene == "displaced_sphere"
    if args.spp_per_step is None:
        args.spp_per_step = 256 if big else 64
    if args.steps is None:
        args.steps = 1 if big else 4
    if args.summarize_stats:
        summarize_stats(args.summarize_stats[0], args.summarize_stats[1], args.command, args.title)
        return
    if args.summarize_counters:
        summarize_counters(args.summarize_counters[:-1], args.summarize_counters[-1], args.command,
                           ["k_trace_queue<false>", "k_shadow_queue<false>", "k_post<false>", "k_shade<0, false>",
                            "k_shade<5, false>", "k_film_tiled<2>"])
        return
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, copy_bytes=args.copy_mib << 20, commands=args.command)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has a single MI355X, so --gpus %d is not run "
              "(the frame would shard by sample index across GPUs)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "service":
        service_bench(args)
        return
    if args.workload == "guided":
        guided_bench(args)
        return
    if args.workload == "trace":
        trace_bench(args)
        return

    import drawlab_amd
    from drawlab_amd import scenes

    # the plain integrator on one of the configured scenes: the 2M-triangle
    # displaced sphere (the headline workload), the Cornell box or the MIS plates
    desc = {"displaced_sphere": scenes.displaced_sphere, "cornell_box": scenes.cornell_box,
            "mis_plates": scenes.mis_plates}[args.scene]()
    workload = WORKLOAD if args.scene == "displaced_sphere" else args.scene
    W, H = desc.camera.width, desc.camera.height
    S = args.spp_per_step
    be = drawlab_amd.Backend(0)
    be.upload_scene(desc)
    accel = be.build_accel(drawlab_amd.BUILD_GPU)

    # Untimed pass with in-kernel counters: traversal work per closest-hit
    # and per shadow ray.
    _, cst = be.render(spp=S, spp_per_pass=S, seed=1 << 20, count_traversal=True, want_image=False)
    nodes_per_ray = cst["nodes_closest"] / cst["closest_rays"]
    tris_per_ray = cst["tris_closest"] / cst["closest_rays"]
    bytes_per_ray = RAY_BYTES + HIT_BYTES + NODE_BYTES * nodes_per_ray + TRI_BYTES * tris_per_ray
    nodes_per_shadow = cst["nodes_shadow"] / max(1.0, cst["shadow_rays"])
    tris_per_shadow = cst["tris_shadow"] / max(1.0, cst["shadow_rays"])
    bytes_per_shadow = RAY_BYTES + OCC_BYTES + NODE_BYTES * nodes_per_shadow + TRI_BYTES * tris_per_shadow

    for i in range(args.warmup):
        be.render(spp=S, spp_per_pass=S, first_sample=i * S, accumulate=i > 0, want_image=False)
    be.sync()

    agg = {}
    step_ms = []  # HIP-event time of each step (each render call synchronises)
    t0 = time.perf_counter()
    for k in range(args.steps):
        _, st = be.render(spp=S, spp_per_pass=S, first_sample=(args.warmup + k) * S, accumulate=True,
                          want_image=False)
        step_ms.append(st["ms_total"])
        for key, v in st.items():
            agg[key] = agg.get(key, 0.0) + v
    be.sync()
    t1 = time.perf_counter()
    seconds = t1 - t0
    paths = agg["paths"]
    value = paths / seconds / 1e6
    launches = max(1.0, agg["launches_closest"])
    bytes_per_launch = agg["closest_rays"] * bytes_per_ray / launches
    ms_per_launch = agg["ms_closest"] / launches
    achieved = bytes_per_launch / (ms_per_launch * 1e-3) / 1e9
    # counter profiles exist for the headline workload only
    headline = args.scene == "displaced_sphere"
    traffic = traffic_from_profiles() if headline else None
    sh_launches = max(1.0, agg["launches_shadow"])
    sh_bytes_per_launch = agg["shadow_rays"] * bytes_per_shadow / sh_launches
    sh_ms_per_launch = agg["ms_shadow"] / sh_launches
    sh_achieved = sh_bytes_per_launch / (sh_ms_per_launch * 1e-3) / 1e9
    # rays: average algorithmic bytes per traversal query; paths: per path
    rays = agg["closest_rays"] + agg["shadow_rays"]
    bytes_per_query = (agg["closest_rays"] * bytes_per_ray + agg["shadow_rays"] * bytes_per_shadow) / rays
    mrays = rays / seconds / 1e6
    bpath = path_bytes(agg, bytes_per_ray, bytes_per_shadow)

    if args.calibration_copy_mib > 0:
        n = args.calibration_copy_mib << 20
        a, b = drawlab_amd.DeviceBuffer(be, n), drawlab_amd.DeviceBuffer(be, n)
        for _ in range(2):
            b.copy_from(a)
        be.sync()
        a.free()
        b.free()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(desc, args.cpu_seconds, host_cpus()[0])

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * seconds / args.steps, 3),
        "step_ms": [round(x, 3) for x in step_ms],
        "mpaths_per_s_median_step": round(W * H * S / (float(np.median(step_ms)) * 1e-3) / 1e6, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene generated in code)",
        "config": {
            "workload": workload,
            "triangles": int(desc.num_triangles),
            "width": W,
            "height": H,
            "spp": args.steps * S,
            "spp_per_step": S,
            "max_depth": desc.max_depth,
            "rr_depth": desc.rr_depth,
            "env_map": "%dx%d" % (desc.envmap.rgb.shape[1], desc.envmap.rgb.shape[0]) if desc.envmap is not None
            else None,
            "bvh": "gpu-ploc",
            "bvh_build_ms": round(accel["build_ms"], 2),
        },
        "mrays_per_s": round(mrays, 2),
        # Mrays/s and Mpaths/s against the rates the HBM roofline allows at
        # the algorithmic bytes of a query and of a path
        "mrays_roofline": {"bytes_per_query": round(bytes_per_query, 1),
                           "roofline_mrays_per_s": round(HBM_PEAK_GBS * 1e3 / bytes_per_query, 1),
                           "frac": round(mrays / (HBM_PEAK_GBS * 1e3 / bytes_per_query), 4)},
        "mpaths_roofline": {"bytes_per_path": round(bpath["total"], 1), "breakdown": bpath,
                            "roofline_mpaths_per_s": round(HBM_PEAK_GBS * 1e3 / bpath["total"], 1),
                            "frac": round(value / (HBM_PEAK_GBS * 1e3 / bpath["total"]), 4)},
        "roofline": {
            "bound": "hbm",
            "limiter": limiter_from_profiles(TRACE_KERNEL, traffic, ms_per_launch) if headline else None,
            "kernel": TRACE_KERNEL + " (closest hit)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": round(bytes_per_launch),
            "bytes_per_ray": round(bytes_per_ray, 1),
            "node_visits_per_ray": round(nodes_per_ray, 2),
            "tri_tests_per_ray": round(tris_per_ray, 2),
            "launches": int(launches),
            "avg_launch_ms": round(ms_per_launch, 4),
        },
        "roofline_shadow": {
            "bound": "hbm",
            "limiter": limiter_from_profiles("k_shadow_queue<false>", None, sh_ms_per_launch) if headline else None,
            "kernel": "k_shadow_queue<false> (any hit)",
            "achieved": round(sh_achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(sh_achieved / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": round(sh_bytes_per_launch),
            "bytes_per_ray": round(bytes_per_shadow, 1),
            "node_visits_per_ray": round(nodes_per_shadow, 2),
            "tri_tests_per_ray": round(tris_per_shadow, 2),
            "launches": int(sh_launches),
            "avg_launch_ms": round(sh_ms_per_launch, 4),
        },
        "stage_ms": {k: round(agg[k], 1) for k in ("ms_closest", "ms_shadow", "ms_shade", "ms_film", "ms_other")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    be.close()


if __name__ == "__main__":
    main()
