"""Render service (dl_hip_service_*): every job's image is bit-identical to
rendering the job alone with dl_hip_render; queued jobs can be cancelled;
device memory reserved by all lanes never exceeds the budget, and a job
that can never fit fails with a message."""
import ctypes
import time

import numpy as np
import pytest

import drawlab_amd
from drawlab_amd import scenes

pytestmark = pytest.mark.gpu


def _fixed_context_bytes():
    be = drawlab_amd.Backend(0)
    cur, peak = ctypes.c_uint64(), ctypes.c_uint64()
    drawlab_amd._service_lib().dl_hip_memory_usage(be.h, ctypes.byref(cur), ctypes.byref(peak))
    be.close()
    return cur.value


def test_jobs_match_standalone_renders(backend):
    cbox = scenes.cornell_box(64, 64, spp=8, max_depth=8, rr_depth=3)
    plates = scenes.mis_plates(64, 36, max_depth=6)
    jobs = [(cbox, 8, 1, 0), (plates, 5, 2, 0), (cbox, 3, 3, 2), (plates, 7, 4, 3), (cbox, 16, 5, 0),
            (cbox, 1, 6, 0)]
    svc = drawlab_amd.Service(lanes=3, policy=drawlab_amd.POLICY_FIFO)
    try:
        sid = {id(cbox): svc.add_scene(cbox), id(plates): svc.add_scene(plates)}
        ids = [svc.submit(sid[id(d)], spp, seed=seed, max_depth=d.max_depth, rr_depth=d.rr_depth, spp_per_pass=k)
               for d, spp, seed, k in jobs]
        print("submitted", ids, flush=True)
        infos = [svc.wait(j) for j in ids]
        assert all(i["state"] == drawlab_amd.JOB_DONE for i in infos), infos
        imgs = [svc.result(j, sid[id(d)]) for j, (d, _, _, _) in zip(ids, jobs)]
        st = svc.stats()
        print("stats", st, flush=True)
        assert st["jobs_done"] == len(jobs) and st["estimate_misses"] == 0
        assert st["peak_reserved_bytes"] <= st["budget_bytes"]
        # each lane uploads each scene at most once (then keeps it resident)
        assert st["scene_uploads"] <= 2 * 3
    finally:
        svc.stop()
    for (d, spp, seed, k), img, info in zip(jobs, imgs, infos):
        backend.upload_scene(d)
        backend.build_accel(drawlab_amd.BUILD_GPU)
        ref, rst = backend.render(spp=spp, seed=seed, spp_per_pass=k)
        assert info["paths"] == rst["paths"]
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (spp, seed, k)
        assert info["t_submit"] <= info["t_start"] <= info["t_finish"]


def test_resident_scenes(backend):
    """Scenes kept resident (dl_hip_scene_keep) and selected again render
    bit-identically to fresh uploads; dropping frees their memory."""
    a = scenes.cornell_box(64, 64, spp=8, max_depth=8, rr_depth=3)
    b = scenes.mis_plates(64, 36, max_depth=6)
    refs = []
    for d in (a, b):
        backend.upload_scene(d)
        backend.build_accel(drawlab_amd.BUILD_GPU)
        refs.append(backend.render(spp=8, seed=3)[0])
    backend.trim()
    base, _ = backend.memory_usage()
    for key, d in ((11, a), (12, b)):
        backend.upload_scene(d)
        backend.build_accel(drawlab_amd.BUILD_GPU)
        backend.scene_keep(key)
    for key, ref in ((11, refs[0]), (12, refs[1]), (11, refs[0])):
        assert backend.scene_select(key)
        img, _ = backend.render(spp=8, seed=3)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), key
    both, _ = backend.memory_usage()
    backend.scene_drop(12)
    assert not backend.scene_select(12)
    one, _ = backend.memory_usage()
    assert one < both
    img, _ = backend.render(spp=8, seed=3)          # scene 11 is still current
    assert np.array_equal(img.view(np.uint32), refs[0].view(np.uint32))
    backend.trim()
    assert backend.memory_usage()[0] <= base
    assert not backend.scene_select(11)


def test_job_camera_override(backend):
    """A job with its own camera (another resolution) equals rendering the
    scene with that camera alone."""
    import dataclasses
    cbox = scenes.cornell_box(64, 64, spp=8, max_depth=8, rr_depth=3)
    cam = dataclasses.replace(cbox.camera, width=48, height=40)
    svc = drawlab_amd.Service(lanes=2)
    try:
        s = svc.add_scene(cbox)
        a = svc.submit(s, 6, seed=3, max_depth=8, rr_depth=3, camera=cam)
        b = svc.submit(s, 6, seed=3, max_depth=8, rr_depth=3)
        assert svc.wait(a)["state"] == drawlab_amd.JOB_DONE and svc.wait(b)["state"] == drawlab_amd.JOB_DONE
        ia, ib = svc.result(a), svc.result(b)
    finally:
        svc.stop()
    assert ia.shape == (40, 48, 3) and ib.shape == (64, 64, 3)
    small = dataclasses.replace(cbox, camera=cam)
    backend.upload_scene(small)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    ref, _ = backend.render(spp=6, seed=3)
    assert np.array_equal(ia.view(np.uint32), ref.view(np.uint32))


def test_cancel_and_errors():
    cbox = scenes.cornell_box(64, 64, spp=8, max_depth=8, rr_depth=3)
    svc = drawlab_amd.Service(lanes=1)
    try:
        s = svc.add_scene(cbox)
        a = svc.submit(s, 256, max_depth=8, rr_depth=3)
        b = svc.submit(s, 4, max_depth=8, rr_depth=3)
        svc.cancel(b)
        assert svc.wait(b)["state"] == drawlab_amd.JOB_CANCELLED
        with pytest.raises(drawlab_amd.DlError, match="not done"):
            svc.result(b, s)
        assert svc.wait(a)["state"] == drawlab_amd.JOB_DONE
        img = svc.result(a, s)
        assert img.shape == (64, 64, 3) and np.isfinite(img).all()
        with pytest.raises(drawlab_amd.DlError, match="already taken"):
            svc.result(a, s)
        with pytest.raises(drawlab_amd.DlError, match="unknown scene id"):
            svc.submit(999, 4)
        with pytest.raises(drawlab_amd.DlError, match="spp must be positive"):
            svc.submit(s, 0)
        with pytest.raises(drawlab_amd.DlError, match="unknown job id"):
            svc.poll(12345)
    finally:
        svc.stop()


def _state_bytes_per_path():
    """Path-state bytes per path that the render service reserves (the
    constant of drawlab_amd/csrc/render_sizes.h)."""
    import os
    import re
    text = open(os.path.join(os.path.dirname(drawlab_amd.__file__), "csrc", "render_sizes.h")).read()
    segs = int(re.search(r"kMatQueueSegs = (\d+);", text).group(1))
    expr = re.search(r"kStateBytesPerPath = ([^;]+);", text).group(1).replace("kMatQueueSegs", str(segs))
    assert re.fullmatch(r"[\d\s+*]+", expr), expr
    return eval(expr)


def test_memory_budget_is_never_exceeded():
    fixed = _fixed_context_bytes()
    cbox = scenes.cornell_box(96, 96, spp=8, max_depth=8, rr_depth=3)
    # room for the two lanes' fixed buffers and one running job, not two,
    # whether on two lanes or sharing a pass (the service's estimate: scene +
    # state + slack, see service.cpp; two jobs in one pass hold about 1.7
    # times one job's bytes)
    one_job = (4 << 20) + 96 * 96 * 8 * _state_bytes_per_path() + 96 * 96 * 28 + (4 << 20) + 64 * 36 + 40 * 384
    budget = 2 * fixed + int(1.25 * one_job)
    svc = drawlab_amd.Service(lanes=2, budget_bytes=budget)
    try:
        s = svc.add_scene(cbox)
        ids = [svc.submit(s, 8, seed=i, max_depth=8, rr_depth=3) for i in range(6)]
        infos = [svc.wait(j) for j in ids]
        st = svc.stats()
        print("budget", budget, "fixed", fixed, "stats", st, flush=True)
        assert all(i["state"] == drawlab_amd.JOB_DONE for i in infos)
        assert st["peak_reserved_bytes"] <= budget and st["estimate_misses"] == 0
        # at most one job in flight at any time
        spans = sorted((i["t_start"], i["t_finish"]) for i in infos)
        for (s0, f0), (s1, f1) in zip(spans, spans[1:]):
            assert s1 >= f0 - 1e-4, spans
        for i in infos:
            assert i["peak_bytes"] <= i["reserved_bytes"]
        # a job larger than the budget can ever hold fails with a message
        big = svc.add_scene(scenes.cornell_box(1920, 1080, spp=1))
        j = svc.submit(big, 64)
        info = svc.wait(j)
        assert info["state"] == drawlab_amd.JOB_FAILED
        with pytest.raises(drawlab_amd.DlError, match="budget"):
            svc.result(j, big)
    finally:
        svc.stop()


def test_sjf_runs_short_jobs_first():
    cbox = scenes.cornell_box(256, 256, spp=8, max_depth=8, rr_depth=3)
    svc = drawlab_amd.Service(lanes=1, policy=drawlab_amd.POLICY_SJF)
    try:
        s = svc.add_scene(cbox)
        first = svc.submit(s, 1024, max_depth=8, rr_depth=3)   # occupies the only lane
        t0 = time.time()
        while svc.poll(first)["state"] == drawlab_amd.JOB_QUEUED and time.time() - t0 < 30:
            time.sleep(0.001)
        big = svc.submit(s, 64, max_depth=8, rr_depth=3)
        small = svc.submit(s, 2, max_depth=8, rr_depth=3)
        ib, is_, i0 = svc.wait(big), svc.wait(small), svc.wait(first)
        print("first", i0, "big", ib, "small", is_, flush=True)
        if i0["t_finish"] > ib["t_submit"]:       # both were queued while the lane was busy
            assert is_["t_start"] <= ib["t_start"]
    finally:
        svc.stop()


def test_service_workload_mix(backend):
    """The service's own job mix, bounded: jobs drawn by scenes.service_jobs
    over the service scene pool (the Cornell box, the MIS plates, the
    2,097,152-triangle displaced sphere and procedural scenes of 10k to
    500k triangles), every pool scene at least once and a 1920x1080 job
    among them, at a few samples per pixel, spread over four lanes.  Every
    job's image is bit-identical to its standalone render, and no lane's
    device memory exceeds its reservation or the budget."""
    import dataclasses
    pool = scenes.service_scene_pool()
    picked, seen, full_hd = [], set(), False
    for si, w, h, spp in scenes.service_jobs(512):
        if si in seen and (full_hd or (w, h) != (1920, 1080)):
            continue
        seen.add(si)
        full_hd |= (w, h) == (1920, 1080)
        picked.append((si, w, h, min(spp, 4)))
        if len(seen) == len(pool) and full_hd:
            break
    assert len(seen) == len(pool) and full_hd
    print("jobs", picked, flush=True)
    svc = drawlab_amd.Service(lanes=4, policy=drawlab_amd.POLICY_FIFO, budget_bytes=200 << 30)
    try:
        sids = [svc.add_scene(d) for d in pool]
        ids = []
        for k, (si, w, h, spp) in enumerate(picked):
            d = pool[si]
            cam = dataclasses.replace(d.camera, width=w, height=h)
            ids.append(svc.submit(sids[si], spp, seed=100 + k, max_depth=d.max_depth, rr_depth=d.rr_depth,
                                  camera=cam))
        infos = [svc.wait(j) for j in ids]
        assert all(i["state"] == drawlab_amd.JOB_DONE for i in infos), infos
        imgs = [svc.result(j) for j in ids]
        st = svc.stats()
    finally:
        svc.stop()
    print("stats", st, flush=True)
    assert st["estimate_misses"] == 0 and st["peak_reserved_bytes"] <= st["budget_bytes"]
    assert len({i["lane"] for i in infos}) > 1
    for i in infos:
        assert i["peak_bytes"] <= i["reserved_bytes"], i
    for k, ((si, w, h, spp), img, info) in enumerate(zip(picked, imgs, infos)):
        d = pool[si]
        print("standalone", k, d.name, w, h, spp, flush=True)
        backend.upload_scene(dataclasses.replace(d, camera=dataclasses.replace(d.camera, width=w, height=h)))
        backend.build_accel(drawlab_amd.BUILD_GPU)
        ref, rst = backend.render(spp=spp, seed=100 + k, max_depth=d.max_depth, rr_depth=d.rr_depth)
        assert img.shape == (h, w, 3) and info["paths"] == rst["paths"] == w * h * spp
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (k, d.name)


def test_guided_job_within_its_reservation():
    """A guided job (which renders alone and allocates mixtures, training
    records, the network and a training batch beside the path state) stays
    within the device bytes the service reserved for it."""
    cbox = scenes.cornell_box(64, 64, spp=16, max_depth=8, rr_depth=3)
    svc = drawlab_amd.Service(lanes=2)
    try:
        s = svc.add_scene(cbox)
        g = svc.submit(s, 16, seed=4, max_depth=8, rr_depth=3, guided=True)
        p = svc.submit(s, 16, seed=5, max_depth=8, rr_depth=3)
        ig, ip = svc.wait(g), svc.wait(p)
        st = svc.stats()
    finally:
        svc.stop()
    print("guided job", ig, "plain job", ip, "stats", st, flush=True)
    assert ig["state"] == drawlab_amd.JOB_DONE and ip["state"] == drawlab_amd.JOB_DONE
    assert ig["peak_bytes"] <= ig["reserved_bytes"] and ip["peak_bytes"] <= ip["reserved_bytes"]
    assert st["estimate_misses"] == 0
    assert ig["reserved_bytes"] > ip["reserved_bytes"] + (1 << 30)  # records + training batch
