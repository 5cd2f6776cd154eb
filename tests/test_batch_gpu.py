"""Shared wavefront passes (dl_hip_render_batch): several images of one
scene -- different cameras, sizes, sample counts and seeds -- travel through
one set of queues and kernels, and each image equals its standalone
dl_hip_render with the same pass sizes bit for bit.  The render service
packs queued jobs of a scene into such passes (continuous batching) and its
images stay bit-identical to standalone renders."""
import dataclasses

import numpy as np
import pytest

import drawlab_amd
from drawlab_amd import scenes
from drawlab_amd.scene import Camera

pytestmark = pytest.mark.gpu


def _cams():
    base = scenes.cornell_box(64, 48).camera
    return [
        base,
        dataclasses.replace(base, width=40, height=40, eye=(0.3, 0.3, -0.7)),
        scenes.opencv_camera((0.278, 0.273, -0.8), (0.278, 0.273, 0.0), (0, 1, 0), 39.3, 32, 24),
        Camera(type="perspective", width=48, height=32, fov=45.0,
               to_world=np.array([[-1, 0, 0, 0.278], [0, 1, 0, 0.273], [0, 0, 1, -0.8], [0, 0, 0, 1]], np.float32)),
    ]


def _standalone(backend, desc, cam, spp, seed, k, normals=False):
    backend.set_camera(cam)
    img, st = backend.render(spp=spp, seed=seed, spp_per_pass=k, normals=normals)
    return img, st


def test_shared_pass_images_equal_standalone_renders(backend):
    d = scenes.cornell_box(64, 48, spp=4, max_depth=6, rr_depth=3)
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    cams = _cams()
    spps, seeds = [4, 7, 2, 5], [3, 11, 0, 9]
    items = [{"camera": c, "spp": s, "seed": e, "film": i} for i, (c, s, e) in enumerate(zip(cams, spps, seeds))]
    imgs, st = backend.render_batch(items)
    print("batch", st["paths"], st["ms_total"], flush=True)
    assert st["paths"] == sum(c.width * c.height * s for c, s in zip(cams, spps))
    for c, s, e, img in zip(cams, spps, seeds, imgs):
        ref, _ = _standalone(backend, d, c, s, e, s)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (c.type, c.width, s, e)
        assert img.mean() > 0.01


def test_shared_passes_accumulate_like_render_passes(backend):
    """Two passes of 4 samples into a film slot equal render(spp=8,
    spp_per_pass=4); the other slot's image in the same passes is untouched
    by it."""
    d = scenes.mis_plates(64, 36, max_depth=5)
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    a, b = d.camera, dataclasses.replace(d.camera, width=32, height=18)
    backend.render_batch([{"camera": a, "spp": 4, "seed": 2, "film": 5, "image": False},
                          {"camera": b, "spp": 3, "seed": 1, "film": 0}])
    imgs, _ = backend.render_batch([{"camera": a, "spp": 4, "first_sample": 4, "seed": 2, "film": 5,
                                     "accumulate": True},
                                    {"camera": b, "spp": 6, "seed": 8, "film": 1}])
    ref, _ = _standalone(backend, d, a, 8, 2, 4)
    assert np.array_equal(imgs[0].view(np.uint32), ref.view(np.uint32))
    ref_b, _ = _standalone(backend, d, b, 6, 8, 6)
    assert np.array_equal(imgs[1].view(np.uint32), ref_b.view(np.uint32))


def test_shared_pass_normals_and_hidden_lights(backend):
    d = scenes.cornell_box(48, 48, spp=4, max_depth=5, rr_depth=3)
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    cams = _cams()[:2]
    imgs, st = backend.render_batch([{"camera": c, "spp": 2, "seed": i, "film": i} for i, c in enumerate(cams)],
                                    normals=True)
    assert st["shadow_rays"] == 0
    for i, (c, img) in enumerate(zip(cams, imgs)):
        ref, _ = _standalone(backend, d, c, 2, i, 2, normals=True)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    backend.set_view_options(hide_lights=True)
    imgs, _ = backend.render_batch([{"camera": c, "spp": 3, "seed": i, "film": i} for i, c in enumerate(cams)])
    for i, (c, img) in enumerate(zip(cams, imgs)):
        ref, _ = _standalone(backend, d, c, 3, i, 3)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    backend.reset_view_options()


def test_shared_pass_errors(backend):
    d = scenes.cornell_box(32, 32, spp=4, max_depth=5, rr_depth=3)
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    c = d.camera
    with pytest.raises(drawlab_amd.DlError, match="share a film slot"):
        backend.render_batch([{"camera": c, "spp": 1, "film": 3}, {"camera": c, "spp": 1, "film": 3}])
    backend.render_batch([{"camera": c, "spp": 1, "film": 2}])
    with pytest.raises(drawlab_amd.DlError, match="another size"):
        backend.render_batch([{"camera": dataclasses.replace(c, width=16), "spp": 1, "film": 2,
                               "accumulate": True}])
    with pytest.raises(drawlab_amd.DlError, match="film slot out of range"):
        backend.render_batch([{"camera": c, "spp": 1, "film": drawlab_amd.BATCH_MAX}])
    with pytest.raises(drawlab_amd.DlError, match="pass too large"):
        backend.render_batch([{"camera": dataclasses.replace(c, width=4096, height=4096), "spp": 9, "film": 0}])


def test_service_packs_jobs_into_shared_passes(backend):
    """Many small jobs of one scene on one lane: they join the running
    group (stats), and every image is bit-identical to its standalone
    render, including jobs of several passes and a second scene's jobs."""
    cbox = scenes.cornell_box(48, 48, spp=8, max_depth=6, rr_depth=3)
    plates = scenes.mis_plates(48, 27, max_depth=5)
    jobs = [(cbox, 4 + i % 5, i, 0, 48 - 4 * (i % 3)) for i in range(10)]
    jobs += [(plates, 6, 50, 2, 48), (cbox, 9, 60, 4, 40), (plates, 3, 70, 0, 32)]
    svc = drawlab_amd.Service(lanes=1, policy=drawlab_amd.POLICY_FIFO)
    try:
        sid = {id(cbox): svc.add_scene(cbox), id(plates): svc.add_scene(plates)}
        ids = []
        for d, spp, seed, k, w in jobs:
            cam = dataclasses.replace(d.camera, width=w, height=d.camera.height * w // 48)
            ids.append(svc.submit(sid[id(d)], spp, seed=seed, max_depth=d.max_depth, rr_depth=d.rr_depth,
                                  spp_per_pass=k, camera=cam))
        infos = [svc.wait(j) for j in ids]
        assert all(i["state"] == drawlab_amd.JOB_DONE for i in infos), infos
        imgs = [svc.result(j) for j in ids]
        st = svc.stats()
        print("stats", st, flush=True)
        assert st["batched_jobs"] >= 5 and st["shared_passes"] >= 1
        assert st["estimate_misses"] == 0 and st["peak_reserved_bytes"] <= st["budget_bytes"]
    finally:
        svc.stop()
    for (d, spp, seed, k, w), img, info in zip(jobs, imgs, infos):
        cam = dataclasses.replace(d.camera, width=w, height=d.camera.height * w // 48)
        backend.upload_scene(dataclasses.replace(d, camera=cam))
        backend.build_accel(drawlab_amd.BUILD_GPU)
        ref, rst = backend.render(spp=spp, seed=seed, spp_per_pass=k)
        assert info["paths"] == rst["paths"]
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (spp, seed, k, w)


def test_pass_split_does_not_change_the_image(backend):
    """The film adds each pixel's samples in sample order whatever the pass
    size (the per-path pcg32 streams depend only on pixel and sample), so a
    plain render is bit-identical over any split into passes."""
    d = scenes.cornell_box(48, 40, spp=9, max_depth=6, rr_depth=3)
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    ref, _ = backend.render(spp=9, seed=4, spp_per_pass=9)
    for k in (4, 1, 0):
        img, _ = backend.render(spp=9, seed=4, spp_per_pass=k)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), k


def _resident(backend, descs, builders=None):
    """Upload and keep each scene under keys 1, 2, ...; the last stays current."""
    for k, d in enumerate(descs, 1):
        backend.upload_scene(d)
        backend.build_accel(builders[k - 1] if builders else drawlab_amd.BUILD_GPU)
        backend.scene_keep(k)


def _reference(backend, key, cam, spp, seed, max_depth, rr_depth, k=None):
    assert backend.scene_select(key)
    backend.set_camera(cam)
    img, st = backend.render(spp=spp, seed=seed, spp_per_pass=spp if k is None else k, max_depth=max_depth,
                             rr_depth=rr_depth)
    return img


def _cross_scene_set():
    return [scenes.cornell_box(48, 40, spp=4, max_depth=8, rr_depth=3),
            scenes.material_zoo(40, 30),
            scenes.textured_box(32, 32),
            scenes.displaced_sphere(48, 27, spp=4, max_depth=8, grid=48, glass_level=2, env_size=(64, 32)),
            scenes.mis_plates(48, 27, max_depth=6)]


def test_pass_over_several_scenes_equals_each_scenes_own_render(backend):
    """Images of five resident scenes -- area, point, several and
    environment lights, textures and normal maps, every BSDF type, two
    reconstruction filters -- share one pass (their scenes in one layout,
    each path carrying its scene's record), and each image equals the
    render of its scene alone bit for bit.  Two items of one scene and an
    item naming the current scene by key 0 are in the same pass."""
    descs = _cross_scene_set()
    # BVHs of all three builders side by side in the shared layout
    _resident(backend, descs, [drawlab_amd.BUILD_GPU, drawlab_amd.BUILD_HOST_SAH, drawlab_amd.BUILD_GPU_LBVH,
                               drawlab_amd.BUILD_GPU, drawlab_amd.BUILD_HOST_SAH])
    md, rr = 7, 3
    keys = [1, 2, 3, 4, 0, 1]
    cams = [descs[0].camera, descs[1].camera, descs[2].camera, descs[3].camera, descs[4].camera,
            dataclasses.replace(descs[0].camera, width=24, height=20, eye=(0.3, 0.3, -0.7))]
    spps, seeds = [3, 2, 4, 2, 3, 5], [1, 2, 3, 4, 5, 6]
    items = [{"camera": c, "spp": s, "seed": e, "film": i, "scene": k}
             for i, (c, s, e, k) in enumerate(zip(cams, spps, seeds, keys))]
    imgs, st = backend.render_batch(items, max_depth=md, rr_depth=rr)
    print("cross-scene batch", st["paths"], st["ms_total"], flush=True)
    assert st["paths"] == sum(c.width * c.height * s for c, s in zip(cams, spps))
    for key, c, s, e, img in zip(keys, cams, spps, seeds, imgs):
        ref = _reference(backend, key or 5, c, s, e, md, rr)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (key, c.width, s, e)
        assert img.mean() > 0.005, key


def test_cross_scene_layout_follows_resident_scenes(backend):
    """The layout holding several scenes is reused while its scenes stay,
    and rebuilt when one of them changes: a scene replaced under its key
    renders as the new scene, a dropped scene is refused, and a pass with
    counted traversal, normals or hidden lights over several scenes still
    equals the standalone renders."""
    descs = _cross_scene_set()[:3]
    _resident(backend, descs)
    md, rr = 6, 3
    item = lambda k, i, s=2, e=0: {"camera": descs[k - 1].camera, "spp": s, "seed": e, "film": i, "scene": k}
    # one resident scene that is not current: its own layout, no union
    one, _ = backend.render_batch([item(1, 0, 2, 3)], max_depth=md, rr_depth=rr)
    ref = _reference(backend, 1, descs[0].camera, 2, 3, md, rr)
    assert np.array_equal(one[0].view(np.uint32), ref.view(np.uint32))
    imgs, _ = backend.render_batch([item(1, 0), item(2, 1)], max_depth=md, rr_depth=rr)
    imgs2, st = backend.render_batch([item(1, 0), item(2, 1)], max_depth=md, rr_depth=rr, count_traversal=True)
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(imgs, imgs2))
    assert st["nodes_closest"] > 0 and st["nodes_shadow"] > 0
    # replace scene 2 by the MIS plates under the same key
    plates = scenes.mis_plates(40, 24, max_depth=6)
    backend.upload_scene(plates)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    backend.scene_keep(2)
    descs[1] = plates
    imgs, _ = backend.render_batch([item(1, 0, 3, 7), item(2, 1, 3, 8), item(3, 2, 2, 9)], max_depth=md, rr_depth=rr)
    for k, img, s, e in ((1, imgs[0], 3, 7), (2, imgs[1], 3, 8), (3, imgs[2], 2, 9)):
        ref = _reference(backend, k, descs[k - 1].camera, s, e, md, rr)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), k
    imgs, _ = backend.render_batch([item(1, 0), item(3, 1)], max_depth=md, rr_depth=rr, normals=True)
    for k, img in ((1, imgs[0]), (3, imgs[1])):
        assert backend.scene_select(k)
        backend.set_camera(descs[k - 1].camera)
        ref, _ = backend.render(spp=2, seed=0, spp_per_pass=2, max_depth=md, rr_depth=rr, normals=True)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), k
    for k in (1, 2, 3):
        assert backend.scene_select(k)
        backend.set_view_options(hide_lights=True)
    imgs, _ = backend.render_batch([item(1, 0, 2, 4), item(2, 1, 2, 5)], max_depth=md, rr_depth=rr)
    for k, img, e in ((1, imgs[0], 4), (2, imgs[1], 5)):
        ref = _reference(backend, k, descs[k - 1].camera, 2, e, md, rr)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), k
    assert backend.scene_select(3)
    backend.reset_view_options()
    with pytest.raises(drawlab_amd.DlError, match="differ in hide_lights or epsilon"):
        backend.render_batch([item(1, 0), item(3, 1)], max_depth=md, rr_depth=rr)
    backend.scene_drop(1)
    with pytest.raises(drawlab_amd.DlError, match="no resident scene with this key"):
        backend.render_batch([item(1, 0), item(3, 1)], max_depth=md, rr_depth=rr)
    backend.scene_drop(0)


@pytest.mark.parametrize("single", [False, True])
def test_service_shares_passes_across_resident_scenes(backend, single):
    """Once a lane holds three scenes, queued jobs of all three join one
    group and share passes (cross_scene_passes), unless
    single_scene_passes; every image stays bit-identical to its
    standalone render and no lane exceeds its reservation."""
    descs = [scenes.cornell_box(40, 32, spp=4, max_depth=6, rr_depth=3), scenes.mis_plates(40, 24, max_depth=6),
             scenes.material_zoo(32, 24, max_depth=6)]
    svc = drawlab_amd.Service(lanes=1, policy=drawlab_amd.POLICY_FIFO, single_scene_passes=single)
    jobs, ids = [], []
    try:
        sids = [svc.add_scene(d) for d in descs]
        # warm-up: one job per scene makes all three resident on the lane
        for i, (d, s) in enumerate(zip(descs, sids)):
            jobs.append((i, 2, 100 + i))
            ids.append(svc.submit(s, 2, seed=100 + i, max_depth=6, rr_depth=3))
        for j in ids:
            svc.wait(j)
        for r in range(3):
            for i, s in enumerate(sids):
                spp, seed = 3 + (i + r) % 3, 10 * r + i
                jobs.append((i, spp, seed))
                ids.append(svc.submit(s, spp, seed=seed, max_depth=6, rr_depth=3, spp_per_pass=2))
        infos = [svc.wait(j) for j in ids]
        assert all(i["state"] == drawlab_amd.JOB_DONE for i in infos), infos
        imgs = [svc.result(j) for j in ids]
        st = svc.stats()
        print("stats", single, st, flush=True)
        if single:
            assert st["cross_scene_passes"] == 0
        else:
            assert st["cross_scene_passes"] >= 1 and st["shared_passes"] >= 1
        assert st["scene_uploads"] == 3
        assert st["estimate_misses"] == 0 and st["peak_reserved_bytes"] <= st["budget_bytes"]
    finally:
        svc.stop()
    for (i, spp, seed), img in zip(jobs, imgs):
        backend.upload_scene(descs[i])
        backend.build_accel(drawlab_amd.BUILD_GPU)
        ref, _ = backend.render(spp=spp, seed=seed, spp_per_pass=2 if seed < 100 else 0, max_depth=6, rr_depth=3)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (i, spp, seed)
