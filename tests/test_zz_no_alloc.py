"""dl_hip_trace_closest_noalloc: the closest-hit entry point whose
traversal scratch is a static per-device array instead of buffers of the
context.  Through the C ABI, its hits equal dl_hip_trace_closest's bit for
bit on two calls in a row (the second on a smaller batch, so stale scratch
from the first call must not matter), the context's device memory does not
change across the calls, and a batch beyond its limit is refused before any
kernel runs."""
import ctypes

import numpy as np
import pytest

import drawlab_amd
import oracle
from drawlab_amd import scenes
from helpers import ray_batches

pytestmark = pytest.mark.gpu


def _trace(be, fn, rays):
    n = rays.shape[1]
    d_rays = be.alloc(rays.nbytes)
    d_rays.upload(np.ascontiguousarray(rays))
    bufs = [be.alloc(4 * max(1, n)) for _ in range(4)]
    for b in bufs:
        b.upload(np.full(n, 0xA5A5A5A5, np.uint32))
    be._check(fn(be.h, d_rays.ptr, n, *(b.ptr for b in bufs)))
    out = (bufs[0].download(np.uint32, n), bufs[1].download(np.float32, n), bufs[2].download(np.float32, n),
           bufs[3].download(np.float32, n))
    for b in bufs + [d_rays]:
        b.free()
    return out


def _same(a, b):
    return all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))


def test_noalloc_closest_hits_equal_the_allocating_entry_point(backend):
    L = backend.L
    desc = scenes.displaced_sphere(width=64, height=36, grid=256, glass_level=3, env_size=(64, 32))
    backend.upload_scene(desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    osc = oracle.from_desc(desc)
    big = np.concatenate([r for k, r in ray_batches(desc, osc, 40000, seed=3).items() if k != "shadow"], axis=1)
    small = np.concatenate([r[:, :3000] for k, r in ray_batches(desc, osc, 3000, seed=4).items()], axis=1)
    for name, rays in (("first call, %d rays" % big.shape[1], big), ("second call, %d rays" % small.shape[1], small)):
        print(name, flush=True)
        ref = _trace(backend, L.dl_hip_trace_closest, rays)
        cur0, _ = backend.memory_usage()
        got = _trace(backend, L.dl_hip_trace_closest_noalloc, rays)
        cur1, _ = backend.memory_usage()
        assert _same(got, ref), name
        assert (ref[0] != 0xFFFFFFFF).mean() > 0.2
        assert cur1 == cur0


def test_noalloc_refuses_oversized_batches(backend):
    d_buf = backend.alloc(64)
    rc = backend.L.dl_hip_trace_closest_noalloc(backend.h, d_buf.ptr, (1 << 31) + 1, d_buf.ptr, d_buf.ptr, d_buf.ptr,
                                                d_buf.ptr)
    assert rc == 1  # DL_ERR_INVALID
    assert b"2^31" in backend.L.dl_hip_last_error(backend.h)
    d_buf.free()
