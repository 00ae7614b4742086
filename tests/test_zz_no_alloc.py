"""lux_trace_rays_device_noalloc against lux_trace_rays_device.

The allocation-free entry point keeps the traversal's ray cursor in a
library-owned device array instead of context memory.  Two calls in a row,
the second over fewer rays, must each give the hits (and barycentrics) of the
allocating entry point bit for bit; more than 2^31 rays are refused.
"""
import numpy as np
import pytest

from bevy_penumbra_amd import lux, scenes
from bevy_penumbra_amd.lux import DeviceBuffer

pytestmark = pytest.mark.gpu


def _trace(ctx, rays, noalloc):
    n = rays.shape[0]
    r = DeviceBuffer.from_array(ctx, np.ascontiguousarray(rays, np.float32))
    h = DeviceBuffer.empty(ctx, (n, 4), np.uint32)
    b = DeviceBuffer.empty(ctx, (n, 2), np.float32)
    if noalloc:
        ctx.trace_device_noalloc(r, n, h, b)
    else:
        ctx.trace_device(r, n, h, b)
    ctx.synchronize()
    return h.numpy(), b.numpy()


def test_gpu_noalloc_trace_equals_trace_device():
    sc, _ = scenes.colonnade()
    with lux.Context(0) as ctx:
        ctx.upload_scene(sc)
        ctx.build_bvh(lux.LUX_BVH_SAH)
        for n, seed in ((300_000, 3), (40_000, 4)):                 # the second call is smaller
            rays = scenes.microbench_rays(sc.bounds, n=n, seed=seed)
            want_h, want_b = _trace(ctx, rays, noalloc=False)
            got_h, got_b = _trace(ctx, rays, noalloc=True)
            assert np.array_equal(got_h, want_h), n
            assert np.array_equal(got_b.view(np.uint32), want_b.view(np.uint32)), n
            assert np.any(got_h[:, 1] != lux.LUX_EMPTY)


def test_gpu_noalloc_trace_refuses_too_many_rays():
    sc = scenes.traversal_fixture()
    with lux.Context(0) as ctx:
        ctx.upload_scene(sc)
        ctx.build_bvh()
        r = DeviceBuffer.empty(ctx, (4, 8), np.float32)
        h = DeviceBuffer.empty(ctx, (4, 4), np.uint32)
        with pytest.raises(lux.LuxError, match="2\\^31"):
            ctx.trace_device_noalloc(r, (1 << 31) + 1, h)
