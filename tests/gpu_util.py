"""Helpers for GPU tests: device buffers through the lux ABI itself.

PyTorch is not used here: it carries its own HIP runtime, and a second
runtime in the same process cannot open the device after the library's.
"""
import numpy as np

from bevy_penumbra_amd.lux import DeviceBuffer


def trace_gpu(ctx, rays: np.ndarray, bary: bool = False):
    n = rays.shape[0]
    r = DeviceBuffer.from_array(ctx, np.ascontiguousarray(rays, np.float32))
    h = DeviceBuffer.empty(ctx, (n, 4), np.uint32)
    b = DeviceBuffer.empty(ctx, (n, 2), np.float32) if bary else None
    ctx.trace_device(r, n, h, b)
    ctx.synchronize()
    out = h.numpy()
    return (out, b.numpy()) if bary else out
