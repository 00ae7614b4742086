"""Helpers for GPU tests: device buffers via torch (plumbing only)."""
import numpy as np


def dev(a: np.ndarray):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    torch.cuda.synchronize()
    return t


def empty_u32(*shape):
    import torch
    return torch.zeros(*shape, dtype=torch.int32, device="cuda:0")


def host(t) -> np.ndarray:
    import torch
    torch.cuda.synchronize()
    a = t.cpu().numpy()
    return a.view(np.uint32) if a.dtype == np.int32 else a


def trace_gpu(ctx, rays: np.ndarray, bary: bool = False):
    import torch
    n = rays.shape[0]
    r = dev(rays)
    h = empty_u32(n, 4)
    b = torch.zeros(n, 2, dtype=torch.float32, device="cuda:0") if bary else None
    ctx.trace_device(r, n, h, b)
    ctx.synchronize()
    out = host(h)
    return (out, b.cpu().numpy()) if bary else out
