"""Device-buffer plumbing for the GPU tests (torch as allocator only)."""
from __future__ import annotations

import ctypes as C

import numpy as np

import lyapunov3d_amd as lyap


def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


def empty(n_bytes: int):
    torch = torch_cuda()
    return torch.empty(max(int(n_bytes), 4), dtype=torch.uint8, device="cuda")


def device_lights(lights, n):
    torch = torch_cuda()
    raw = bytes(C.string_at(C.addressof(lights), C.sizeof(lyap.LyapLight) * max(n, 1)))
    t = torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda()
    torch.cuda.synchronize()
    return t


def to_numpy(t, dtype, shape=None):
    a = t.cpu().numpy().view(dtype)
    return a.reshape(shape) if shape is not None else a


class RenderBuffers:
    def __init__(self, w: int, h: int):
        self.w, self.h = w, h
        n = w * h
        self.rgba = empty(4 * n)
        self.points = empty(36 * n)
        self.depth = empty(4 * n)
        self.lam = empty(4 * n)
        self.evals = empty(4 * n)

    def render(self, ctx: lyap.Context, cam, prm, seq, lights_dev, n_lights, mode):
        ctx.calc_render(self.rgba.data_ptr(), self.points.data_ptr(), cam, prm, seq,
                        lights_dev.data_ptr(), n_lights, self.w, self.h, self.depth.data_ptr(),
                        self.lam.data_ptr(), self.evals.data_ptr(), mode)

    def fetch(self):
        h, w = self.h, self.w
        return {
            "rgba": to_numpy(self.rgba, np.uint8, (h, w, 4)),
            "points": to_numpy(self.points, np.float32, (h, w, 9)),
            "depth": to_numpy(self.depth, np.float32, (h, w)),
            "lambda": to_numpy(self.lam, np.float32, (h, w)),
            "evals": to_numpy(self.evals, np.uint32, (h, w)),
        }
