"""Python handle over libaurora's learned denoiser (aurora_denoiser_*).

Inputs are device tensors holding per-pixel sums over `spp` samples, as
Scene.render_device writes them (H x W x 3 float32, or B x H x W x 3 for
training); outputs are linear RGB means.
"""
import ctypes as C

import numpy as np

from . import abi

# (A slot, B slot, out slot, relu) and slot channels / levels, as the
# library's network (parameter layout: per layer W[Cout][9][Cin] then b)
SLOT_C = [16, 32, 32, 32, 48, 48, 48, 64, 64, 64, 64, 64, 64, 64, 48, 48, 32, 32, 16]
SLOT_L = [0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0, 0]
LAYERS = [(0, -1, 1, 1), (1, -1, 2, 1), (3, -1, 4, 1), (4, -1, 5, 1), (6, -1, 7, 1), (7, -1, 8, 1),
          (9, -1, 10, 1), (10, -1, 11, 1), (8, 11, 12, 1), (12, -1, 13, 1), (5, 13, 14, 1), (14, -1, 15, 1),
          (2, 15, 16, 1), (16, -1, 17, 1), (17, -1, 18, 0)]
POOLS = [(2, 3), (5, 6), (8, 9)]


def layer_shapes():
    """[(cout, cin)] per layer."""
    out = []
    for a, b, o, _ in LAYERS:
        out.append((SLOT_C[o], SLOT_C[a] + (SLOT_C[b] if b >= 0 else 0)))
    return out


class Denoiser:
    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        ctx.check(self.lib.aurora_denoiser_create(ctx.h, seed, C.byref(h)))
        self.h = h
        self.n = int(self.lib.aurora_denoiser_param_count(h))
        ctx.adopt(self)

    def close(self):
        if self.h:
            self.lib.aurora_denoiser_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.n, np.float32)
        self.ctx.check(self.lib.aurora_denoiser_get_params(self.h, p.ctypes.data, self.n))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n
        self.ctx.check(self.lib.aurora_denoiser_set_params(self.h, p.ctypes.data, self.n))

    def save(self, path):
        self.get_params().tofile(path)

    def load(self, path):
        self.set_params(np.fromfile(path, np.float32))

    def denoise(self, rgb_sum, alb_sum, nrm_sum, spp, out=None):
        import torch
        H, W = rgb_sum.shape[0], rgb_sum.shape[1]
        if out is None:
            out = torch.empty((H, W, 3), dtype=torch.float32, device=rgb_sum.device)
        self.ctx.check(self.lib.aurora_denoise(self.h, rgb_sum.data_ptr(), alb_sum.data_ptr(), nrm_sum.data_ptr(),
                                               spp, W, H, out.data_ptr()))
        self.ctx.synchronize()
        return out

    def set_loss(self, kind):
        """"log" (L2 in log(1 + radiance)) or "rel" (relative L2, the error
        metric's form)."""
        code = {"log": abi.LOSS_LOG_L2, "rel": abi.LOSS_REL_L2}[kind]
        self.ctx.check(self.lib.aurora_denoiser_set_loss(self.h, code))

    def loss_grad(self, rgb, alb, nrm, target, spp, grads=True):
        B, H, W = rgb.shape[0], rgb.shape[1], rgb.shape[2]
        loss = C.c_float(0)
        g = np.zeros(self.n, np.float32) if grads else None
        self.ctx.check(self.lib.aurora_denoiser_loss_grad(self.h, rgb.data_ptr(), alb.data_ptr(), nrm.data_ptr(),
                                                          target.data_ptr(), B, spp, W, H, C.byref(loss),
                                                          g.ctypes.data if grads else None))
        return loss.value, g

    def train_step(self, rgb, alb, nrm, target, spp, lr):
        B, H, W = rgb.shape[0], rgb.shape[1], rgb.shape[2]
        loss = C.c_float(0)
        self.ctx.check(self.lib.aurora_denoiser_train_step(self.h, rgb.data_ptr(), alb.data_ptr(), nrm.data_ptr(),
                                                           target.data_ptr(), B, spp, W, H, lr, C.byref(loss)))
        return loss.value


def mfma_selftest(ctx, A_bits, B_bits):
    D = np.zeros((16, 16), np.float32)
    A = np.ascontiguousarray(A_bits, np.uint16)
    Bm = np.ascontiguousarray(B_bits, np.uint16)
    ctx.check(ctx.lib.aurora_selftest_mfma(ctx.h, A.ctypes.data, Bm.ctypes.data, D.ctypes.data))
    return D


def torch_reference(params, rgb_sum, alb_sum, nrm_sum, spp, emulate_bf16=True, dtype=None):
    """The same network in PyTorch (CPU): a plain reference for the HIP
    kernels' numerics.  Inputs B x H x W x 3 tensors; returns (out16 in NCHW,
    noisy_log NCHW).  With emulate_bf16 the weights, the network input and
    every stored activation are rounded to bfloat16 as on the GPU."""
    import torch
    import torch.nn.functional as F
    dtype = dtype or torch.float64

    def rb(t):
        return t.to(torch.bfloat16).to(dtype) if emulate_bf16 else t

    inv = 1.0 / spp
    r = rgb_sum.to(dtype) * inv
    noisy_log = torch.log1p(torch.clamp(r, min=0.0)).permute(0, 3, 1, 2)
    x = torch.cat([noisy_log, (alb_sum.to(dtype) * inv).permute(0, 3, 1, 2),
                   (nrm_sum.to(dtype) * inv).permute(0, 3, 1, 2)], dim=1)
    x = torch.cat([x, torch.zeros(x.shape[0], 7, x.shape[2], x.shape[3], dtype=dtype)], dim=1)
    acts = [None] * 19
    acts[0] = rb(x)
    p = torch.as_tensor(params).to(dtype)
    off = 0
    for (a, b, o, relu), (cout, cin) in zip(LAYERS, layer_shapes()):
        for src, dst in POOLS:
            if dst == a:
                acts[dst] = rb(F.avg_pool2d(acts[src], 2))
        w = p[off:off + cout * 9 * cin].reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
        off += cout * 9 * cin
        bias = p[off:off + cout]
        off += cout
        inp = acts[a]
        if b >= 0:
            inp = torch.cat([inp, F.interpolate(acts[b], scale_factor=2, mode="nearest")], dim=1)
        y = F.conv2d(inp, rb(w), bias, padding=1)
        if relu:
            y = rb(F.relu(y))
        acts[o] = y
    return acts[18], noisy_log
