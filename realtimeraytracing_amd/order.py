"""Learned child ordering: log -> train -> apply, on device buffers.

A thin holder of the model's device state (parameters and Adam moments in
torch tensors) over the rtq_order_* entry points; every step runs in librtq's
HIP kernels.
"""
from __future__ import annotations

from .rtq import ORDER_PARAMS, ORDER_SAMPLE, RTQ_TRACE_ANY, Context, Scene, TrainParams


class OrderModel:
    def __init__(self, ctx: Context, seed: int = 1):
        import torch
        self.ctx = ctx
        self.params = torch.zeros(ORDER_PARAMS, dtype=torch.float32, device="cuda")
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        ctx.order_init(seed, self.params.data_ptr())

    def log(self, scene: Scene, rays, n: int, any_hit: bool, stride: int = 1, capacity: int = 1 << 20):
        """Samples from every `stride`-th ray of a device ray buffer: (tensor [k, ORDER_SAMPLE], k).

        The kernel appends samples in the order its atomics hand out slots;
        they are returned sorted by a key computed from their contents, so the
        same rays give the same sample order and training is reproducible."""
        import torch
        buf = torch.empty((capacity, ORDER_SAMPLE), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        k = scene.order_log(rays.data_ptr(), n, buf.data_ptr(), capacity, RTQ_TRACE_ANY if any_hit else 0, stride)
        out = buf[:k]
        if k > 1:
            g = torch.Generator(device="cpu").manual_seed(0x5EED)
            w = torch.rand(ORDER_SAMPLE, generator=g, dtype=torch.float64).to("cuda") + 0.5
            key = (out.double() * w).sum(dim=1)
            out = out[torch.argsort(key, stable=True)]
        return out.contiguous(), k

    def train(self, samples, epochs: int = 4, batch: int = 1024, lr: float = 2e-3, seed: int = 7):
        """Adam epochs over a [k, ORDER_SAMPLE] device tensor; returns the per-epoch mean losses."""
        import torch
        torch.cuda.synchronize()
        tp = TrainParams(epochs, batch, lr, 0.9, 0.999, 0, seed)
        self.step, losses = self.ctx.order_train(samples.data_ptr(), len(samples), tp, self.params.data_ptr(),
                                                 self.m.data_ptr(), self.v.data_ptr(), self.step)
        return losses

    def apply(self, scene: Scene, confidence_any: float = 0.5, confidence_closest: float = 1.0) -> None:
        """Order words for the scene.  By default only any-hit rays use the
        learned order: for closest-hit rays the nearer-child order is as good
        and keeps lanes more coherent."""
        import torch
        torch.cuda.synchronize()
        scene.order_apply(self.params.data_ptr(), confidence_any, confidence_closest)

    def save(self, path) -> None:
        import torch
        torch.cuda.synchronize()
        self.ctx.order_save(path, self.params.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self.step)

    def load(self, path) -> None:
        self.step = self.ctx.order_load(path, self.params.data_ptr(), self.m.data_ptr(), self.v.data_ptr())
