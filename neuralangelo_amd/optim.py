"""Fused AdamW on the GPU (torch.optim.AdamW semantics, one kernel pass per tensor).

Replaces the optimizer the reference builds at
imaginaire/trainers/utils/get_trainer.py:106-150 (AdamW, lr 1e-3, wd 1e-2,
projects/neuralangelo/configs/base.yaml:117-121).  Parameters registered with
``set_lowp_shadow`` (the hash table) get their bf16 copy refreshed in the same
pass, and their gradient zeroed in it too: the holder keeps the zeroed buffer for
the next backward, so the step's zero_grad (imaginaire/trainers/base.py:447-457,
with set_to_none) costs no fill of the 1.46 GB table gradient.  The parameter's
.grad still holds that buffer until zero_grad; it reads as zero after step().
"""
import ctypes

import torch

from . import _lib


class AdamW(torch.optim.Optimizer):

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=1e-2):
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay))
        self._shadows = {}

    def set_lowp_shadow(self, param, holder):
        """holder.lowp_shadow() -> (bf16 tensor) / holder.mark_shadow_fresh(): kept in sync."""
        self._shadows[id(param)] = holder

    @torch.no_grad()
    def step(self, closure=None):
        loss = closure() if closure is not None else None
        for group in self.param_groups:
            b1, b2 = group["betas"]
            hyper = (float(group["lr"]), float(b1), float(b2), float(group["eps"]),
                     float(group["weight_decay"]))
            batch = {}  # (step, device) -> [(p, g, state)]: tensors without a bf16 copy
            for p in group["params"]:
                if p.grad is None:
                    continue
                _lib.require_cuda(p)
                if p.dtype != torch.float32 or not p.is_contiguous():
                    raise RuntimeError("fused AdamW expects contiguous float32 parameters")
                g = p.grad
                if g.dtype != torch.float32 or not g.is_contiguous():
                    g = g.float().contiguous()
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                st["step"] += 1
                holder = self._shadows.get(id(p))
                if holder is None:
                    batch.setdefault((int(st["step"]), p.device), []).append((p, g, st))
                    continue
                lowp = holder.lowp_shadow()
                # the gradient is zeroed in the same pass when it is the parameter's own
                # buffer; the holder keeps it for the next backward (no separate fill)
                clear = g is p.grad and hasattr(holder, "take_cleared_grad")
                _lib.call("na_adamw_step", _lib.ptr(p), _lib.ptr(g), _lib.ptr(st["exp_avg"]),
                          _lib.ptr(st["exp_avg_sq"]), p.numel(), *hyper, int(st["step"]),
                          _lib.ptr(lowp), _lib.NA_BF16, int(clear), _lib.stream_of(p))
                if clear:
                    holder.take_cleared_grad(g)
                # the kernel wrote through a raw pointer: bump the version counter so
                # caches keyed on it (the encoder's bf16 copy) see the update
                torch.autograd.graph.increment_version(p)
                holder.mark_shadow_fresh()
            for (step, _), items in batch.items():
                self._step_multi(items, hyper, step)
        return loss

    @staticmethod
    def _step_multi(items, hyper, step):
        """One launch per NA_ADAMW_MULTI_MAX tensors (na_adamw_step_multi)."""
        for k in range(0, len(items), _lib.NA_ADAMW_MULTI_MAX):
            chunk = items[k:k + _lib.NA_ADAMW_MULTI_MAX]
            arr = (_lib.AdamwTensor * len(chunk))()
            for e, (p, g, st) in zip(arr, chunk):
                e.param, e.grad = p.data_ptr(), g.data_ptr()
                e.exp_avg, e.exp_avg_sq = st["exp_avg"].data_ptr(), st["exp_avg_sq"].data_ptr()
                e.n = p.numel()
            _lib.call("na_adamw_step_multi", arr, len(chunk), *hyper, step,
                      _lib.stream_of(chunk[0][0]))
            for p, _, _ in chunk:
                torch.autograd.graph.increment_version(p)
