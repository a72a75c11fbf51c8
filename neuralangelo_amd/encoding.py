"""Hash-grid encoding module, a drop-in for ``tcnn.Encoding`` on gfx950.

Reference call sites: the module is constructed at
projects/neuralangelo/utils/modules.py:42-50 as ``tcnn.Encoding(3, config)``
with ``config = dict(otype="HashGrid", n_levels, n_features_per_level,
log2_hashmap_size, base_resolution, per_level_scale)`` and called at
modules.py:84-86 on ``[P,3]`` float32 points normalised to the unit cube.
Only ``params`` (one flat parameter, tcnn layout) enters the state dict, so
checkpoints keyed ``neural_sdf.tcnn_encoding.params`` load unchanged.

Storage precision: ``params`` is always the FP32 master.  With
``dtype=torch.bfloat16`` the forward pass gathers from a BF16 copy of the
table (refreshed whenever the master changes) and returns BF16 features; the
backward pass always scatters FP32 gradients into ``params.grad``.
"""
import ctypes
import math

import torch

from . import _lib


def grid_layout(n_levels, n_features, log2_hashmap_size, base_resolution, per_level_scale):
    """Host level table (C ABI ``na_hashgrid_layout``).  Returns (desc, n_params)."""
    lib = _lib.load()
    desc = _lib.GridDesc()
    n_params = lib.na_hashgrid_layout(int(n_levels), int(n_features), int(log2_hashmap_size),
                                      int(base_resolution), float(per_level_scale),
                                      ctypes.byref(desc))
    _lib.check(n_params, "na_hashgrid_layout")
    return desc, int(n_params)


def _parse_config(encoding_config):
    otype = encoding_config.get("otype", "HashGrid")
    if otype not in ("HashGrid", "Grid"):
        raise NotImplementedError(f"encoding otype {otype!r} is not supported (HashGrid only)")
    if otype == "Grid" and encoding_config.get("type", "Hash") != "Hash":
        raise NotImplementedError("only the Hash grid type is supported")
    interp = encoding_config.get("interpolation", "Linear")
    if interp != "Linear":
        raise NotImplementedError(f"interpolation {interp!r} is not supported (Linear only)")
    n_levels = int(encoding_config.get("n_levels", 16))
    n_features = int(encoding_config.get("n_features_per_level", 2))
    log2_T = int(encoding_config.get("log2_hashmap_size", 19))
    base_res = int(encoding_config.get("base_resolution", 16))
    per_level_scale = float(encoding_config.get("per_level_scale", 2.0))
    return n_levels, n_features, log2_T, base_res, per_level_scale


class _HashGridFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, params, enc):
        out = enc._forward_raw(x, params)
        ctx.save_for_backward(x)
        ctx.enc = enc
        return out

    @staticmethod
    def backward(ctx, dy):
        (x,) = ctx.saved_tensors
        enc = ctx.enc
        if ctx.needs_input_grad[0]:
            raise NotImplementedError(
                "gradient w.r.t. the encoder input (analytical SDF gradients) is not "
                "implemented; use numerical gradients")
        grad_params = None
        if ctx.needs_input_grad[1]:
            grad_params = enc.table_grad_buffer()
            enc.backward_into(x, dy, grad_params)
        return None, grad_params, None


class Encoding(torch.nn.Module):
    """Drop-in for ``tinycudann.Encoding(n_input_dims, encoding_config, seed, dtype)``.

    forward(x): x [P, 3] float32 in [0,1]^3 -> [P, n_levels * n_features_per_level].
    """

    def __init__(self, n_input_dims, encoding_config, seed=1337, dtype=None):
        super().__init__()
        if n_input_dims != 3:
            raise NotImplementedError("only 3-D hash grids are supported")
        self.n_input_dims = 3
        self.encoding_config = dict(encoding_config)
        (self.n_levels, self.n_features_per_level, self.log2_hashmap_size,
         self.base_resolution, self.per_level_scale) = _parse_config(encoding_config)
        self.desc, self.n_params = grid_layout(self.n_levels, self.n_features_per_level,
                                               self.log2_hashmap_size, self.base_resolution,
                                               self.per_level_scale)
        self.n_output_dims = self.n_levels * self.n_features_per_level
        self.seed = seed
        self.dtype = torch.float32 if dtype is None else dtype
        if self.dtype not in (torch.float32, torch.bfloat16):
            raise NotImplementedError(f"table dtype {self.dtype} not supported (float32, bfloat16)")
        # tcnn initialises the grid uniformly in [-1e-4, 1e-4].
        gen = torch.Generator().manual_seed(int(seed))
        init = torch.rand(self.n_params, generator=gen, dtype=torch.float32) * 2e-4 - 1e-4
        self.params = torch.nn.Parameter(init)
        self.active_levels = self.n_levels
        # table gradient by the sorted reduction (bitwise repeatable) instead of
        # float atomics; torch.use_deterministic_algorithms(True) selects it too
        self.deterministic = False
        self._lowp_cache = None  # (data_ptr, version, tensor)

    # -- level table accessors -------------------------------------------------
    @property
    def resolutions(self):
        return [int(self.desc.resolution[i]) for i in range(self.n_levels)]

    @property
    def offsets(self):
        return [int(self.desc.offset[i]) for i in range(self.n_levels + 1)]

    @property
    def scales(self):
        return [float(self.desc.scale[i]) for i in range(self.n_levels)]

    def extra_repr(self):
        return (f"levels={self.n_levels}, F={self.n_features_per_level}, "
                f"log2_T={self.log2_hashmap_size}, base={self.base_resolution}, "
                f"scale={self.per_level_scale:.6f}, params={self.n_params}, dtype={self.dtype}")

    # -- table views ----------------------------------------------------------
    def table_for_forward(self, params=None):
        """Table tensor the gather reads: the FP32 master or its BF16 copy."""
        p = self.params if params is None else params
        if self.dtype == torch.float32:
            return p.detach()
        key = (p.data_ptr(), p._version)
        if self._lowp_cache is None or self._lowp_cache[0] != key:
            self._lowp_cache = (key, p.detach().to(self.dtype))
        return self._lowp_cache[1]

    def lowp_shadow(self):
        """The bf16 gather copy (allocated on first use), or None for an fp32 table."""
        if self.dtype == torch.float32:
            return None
        p = self.params
        if self._lowp_cache is None or self._lowp_cache[1].device != p.device:
            self._lowp_cache = ((p.data_ptr(), p._version), p.detach().to(self.dtype))
        return self._lowp_cache[1]

    def table_grad_buffer(self):
        """A zeroed fp32 gradient buffer of the table for a backward pass.  When the
        optimizer cleared the previous gradient in its update pass (na_adamw_step with
        clear_grad) and handed it back, that buffer is reused -- it is zero already, so
        no fill runs; else a new zero-filled buffer.  The spare is taken (not kept), so
        autograd can adopt it as params.grad without a copy; it is only reused while
        params.grad is None (a gradient still attached to the parameter is never
        aliased)."""
        spare, self._spare_grad = getattr(self, "_spare_grad", None), None
        p = self.params
        if (spare is not None and p.grad is None and spare.device == p.device
                and spare.numel() == self.n_params):
            return spare
        return torch.zeros(self.n_params, dtype=torch.float32, device=p.device)

    def sorted_grad(self):
        """True when the table gradient takes the sorted reduction: per table entry the
        contributions are summed in a fixed (point, evaluation, corner) order, so the
        gradient is bitwise repeatable (``deterministic = True`` on this encoder, or
        torch.use_deterministic_algorithms(True)).  The default scatter uses float
        atomics, as tcnn's backward does."""
        return bool(self.deterministic) or torch.are_deterministic_algorithms_enabled()

    def sorted_workspace(self, n, evals, device):
        """Workspace (uint8) of the sorted reduction for n points of `evals` (1 or 5)
        evaluations each."""
        nbytes = _lib.call("na_hashgrid_sorted_workspace", ctypes.byref(self.desc), int(n),
                           int(evals))
        return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)

    def take_cleared_grad(self, grad):
        """Keep `grad` (zeroed by the optimizer's update pass) for the next backward."""
        self._spare_grad = grad

    def mark_shadow_fresh(self):
        """Declare the bf16 copy up to date with the master (written by the optimizer)."""
        if self._lowp_cache is not None:
            p = self.params
            self._lowp_cache = ((p.data_ptr(), p._version), self._lowp_cache[1])

    def _forward_raw(self, x, params):
        _lib.require_cuda(x, params)
        if x.dim() != 2 or x.shape[1] != 3:
            raise ValueError(f"expected input of shape [P, 3], got {tuple(x.shape)}")
        x = x.contiguous().float()
        n = x.shape[0]
        table = self.table_for_forward(params)
        out = torch.empty(n, self.n_output_dims, dtype=self.dtype, device=x.device)
        _lib.call("na_hashgrid_fwd", ctypes.byref(self.desc), _lib.ptr(x), n, _lib.ptr(table),
                  _lib.dtype_code(table.dtype), _lib.ptr(out), _lib.dtype_code(out.dtype),
                  int(self.active_levels), _lib.stream_of(x))
        return out

    def backward_into(self, x, dy, grad_params):
        """Accumulate d(loss)/d(params) for upstream gradient dy [P, L*F] into grad_params."""
        x = x.contiguous().float()
        dy = dy.contiguous()
        if dy.dtype not in (torch.float32, torch.bfloat16):
            dy = dy.float()
        if self.sorted_grad():
            ws = self.sorted_workspace(x.shape[0], 1, x.device)
            _lib.call("na_hashgrid_bwd_sorted", ctypes.byref(self.desc), _lib.ptr(x), x.shape[0],
                      _lib.ptr(dy), _lib.dtype_code(dy.dtype), _lib.ptr(grad_params),
                      int(self.active_levels), _lib.ptr(ws), ws.numel(), _lib.stream_of(x))
            return
        _lib.call("na_hashgrid_bwd", ctypes.byref(self.desc), _lib.ptr(x), x.shape[0],
                  _lib.ptr(dy), _lib.dtype_code(dy.dtype), _lib.ptr(grad_params),
                  int(self.active_levels), _lib.stream_of(x))

    def corners(self, x):
        """Per (point, level, corner) table entry index and trilinear weight (parity probe)."""
        _lib.require_cuda(x)
        x = x.contiguous().float()
        n = x.shape[0]
        idx = torch.empty(n, self.n_levels, 8, dtype=torch.int32, device=x.device)
        w = torch.empty(n, self.n_levels, 8, dtype=torch.float32, device=x.device)
        _lib.call("na_hashgrid_corners", ctypes.byref(self.desc), _lib.ptr(x), n, _lib.ptr(idx),
                  _lib.ptr(w), _lib.stream_of(x))
        return idx, w

    def forward(self, x):
        return _HashGridFunction.apply(x, self.params, self)


def growth_rate(min_logres, max_logres, num_levels):
    """Per-level scale exactly as projects/neuralangelo/utils/modules.py:38-41 (float64)."""
    import numpy as np
    r_min, r_max = 2 ** min_logres, 2 ** max_logres
    return float(np.exp((np.log(r_max) - np.log(r_min)) / (num_levels - 1)))


def reference_resolutions(min_logres, num_levels, rate):
    """``NeuralSDF.resolutions`` (modules.py:51-54): floor(2^min_logres * rate^l) + 1."""
    r_min = 2 ** min_logres
    return [int(math.floor(r_min * rate ** lv)) + 1 for lv in range(num_levels)]
