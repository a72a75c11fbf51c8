"""Fused AdamW vs torch.optim.AdamW (CPU, float64 reference)."""
import pytest
import torch

from neuralangelo_amd.optim import AdamW

pytestmark = pytest.mark.gpu


def test_adamw_matches_torch(cuda):
    g = torch.Generator().manual_seed(0)
    p0 = torch.randn(100003, generator=g)
    ref = p0.clone().requires_grad_(True)  # torch's own float32 AdamW on the CPU
    opt_ref = torch.optim.AdamW([ref], lr=1e-3, weight_decay=1e-2)
    p = torch.nn.Parameter(p0.to(cuda))
    opt = AdamW([p], lr=1e-3, weight_decay=1e-2)
    for it in range(5):
        grad = torch.randn(100003, generator=g) * (0.1 if it % 2 else 10.0)
        grad[::7] = 0.0
        ref.grad = grad.clone()
        opt_ref.step()
        p.grad = grad.to(cuda)
        opt.step()
    # same update formula; only float32 rounding of the op sequence differs (a few
    # ulps of O(1) parameters); a formula error would show at the lr scale (1e-4)
    err = (p.detach().cpu() - ref.detach()).abs().max().item()
    assert err < 2e-6, err


def test_adamw_refreshes_bf16_table(cuda):
    from neuralangelo_amd.encoding import Encoding
    enc = Encoding(3, dict(otype="HashGrid", n_levels=4, n_features_per_level=2,
                           log2_hashmap_size=12, base_resolution=8, per_level_scale=1.5),
                   dtype=torch.bfloat16).to(cuda)
    x = torch.rand(500, 3, device=cuda)
    enc(x).float().sum().backward()
    opt = AdamW(enc.parameters(), lr=1e-2)
    opt.set_lowp_shadow(enc.params, enc)
    opt.step()
    shadow = enc.table_for_forward()
    assert torch.equal(shadow, enc.params.detach().to(torch.bfloat16))
    # a later in-place edit of the master still invalidates the copy
    with torch.no_grad():
        enc.params.add_(1.0)
    assert torch.equal(enc.table_for_forward(), enc.params.detach().to(torch.bfloat16))


def test_adamw_multi_tensor_equals_single(cuda):
    """40 tensors of ragged sizes (one empty) through the optimizer's batched launches
    (32 + 8 tensors) against na_adamw_step per tensor: bit-identical parameters and
    moments over three steps."""
    from neuralangelo_amd import _lib
    g = torch.Generator().manual_seed(1)
    sizes = [1, 3, 4, 5, 7, 64, 1023, 1024, 1025, 0, 100003] + [257 * (i + 1) for i in range(29)]
    params = [torch.nn.Parameter(torch.randn(n, generator=g).to(cuda)) for n in sizes]
    single = [p.detach().clone() for p in params]
    m1 = [torch.zeros_like(p) for p in single]
    v1 = [torch.zeros_like(p) for p in single]
    opt = AdamW(params, lr=1e-3, weight_decay=1e-2)
    for step in range(1, 4):
        grads = [torch.randn(n, generator=g).to(cuda) * (10.0 if step % 2 else 0.1) for n in sizes]
        for p, gr in zip(params, grads):
            p.grad = gr.clone()
        opt.step()
        for p, gr, m, v in zip(single, grads, m1, v1):
            _lib.call("na_adamw_step", _lib.ptr(p), _lib.ptr(gr), _lib.ptr(m), _lib.ptr(v), p.numel(),
                      1e-3, 0.9, 0.999, 1e-8, 1e-2, step, None, _lib.NA_BF16, 0,
                      _lib.stream_of(p))
    for p, s, m, v in zip(params, single, m1, v1):
        st = opt.state[p]
        assert torch.equal(p.detach(), s)
        assert torch.equal(st["exp_avg"], m) and torch.equal(st["exp_avg_sq"], v)


def test_table_gradient_cleared_in_the_update_pass(cuda):
    """The table gradient is zeroed by the AdamW pass and handed back to the encoder,
    which gives the same (zero) buffer to the next backward: no fill, same results as
    fresh zero-filled gradients (bit for bit, with the sorted table-gradient
    reduction), and no aliasing while .grad is still attached."""
    from neuralangelo_amd.encoding import Encoding
    cfg = dict(otype="HashGrid", n_levels=8, n_features_per_level=4, log2_hashmap_size=14,
               base_resolution=8, per_level_scale=1.5)
    torch.manual_seed(0)
    runs = []
    for fused in (True, False):
        enc = Encoding(3, cfg, dtype=torch.bfloat16).to(cuda)
        enc.deterministic = True
        with torch.no_grad():
            enc.params.copy_(torch.linspace(-1e-2, 1e-2, enc.n_params, device=cuda))
        opt = AdamW(enc.parameters(), lr=1e-2)
        opt.set_lowp_shadow(enc.params, enc)
        g = torch.Generator().manual_seed(1)
        ptrs = []
        for step in range(3):
            x = torch.rand(4000, 3, generator=g).to(cuda)
            opt.zero_grad(set_to_none=True)
            if not fused:
                enc._spare_grad = None  # the reference run: a fresh zero fill every step
            enc(x).float().square().sum().backward()
            ptrs.append(enc.params.grad.data_ptr())
            opt.step()
            if fused:
                assert enc.params.grad.abs().max().item() == 0.0  # cleared in the pass
        runs.append((enc.params.detach().clone(), ptrs))
    (p_fused, ptr_fused), (p_ref, _) = runs
    assert torch.equal(p_fused, p_ref)
    assert ptr_fused[1] == ptr_fused[0] and ptr_fused[2] == ptr_fused[0]  # the buffer is reused
    # with the gradient still attached (no zero_grad), a fresh buffer is used and the
    # attached one accumulates as autograd defines it
    enc = Encoding(3, cfg).to(cuda)
    opt = AdamW(enc.parameters(), lr=1e-2)
    opt.set_lowp_shadow(enc.params, enc)
    x = torch.rand(1000, 3, device=cuda)
    enc(x).square().sum().backward()
    opt.step()
    enc(x).square().sum().backward()
    once = torch.zeros_like(enc.params)
    enc.backward_into(x, 2 * enc(x).detach(), once)
    assert torch.allclose(enc.params.grad, once, rtol=1e-5, atol=1e-7)
