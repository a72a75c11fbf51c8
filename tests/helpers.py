"""Shared test helpers: copy a product model's parameters into oracle form."""
import torch

from oracle import field as ofield
from oracle import hashgrid as hg


def oracle_params(model, dtype=torch.float64):
    sdf, rgb = model.neural_sdf, model.neural_rgb
    t = {}
    l0, l1 = sdf.mlp.linears
    t["W0v"], t["W0g"], t["b0"] = l0.weight_v, l0.weight_g, l0.bias
    t["W1v"], t["W1g"], t["b1"] = l1.weight_v, l1.weight_g, l1.bias
    t["ws"], t["bs"] = sdf.mlp.linear_sdf.weight, sdf.mlp.linear_sdf.bias
    for i, lin in enumerate(rgb.mlp.linears):
        t[f"R{i}v"], t[f"R{i}g"], t[f"R{i}b"] = lin.weight_v, lin.weight_g, lin.bias
    return ofield.FieldParams({k: v.detach().cpu().to(dtype).requires_grad_(True)
                               for k, v in t.items()})


def oracle_field(model, P, dtype=torch.float64, table_dtype=None):
    enc = model.neural_sdf.tcnn_encoding
    lt = hg.level_table(enc.n_levels, enc.n_features_per_level, enc.log2_hashmap_size,
                        enc.base_resolution, enc.per_level_scale)
    table = enc.params.detach().cpu().to(dtype).requires_grad_(True)
    return ofield.SDFField(lt, table, P, vol_range=enc.vol_range,
                           active_levels=model.neural_sdf.active_levels, dtype=dtype,
                           table_dtype=table_dtype), table


def randomize_model(model, seed=0, table_scale=0.05):
    """Non-degenerate parameters: random table and W0 encoder block (the geometric
    init zeroes it, which would hide encoder gradients)."""
    g = torch.Generator().manual_seed(seed)
    with torch.no_grad():
        enc = model.neural_sdf.tcnn_encoding
        enc.params.copy_((torch.rand(enc.n_params, generator=g) * 2 - 1) * table_scale)
        l0 = model.neural_sdf.mlp.linears[0]
        v = l0.weight_v.detach().cpu().clone()
        v[:, 3:] = torch.randn(v.shape[0], v.shape[1] - 3, generator=g) * 0.05
        l0.weight_v.copy_(v)
        l0.weight_g.copy_(v.norm(dim=1, keepdim=True))
        for lin in model.neural_sdf.mlp.linears:
            lin.bias.copy_(torch.randn(lin.bias.shape, generator=g) * 0.01)


def sphere_cameras(n, H=32, W=32, fov_deg=50.0, radius=3.0, seed=0):
    """Look-at cameras uniform on a sphere of `radius` aimed at the origin, in the
    reference's world-to-camera convention (x right, y down, z forward):
    pose [n,3,4] = [R | t] with t = -R c, intr [n,3,3] pinhole with the principal
    point at the image centre and focal length from the horizontal field of view."""
    import math
    g = torch.Generator().manual_seed(seed)
    f = 0.5 * W / math.tan(math.radians(fov_deg) / 2)
    poses, intrs = [], []
    for _ in range(n):
        c = torch.randn(3, generator=g, dtype=torch.float64)
        c = c / c.norm() * radius
        fwd = -c / c.norm()
        right = torch.linalg.cross(fwd, torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64))
        right = right / right.norm()
        down = torch.linalg.cross(fwd, right)
        R = torch.stack([right, down, fwd], dim=0)   # rows: camera axes in world coords
        poses.append(torch.cat([R, (-R @ c)[:, None]], dim=1))
        intrs.append(torch.tensor([[f, 0.0, W / 2], [0.0, f, H / 2], [0.0, 0.0, 1.0]],
                                  dtype=torch.float64))
    return torch.stack(poses).float(), torch.stack(intrs).float()


def hessian_floor(cdt, table_scale, err_ref32):
    """Absolute part of the tr H error bound (the relative part is 3 % of |tr H|).
    BF16 at the configuration's own table scale (1e-4): 0.2, from the increments'
    precision -- hi/lo parts keep 2^-17 of |D| ~ eps |grad s|, which divided by
    eps^2 (7.9e-8) is ~0.05 per tap.  Otherwise twice the 99th-percentile error of
    the reference's own float32 arithmetic at the same points (the FP32 oracle):
    FP32 is that arithmetic in another order, and on a rough field (table 0.05)
    the bf16 rounding of the table and weights (2^-9 relative) is of that size."""
    if cdt == torch.bfloat16 and table_scale <= 1e-3:
        return 0.2
    return 2.0 * err_ref32.quantile(0.99).item()
