"""Benchmark: Neuralangelo train step (fwd+bwd+AdamW) on one MI355X.

Workload (the 2^22 x 16 x 8 train-step configuration): hash grid 16 levels,
2^22 entries/level, 8 features, resolutions 2^5..2^11; SDF MLP 1x256 softplus
(geometric init), colour MLP 4x256 ReLU; 8192 rays from 8 cameras on a sphere
of radius 3 (32x32 px, FOV 50 deg) x (64 coarse + 2 x 16 resampled) samples;
4-tap numerical gradient; loss L1*3 + 0.1 eikonal + 5e-4 curvature; AdamW
(lr 1e-3, wd 1e-2) over all parameters.  Synthetic inputs, random targets.

Prints one JSON line.  value = train rays/s of the BF16 path; beside it the same
step with every kernel in FP32, and the same step written in stock torch ops
(FP32, no kernel of this package).  roofline = the step's dominant kernel, the
hash-grid backward over the centre point and its 4 taps (HBM/atomic-bound
scatter), timed with HIP events on torch's stream around its launch inside the
timed steps: achieved = its compulsory bytes per launch (each point's centre
corner gradients read and written once, its six dL/dy rows and the point)
over that time; traffic = its HBM bytes per launch from the committed rocprofv3
counter summary; distinct_entry_bytes = the bytes of every table entry the
launch touches, read and written once, counted on the step's own points.
error_vs_fp64 = SDF, gradient, Hessian-trace and colour error of both precisions
against the FP64 oracle on rays of the same workload.  hash_fwd = the encoder
forward kernel (hash GB/s); render = the eval render of one 32 x 32 image;
mc = marching cubes on the 1024^3 lattice.
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

import numpy as np  # noqa: E402
import torch  # noqa: E402
import torch.nn.functional as torch_F  # noqa: E402

HBM_PEAK_GBS = 8000.0


def camera_rays(n_cams, H=32, W=32, fov_deg=50.0, radius=3.0, seed=0):
    """Pixel-centre rays of look-at cameras uniform on a sphere (camera.py:283-311 convention)."""
    g = torch.Generator().manual_seed(seed)
    centers, dirs = [], []
    f = 0.5 * W / math.tan(math.radians(fov_deg) / 2)
    ys, xs = torch.meshgrid(torch.arange(H) + 0.5, torch.arange(W) + 0.5, indexing="ij")
    cam_dirs = torch.stack([(xs - W / 2) / f, (ys - H / 2) / f, torch.ones_like(xs)], -1).view(-1, 3)
    for _ in range(n_cams):
        c = torch.randn(3, generator=g)
        c = c / c.norm() * radius
        fwd = -c / c.norm()
        up = torch.tensor([0.0, 0.0, 1.0])
        right = torch.linalg.cross(fwd, up)
        right = right / right.norm()
        down = torch.linalg.cross(fwd, right)
        Rcw = torch.stack([right, down, fwd], dim=1)  # camera -> world
        d = cam_dirs @ Rcw.t()
        dirs.append(d / d.norm(dim=-1, keepdim=True))
        centers.append(c.expand(H * W, 3))
    return torch.cat(centers).contiguous(), torch.cat(dirs).contiguous()


def build_model(dtype, dev, log2_T=22, F=8):
    from neuralangelo_amd.model import Model, default_config
    torch.manual_seed(0)
    m = Model(default_config(log2_T=log2_T, F=F), compute_dtype=dtype).to(dev).train()
    m.progress = 1.0
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    return m


def train_step(model, opt, center, ray_unit, target):
    from neuralangelo_amd import misc
    out = model.render_rays(center, ray_unit, stratified=True)
    render = torch.nn.functional.l1_loss(out["rgb"], target) * 3
    eik, curv = misc.sdf_losses(out["gradients"], out["hessians"], out["outside"])
    loss = render + 0.1 * eik + 5e-4 * curv
    opt.zero_grad(set_to_none=True)
    loss.backward()
    opt.step()
    return loss


def time_train(dtype, dev, n_rays, steps, warmup):
    from neuralangelo_amd.optim import AdamW
    model = build_model(dtype, dev)
    opt = AdamW(model.parameters(), lr=1e-3, weight_decay=1e-2)
    enc = model.neural_sdf.tcnn_encoding
    opt.set_lowp_shadow(enc.params, enc)
    n_cams = max(1, n_rays // 1024)
    c, v = camera_rays(n_cams)
    c, v = c[:n_rays].to(dev), v[:n_rays].to(dev)
    target = torch.rand(n_rays, 3, generator=torch.Generator().manual_seed(1)).to(dev)
    from neuralangelo_amd import functional as Fn
    for k in range(warmup):
        if k == warmup - 1:
            Fn.SCATTER_CAPTURE = []
        loss = train_step(model, opt, c, v, target)
    captured, Fn.SCATTER_CAPTURE = Fn.SCATTER_CAPTURE, None
    torch.cuda.synchronize()
    Fn.LAUNCH_EVENTS = {ROOF_CALL: []}
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = train_step(model, opt, c, v, target)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    pairs, Fn.LAUNCH_EVENTS = Fn.LAUNCH_EVENTS[ROOF_CALL], None
    launch_ms = sum(a.elapsed_time(b) for a, b in pairs) / max(1, len(pairs))
    assert torch.isfinite(loss).item(), "non-finite loss"
    return model, dt, float(loss.item()), launch_ms, len(pairs), captured


ROOF_CALL = "na_sdf_taps_bwd"
ROOF_KERNEL = "hashgrid_taps_bwd_kernel"
TRAFFIC_FILE = os.path.join(REPO, "profiles", "hashgrid_taps_bwd_train_step_pmc.txt")


def look_at_cameras(n, H=32, W=32, fov_deg=50.0, radius=3.0, seed=0):
    """World-to-camera poses [n,3,4] ([R | -R c], rows = camera x right, y down, z
    forward) and pinhole intrinsics [n,3,3] of cameras on a sphere aimed at the origin."""
    g = torch.Generator().manual_seed(seed)
    f = 0.5 * W / math.tan(math.radians(fov_deg) / 2)
    poses, intrs = [], []
    for _ in range(n):
        c = torch.randn(3, generator=g, dtype=torch.float64)
        c = c / c.norm() * radius
        fwd = -c / c.norm()
        right = torch.linalg.cross(fwd, torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64))
        right = right / right.norm()
        rot = torch.stack([right, torch.linalg.cross(fwd, right), fwd], dim=0)
        poses.append(torch.cat([rot, (-rot @ c)[:, None]], dim=1))
        intrs.append(torch.tensor([[f, 0.0, W / 2], [0.0, f, H / 2], [0.0, 0.0, 1.0]],
                                  dtype=torch.float64))
    return torch.stack(poses).float(), torch.stack(intrs).float()


def render_leg(dev, dtype, iters=10, warmup=3):
    """Eval-mode render of one 32 x 32 image (1024 rays; camera at radius 3, FOV 50 deg)
    through Model.inference on the 2^19 x 16 x 4 model (64 + 2 x 16 samples per ray,
    rgb, opacity, normal and depth maps); HIP events around each call."""
    from neuralangelo_amd.model import Model, default_config
    torch.manual_seed(0)
    size = [32, 32]
    m = Model(default_config(log2_T=19, F=4), compute_dtype=dtype,
              cfg_data=dict(train=dict(image_size=size), val=dict(image_size=size))).to(dev)
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    m.cfg_render["rand_rays"] = 1024
    pose, intr = look_at_cameras(1, seed=7)
    data = dict(pose=pose.to(dev), intr=intr.to(dev), idx=torch.zeros(1, dtype=torch.long, device=dev))
    for _ in range(warmup):
        m.inference(data)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = []
    for _ in range(iters):
        e0.record()
        m.inference(data)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    t = float(np.median(ms))
    del m
    torch.cuda.empty_cache()
    return dict(rays_per_s=round(1024 / (t * 1e-3), 1), ms=round(t, 3), median_of=iters)


def measured_traffic():
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the factor the same
    counter reads for a copy of known size), or None when the summary is absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


TRAFFIC_TABLE = os.path.join(REPO, "profiles", "train_step_hbm_traffic_pmc.md")


def table_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter table (the
    copy-calibrated `total GB` column), or None."""
    try:
        with open(TRAFFIC_TABLE) as f:
            for line in f:
                cells = [c.strip() for c in line.split("|")]
                if len(cells) > 5 and cells[1] == f"`{kernel}`":
                    return float(cells[5]) * 1e9
    except (OSError, ValueError):
        return None
    return None


TAP_DIRS = ((1, -1, -1), (-1, -1, 1), (-1, 1, -1), (1, 1, 1))


@torch.no_grad()
def scatter_distinct_entries(model, captured):
    """Table entries the scatter launch touches: the distinct corners of the five
    evaluations (centre and four taps, as the kernel forms them) of every point of
    the captured launch, over all levels."""
    if not captured:
        return None
    pts, eps = captured[-1][:2]
    enc = model.neural_sdf.tcnn_encoding
    vmin, vmax = enc.vol_range
    total = 0
    for lv in range(enc.n_levels):
        idx = []
        for e in range(5):
            x = pts if e == 0 else pts + torch.tensor(TAP_DIRS[e - 1], device=pts.device,
                                                      dtype=pts.dtype) * eps
            ci, _ = enc.corners(((x - vmin) / (vmax - vmin)).contiguous())
            idx.append(ci[:, lv, :].reshape(-1))
        total += int(torch.unique(torch.cat(idx)).numel())
    return total


def taps_bwd_roofline(model, n_points, launch_ms, n_launches, dy_bytes, captured, rows=6):
    """Roofline of the dominant kernel of the step: the hash-grid backward over the
    centre point and its 4 taps (one launch per step, timed with HIP events on
    torch's stream inside the timed steps).  Compulsory bytes per point with the
    taps' corners taken as the centre's and the scatter fused: the centre's L*8*F
    fp32 corner gradients (4 B each, counted once), the `rows` dL/dy rows of L*F
    values in the compute dtype (centre, four taps and the taps' shared row of the
    increment formulation) and the 12-byte point.  Beside it: the bytes of every
    table entry the launch really touches, read and written once (distinct corners
    of the five evaluations, counted on the step's own points), and the counter
    traffic."""
    enc = model.neural_sdf.tcnn_encoding
    L, F = enc.n_levels, enc.n_features_per_level
    per_point = L * 8 * F * 4 + rows * L * F * dy_bytes + 12
    achieved = per_point * n_points / (launch_ms * 1e-3) / 1e9
    traffic = measured_traffic()
    distinct = scatter_distinct_entries(model, captured)
    rec = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
               frac=round(achieved / HBM_PEAK_GBS, 4),
               traffic=round(traffic) if traffic is not None else None,
               kernel=f"{ROOF_KERNEL} (centre + 4 taps, merged corner scatter)",
               bytes_per_point=per_point,
               bytes_model=f"compulsory: L*8*F*4 centre corner gradients + {rows} dL/dy rows "
                           f"of L*F x {dy_bytes} B + 12 B point",
               points_per_launch=n_points, avg_launch_ms=round(launch_ms, 4),
               launches_timed=n_launches,
               traffic_GBps=round(traffic / (launch_ms * 1e-3) / 1e9, 1) if traffic else None,
               traffic_source=os.path.relpath(TRAFFIC_FILE, REPO) if traffic is not None else None)
    if distinct is not None:
        db = distinct * F * 4 * 2 + n_points * (rows * L * F * dy_bytes + 12)
        rec.update(distinct_entries=distinct, distinct_entry_bytes=db,
                   distinct_entry_GBps=round(db / (launch_ms * 1e-3) / 1e9, 1),
                   distinct_entry_frac=round(db / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
    return rec


def encoder_roofline(model, dev, n_points=786432, iters=20):
    """HIP-event timing of the encoder forward (SDF-input kernel) at the centre-pass shape."""
    from neuralangelo_amd import functional as Fn
    enc = model.neural_sdf.tcnn_encoding
    g = torch.Generator().manual_seed(2)
    c, v = camera_rays(8, seed=3)
    R = n_points // 96
    t = torch.linspace(2.0, 4.0, 96)
    pts = (c[:R, None, :] + v[:R, None, :] * t[None, :, None]).reshape(-1, 3)
    pts = pts.clamp(-1.5, 1.5).contiguous().to(dev)
    cdt = model.compute_dtype
    k0p = Fn.sdf_input_width(enc.n_output_dims, cdt)
    for _ in range(3):
        Fn.sdf_input(enc, pts, cdt, k0p)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        Fn.sdf_input(enc, pts, cdt, k0p)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / iters
    L, F = enc.n_levels, enc.n_features_per_level
    bt = 2 if enc.dtype == torch.bfloat16 else 4
    bo = 2 if cdt == torch.bfloat16 else 4
    # per point: 8 corners x L levels x F features gathered, 3 floats in, one row out
    per_point = L * 8 * F * bt + 12 + k0p * bo
    achieved = per_point * n_points / (ms * 1e-3) / 1e9
    del g
    traffic = table_traffic("hashgrid_fwd_kernel")
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=round(traffic) if traffic is not None else None,
                kernel="hashgrid_fwd_kernel (SDF input)", bytes_per_point=per_point,
                points_per_launch=n_points, avg_launch_ms=round(ms, 4),
                traffic_source=os.path.relpath(TRAFFIC_TABLE, REPO) if traffic else None)


def mc_leg(model, dev, res=1024, block_res=128, iters=5, cpu=True):
    """Mesh extraction at res^3 over [-1,1)^3 in block_res^3 blocks (FP32 network SDF).

    The lattice SDF is evaluated once and stored; marching cubes (all blocks,
    one count + one emit launch chain) is timed on the stored volume with HIP
    events; voxels/s = res^3 / t_MC.  Roofline: the classify kernel reads
    4 B per lattice point (algorithmic), timed on its own."""
    from neuralangelo_amd import _lib, mcubes
    from neuralangelo_amd import mesh as gmesh
    sdf = model.neural_sdf
    intv = 2.0 / res
    axes = gmesh.lattice_axes(((-1.0, 1.0),) * 3, intv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vol = gmesh.lattice_sdf(lambda x: -sdf.sdf_nograd(x), axes, dev, max_points=1 << 23)
    torch.cuda.synchronize()
    t_sdf = time.perf_counter() - t0
    nx, ny, nz = vol.shape
    lib = _lib.load()
    need = _lib.check(lib.na_mc_workspace_size(nx, ny, nz, block_res), "na_mc_workspace_size")
    ws = torch.empty(int(need), dtype=torch.uint8, device=dev)
    axes_d = [a.to(dev) for a in axes]

    def run():
        return mcubes.lattice_marching_cubes(vol, block_res, 0.0, axes=axes_d, intv=intv,
                                             workspace=ws)
    V, F = run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = []
    for _ in range(iters):
        e0.record()
        V, F = run()
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    t_mc = float(np.median(ms)) * 1e-3
    # classify alone (the HBM-bound pass over the volume), HIP events on torch's stream
    totals = torch.empty(2, dtype=torch.int64, device=dev)
    s = _lib.stream_of(vol)
    c_ms = []
    for _ in range(iters):
        e0.record()
        _lib.call("na_mc_count", _lib.ptr(vol), nx, ny, nz, block_res, 0.0, _lib.ptr(ws),
                  int(ws.numel()), _lib.ptr(totals), s)
        e1.record()
        torch.cuda.synchronize()
        c_ms.append(e0.elapsed_time(e1))
    t_count = float(np.median(c_ms)) * 1e-3
    bytes_count = 4.0 * nx * ny * nz
    achieved = bytes_count / t_count / 1e9
    rec = dict(
        voxels_per_s=round(res ** 3 / t_mc, 1), ms=round(t_mc * 1e3, 3),
        lattice=[nx, ny, nz], block_res=block_res, vertices=int(V.shape[0]),
        faces=int(F.shape[0]), sdf_eval_s=round(t_sdf, 3),
        sdf_points_per_s=round(nx * ny * nz / t_sdf, 1),
        roofline=dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(achieved / HBM_PEAK_GBS, 4),
                      traffic=(round(table_traffic("mc_classify"))
                               if table_traffic("mc_classify") else None),
                      kernel="na_mc_count (classify + tile scan), 4 B per lattice point",
                      avg_launch_ms=round(t_count * 1e3, 4)))
    if cpu:
        from oracle import mesh as omesh
        x0 = max(0, nx // 2 - block_res // 2)
        sample = vol[x0:x0 + block_res + 1].cpu().numpy()
        t0 = time.perf_counter()
        omesh.lattice_marching_cubes(sample, block_res)
        dt = time.perf_counter() - t0
        rec["cpu_baseline"] = dict(
            value=round(sample.size / dt, 1), unit="voxels/s", cores=1, kind="port",
            sample=f"C restatement of the sequential algorithm (single thread, as PyMCubes) on "
                   f"the central {sample.shape[0]}x{ny}x{nz} slab of the same volume, "
                   f"{block_res}^3 blocks ({dt:.2f} s)")
    del vol, ws
    return rec


# ---------------------------------------------------------------------------
# The plain FP32 GPU path: the same train step in stock torch ops (no HIP kernels
# of this package): per-corner gathers for the hash encoding, weight-normed
# torch linear layers, the 4-tap difference quotient, NeuS alpha and cumprod
# compositing, torch.optim.AdamW.  It reads the model's own parameters.
# ---------------------------------------------------------------------------
_PY, _PZ = 2654435761, 805459861
_TAPS = ((1, -1, -1), (-1, -1, 1), (-1, 1, -1), (1, 1, 1))


def _sh_matrix(dev):
    """Degree-3 real SH as a [16 x 20] coefficient matrix over the monomials
    1, x, y, z, xy, yz, xz, x^2, y^2, z^2, x^2 y, y^3, y z^2, xyz, z^3, x^2 z, y^2 z,
    x z^2, x^3, x y^2 (normalisations from the closed forms)."""
    pi = math.pi
    c = lambda a, b: math.sqrt(a) / (b * math.sqrt(pi))  # noqa: E731
    M = torch.zeros(16, 20, dtype=torch.float64)
    M[0, 0] = c(1, 2)
    M[1, 2], M[2, 3], M[3, 1] = -c(3, 2), c(3, 2), -c(3, 2)
    M[4, 4], M[5, 5], M[7, 6] = c(15, 2), -c(15, 2), -c(15, 2)
    M[6, 9], M[6, 7], M[6, 8] = 2 * c(5, 4), -c(5, 4), -c(5, 4)
    M[8, 7], M[8, 8] = c(15, 4), -c(15, 4)
    k0 = -c(35 / 2, 4)
    M[9, 10], M[9, 11] = 3 * k0, -k0                   # y (3x^2 - y^2)
    M[10, 13] = c(105, 2)                              # xyz
    k2 = -c(21 / 2, 4)
    M[11, 12], M[11, 10], M[11, 11] = 4 * k2, -k2, -k2  # y (4z^2 - x^2 - y^2)
    k3 = c(7, 4)
    M[12, 14], M[12, 15], M[12, 16] = 2 * k3, -3 * k3, -3 * k3  # z (2z^2 - 3x^2 - 3y^2)
    k4 = -c(21 / 2, 4)
    M[13, 17], M[13, 18], M[13, 19] = 4 * k4, -k4, -k4  # x (4z^2 - x^2 - y^2)
    k5 = c(105, 4)
    M[14, 15], M[14, 16] = k5, -k5                     # z (x^2 - y^2)
    k6 = -c(35 / 2, 4)
    M[15, 18], M[15, 19] = k6, -3 * k6                 # x (x^2 - 3y^2)
    return M.float().to(dev)


def _sh3(d, M):
    x, y, z = d.unbind(-1)
    mono = torch.stack([torch.ones_like(x), x, y, z, x * y, y * z, x * z, x * x, y * y, z * z,
                        x * x * y, y * y * y, y * z * z, x * y * z, z * z * z, x * x * z,
                        y * y * z, x * z * z, x * x * x, x * y * y], dim=-1)
    return mono @ M.t()


class TorchStep:
    """The train step in stock torch ops (FP32): hash encoding by per-corner gathers,
    weight-normed linear layers, the 4-tap numerical gradient, NeuS alpha, cumprod
    compositing, L1 + eikonal + curvature losses, torch.optim.AdamW."""

    def __init__(self, model):
        enc = model.neural_sdf.tcnn_encoding
        self.enc = enc
        self.L, self.F = enc.n_levels, enc.n_features_per_level
        self.scale = enc.scales
        self.res = enc.resolutions
        self.off = enc.offsets
        self.T = 1 << enc.log2_hashmap_size
        self.vmin, self.vmax = enc.vol_range
        self.table = enc.params
        sdf, rgb = model.neural_sdf, model.neural_rgb
        self.sdf_lin = list(sdf.mlp.linears)
        self.head = sdf.mlp.linear_sdf
        self.rgb_lin = list(rgb.mlp.linears)
        self.s_var = model.s_var
        self.eps = model.neural_sdf.normal_eps / math.sqrt(3)
        self.M = _sh_matrix(self.table.device)
        self.corner = torch.tensor([[(c >> d) & 1 for d in range(3)] for c in range(8)],
                                   device=self.table.device)

    def encode(self, x):
        u = (x - self.vmin) / (self.vmax - self.vmin)
        outs = []
        for lv in range(self.L):
            pos = u * self.scale[lv] + 0.5
            fl = torch.floor(pos)
            t = pos - fl
            cell = fl.to(torch.int64)
            size = self.off[lv + 1] - self.off[lv]
            tab = self.table[self.off[lv] * self.F:self.off[lv + 1] * self.F].view(size, self.F)
            dense = self.res[lv] ** 3 <= size
            acc = 0
            for c in range(8):
                cc = cell + self.corner[c]
                if dense:
                    r = self.res[lv]
                    idx = (cc[:, 0] + cc[:, 1] * r + cc[:, 2] * r * r) % size
                else:
                    idx = ((cc[:, 0] ^ (cc[:, 1] * _PY) ^ (cc[:, 2] * _PZ)) & 0xFFFFFFFF) & (size - 1)
                w = torch.where(self.corner[c].bool(), t, 1 - t).prod(-1, keepdim=True)
                acc = acc + w * tab.index_select(0, idx)
            outs.append(acc)
        return torch.cat(outs, -1)

    @staticmethod
    def _wn(lin):
        return torch._weight_norm(lin.weight_v, lin.weight_g, 0)

    def sdf_mlp(self, x, with_feat):
        h = torch.cat([x, self.encode(x)], -1)
        l0, l1 = self.sdf_lin
        h0 = torch_F.softplus(torch_F.linear(h, self._wn(l0), l0.bias), beta=100)
        s = self.head(h0)
        feat = torch_F.softplus(torch_F.linear(h0, self._wn(l1), l1.bias), beta=100) if with_feat else None
        return s, feat

    def rgb(self, x, n, v, feat):
        h = torch.cat([x, _sh3(v, self.M), n, feat], -1)
        for i, lin in enumerate(self.rgb_lin):
            h = torch_F.linear(h, self._wn(lin), lin.bias)
            h = torch.relu(h) if i < len(self.rgb_lin) - 1 else torch.sigmoid(h)
        return h

    @torch.no_grad()
    def sample(self, c, v, near, far, rands, n_fine=16, n_hier=2):
        N = rands.shape[1]
        d = (rands + torch.arange(N, device=c.device)) / N * (far - near) + near
        s = self.sdf_mlp((c[:, None] + v[:, None] * d[..., None]).reshape(-1, 3), False)[0].view(d.shape)
        for hh in range(n_hier):
            inv_s = 64 * 2 ** hh
            mid = (s[:, :-1] + s[:, 1:]) * 0.5
            cos = (s[:, 1:] - s[:, :-1]) / (d[:, 1:] - d[:, :-1] + 1e-5)
            cos = torch.minimum(torch.cat([torch.zeros_like(cos[:, :1]), cos[:, :-1]], 1), cos)
            dd = d[:, 1:] - d[:, :-1]
            p = torch.sigmoid((mid - cos * dd * 0.5) * inv_s)
            q = torch.sigmoid((mid + cos * dd * 0.5) * inv_s)
            a = ((p - q) / (p + 1e-5)).clip(0, 1)
            w = a * torch.cumprod(torch.cat([torch.ones_like(a[:, :1]), 1 - a[:, :-1]], 1), 1)
            pdf = w / w.sum(-1, keepdim=True).clamp_min(1e-12)
            cdf = torch.cat([torch.zeros_like(pdf[:, :1]), torch.cumsum(pdf, -1)], -1)
            uu = torch.linspace(0, 1, n_fine + 1, device=c.device)
            uu = ((uu[:-1] + uu[1:]) / 2).expand(d.shape[0], n_fine).contiguous()
            idx = torch.searchsorted(cdf, uu, right=True)
            lo, hi = (idx - 1).clamp(min=0), idx.clamp(max=d.shape[1] - 1)
            cl, ch = cdf.gather(1, lo), cdf.gather(1, hi)
            bl, bh = d.gather(1, lo), d.gather(1, hi)
            fine = bl + (uu - cl) / (ch - cl + 1e-8) * (bh - bl)
            d, order = torch.sort(torch.cat([d, fine], 1), 1)
            if hh != n_hier - 1:
                sf = self.sdf_mlp((c[:, None] + v[:, None] * fine[..., None]).reshape(-1, 3),
                                  False)[0].view(fine.shape)
                s = torch.cat([s, sf], 1).gather(1, order)
        return d

    def loss(self, c, v, rands, target, progress=1.0, anneal_end=0.1):
        ctc, ctv = (c * c).sum(-1, keepdim=True), (c * v).sum(-1, keepdim=True)
        disc = ctv ** 2 - (ctc - 1)
        near, far = (-ctv - disc.sqrt()).relu(), -ctv + disc.sqrt()
        outside = near.isnan()
        near = torch.where(outside, torch.ones_like(near), near)
        far = torch.where(outside, torch.full_like(far, 1.2), far)
        d = self.sample(c, v, near, far, rands)
        R, N = d.shape
        x = (c[:, None] + v[:, None] * d[..., None]).reshape(-1, 3)
        s, feat = self.sdf_mlp(x, True)
        s = torch.where(outside.repeat_interleave(N, 0), torch.full_like(s, 1000.0), s)
        st = [self.sdf_mlp(x + self.eps * torch.tensor(k, device=x.device, dtype=x.dtype), False)[0]
              for k in _TAPS]
        kk = torch.tensor(_TAPS, device=x.device, dtype=x.dtype)
        grad = sum(kk[i] * st[i] for i in range(4)) / (4 * self.eps)
        hess = ((st[0] + st[1] + st[2] + st[3]) / 2 - 2 * s) / self.eps ** 2
        vs = v.repeat_interleave(N, 0)
        rgbs = self.rgb(x, torch_F.normalize(grad, dim=-1), vs, feat).view(R, N, 3)
        inv_s = self.s_var.exp()
        cos = (vs * grad).sum(-1).view(R, N)
        ar = min(progress / anneal_end, 1.0)
        it = -((-cos * 0.5 + 0.5).relu() * (1 - ar) + (-cos).relu() * ar)
        dd = torch.cat([d, far], 1)
        dd = dd[:, 1:] - dd[:, :-1]
        sv = s.view(R, N)
        p = torch.sigmoid((sv - it * dd * 0.5) * inv_s)
        q = torch.sigmoid((sv + it * dd * 0.5) * inv_s)
        alpha = ((p - q) / (p + 1e-5)).clip(0, 1)
        w = alpha * torch.cumprod(torch.cat([torch.ones_like(alpha[:, :1]), 1 - alpha[:, :-1]], 1), 1)
        rgb = (rgbs * w[..., None]).sum(1)
        inside = (~outside).repeat_interleave(N, 0).float()[:, 0]
        eik = (((grad.norm(dim=-1) - 1) ** 2).nan_to_num(0, 0, 0) * inside).mean()
        curv = (hess[:, 0].abs().nan_to_num(0, 0, 0) * inside).mean()
        return torch_F.l1_loss(rgb, target) * 3 + 0.1 * eik + 5e-4 * curv


def time_torch_fp32(dev, n_rays, steps, warmup):
    """Train rays/s of the stock-torch FP32 step on the same workload."""
    model = build_model(torch.float32, dev)
    step = TorchStep(model)
    opt = torch.optim.AdamW(model.parameters(), lr=1e-3, weight_decay=1e-2, foreach=True)
    n_cams = max(1, n_rays // 1024)
    c, v = camera_rays(n_cams)
    c, v = c[:n_rays].to(dev), v[:n_rays].to(dev)
    g = torch.Generator().manual_seed(1)
    target = torch.rand(n_rays, 3, generator=g).to(dev)

    def one():
        rands = torch.rand(n_rays, 64, device=dev)
        loss = step.loss(c, v, rands, target)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        return loss
    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = one()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    assert torch.isfinite(loss).item(), "non-finite loss (stock torch path)"
    del model, step, opt
    torch.cuda.empty_cache()
    return dt


@torch.no_grad()
def ragged_rays(n_rays, dev, seed=11, max_coarse=160, n_fine=32, per_unit=160.0):
    """A ragged batch of the packed configuration: rays from cameras on a sphere of
    radius 3 aimed (three in four) at a random union of 1-8 spheres with radii
    U(0.1, 0.5); a ray gets floor(per_unit * chord length) coarse samples (the chords
    of the spheres it crosses added up, at most max_coarse) plus n_fine resampled ones
    when it has at least 2: 0..192 samples, uniform and sorted inside its unit-sphere
    bounds.  Returns center, ray_unit [R,3], offsets int32 [R+1] and dists [S]."""
    g = torch.Generator().manual_seed(seed)
    k = int(torch.randint(1, 9, (1,), generator=g))
    sc = torch.rand(k, 3, generator=g) * 1.2 - 0.6
    sr = torch.rand(k, generator=g) * 0.4 + 0.1
    c = torch.randn(n_rays, 3, generator=g)
    c = c / c.norm(dim=-1, keepdim=True) * 3.0
    pick = torch.randint(0, k, (n_rays,), generator=g)
    jit = (torch.rand(n_rays, 3, generator=g) * 2 - 1) * sr[pick, None]
    tgt = torch.where(torch.rand(n_rays, 1, generator=g) < 0.75, sc[pick] + jit,
                      torch.rand(n_rays, 3, generator=g) * 1.6 - 0.8)
    v = tgt - c
    v = v / v.norm(dim=-1, keepdim=True)
    c, v, sc, sr = c.to(dev), v.to(dev), sc.to(dev), sr.to(dev)
    oc = c[:, None, :] - sc[None]                       # [R, k, 3]
    b = (oc * v[:, None, :]).sum(-1)
    disc = b * b - ((oc * oc).sum(-1) - sr[None] ** 2)
    chord = 2.0 * disc.clamp(min=0.0).sqrt()
    n_c = (chord.sum(-1) * per_unit).floor().clamp(max=max_coarse).long()
    counts = torch.where(n_c >= 2, n_c + n_fine, n_c)
    offsets = torch.zeros(n_rays + 1, dtype=torch.int32, device=dev)
    offsets[1:] = counts.cumsum(0).to(torch.int32)
    cv = (c * v).sum(-1)
    disc1 = (cv * cv - ((c * c).sum(-1) - 1.0)).clamp(min=0.0).sqrt()
    near, far = (-cv - disc1).clamp(min=0.0), -cv + disc1
    ridx = torch.repeat_interleave(torch.arange(n_rays, device=dev), counts)
    u = torch.rand(ridx.numel(), generator=g).to(dev)
    key = ridx.double() + u.double()                    # sorted within each ray
    u = u[torch.argsort(key)]
    d = near[ridx] + u * (far[ridx] - near[ridx])
    return c.contiguous(), v.contiguous(), offsets, d.contiguous()


def packed_leg(dev, n_rays=65536, steps=5, warmup=2):
    """Train rays/s of the ragged packed batch (BF16): packed render of every ray's own
    samples, L1 + eikonal + curvature, backward, AdamW; the batch is fixed and resident
    before the timed steps."""
    from neuralangelo_amd import functional as Fn
    from neuralangelo_amd import misc
    from neuralangelo_amd.optim import AdamW
    model = build_model(torch.bfloat16, dev)
    opt = AdamW(model.parameters(), lr=1e-3, weight_decay=1e-2)
    enc = model.neural_sdf.tcnn_encoding
    opt.set_lowp_shadow(enc.params, enc)
    c, v, off, d = ragged_rays(n_rays, dev)
    target = torch.rand(n_rays, 3, generator=torch.Generator().manual_seed(1)).to(dev)

    def step():
        out = model.render_rays_packed(c, v, d, off, validate=False)
        eik, curv = misc.sdf_losses(out["gradients"], out["hessians"], out["outside"])
        loss = torch.nn.functional.l1_loss(out["rgb"], target) * 3 + 0.1 * eik + 5e-4 * curv
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        return loss
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    names = ("na_neus_composite_packed_fwd", "na_neus_composite_packed_bwd")
    Fn.LAUNCH_EVENTS = {k: [] for k in names}
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ev, Fn.LAUNCH_EVENTS = Fn.LAUNCH_EVENTS, None
    assert torch.isfinite(loss).item(), "non-finite loss (packed batch)"
    counts = (off[1:] - off[:-1]).cpu()
    S = int(off[-1].item())
    # the persistent launches (NeuS alpha + compositing, one each way): algorithmic
    # bytes 32 B read + 8 B written per sample forward (sdf, gradient, distance,
    # colour; alpha, weight), 40 B read + 36 B written + the transmittance's 8 B
    # round trip backward (dsdf, dgrad, dcolour), 37 B per ray each way (offsets,
    # direction, far, outside flag, colour or its gradient)
    ms = sum(sum(a.elapsed_time(b) for a, b in ev[k]) / max(1, len(ev[k])) for k in names)
    nbytes = S * (40 + 84) + n_rays * 74
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    del model, opt
    torch.cuda.empty_cache()
    return dict(rays_per_s=round(n_rays / dt, 1), ms_per_step=round(dt * 1e3, 3),
                samples=S, samples_per_s=round(S / dt, 1),
                empty_rays=int((counts == 0).sum()), max_samples_per_ray=int(counts.max()),
                steps=steps, warmup=warmup,
                roofline=dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                              unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                              kernel="na_neus_composite_packed_fwd + _bwd (persistent, a "
                                     "wave per ray)",
                              bytes_per_sample=124, bytes_per_ray=74, avg_ms_fwd_plus_bwd=round(ms, 4),
                              share_of_step=round(ms / (dt * 1e3), 4)),
                workload=f"{n_rays} rays, 0..192 samples each (random union of spheres), "
                         "packed render, losses, backward, AdamW; BF16, 2^22 x 16 x 8")


def oracle_of(model, dtype):
    """The oracle field (SDF network + hash grid) and colour parameters of a model's
    weights, in `dtype`."""
    from oracle import field as ofield
    from oracle import hashgrid as hg
    enc = model.neural_sdf.tcnn_encoding
    lt = hg.level_table(enc.n_levels, enc.n_features_per_level, enc.log2_hashmap_size,
                        enc.base_resolution, enc.per_level_scale)
    sdf, rgb = model.neural_sdf, model.neural_rgb
    t = {}
    l0, l1 = sdf.mlp.linears
    t["W0v"], t["W0g"], t["b0"] = l0.weight_v, l0.weight_g, l0.bias
    t["W1v"], t["W1g"], t["b1"] = l1.weight_v, l1.weight_g, l1.bias
    t["ws"], t["bs"] = sdf.mlp.linear_sdf.weight, sdf.mlp.linear_sdf.bias
    for i, lin in enumerate(rgb.mlp.linears):
        t[f"R{i}v"], t[f"R{i}g"], t[f"R{i}b"] = lin.weight_v, lin.weight_g, lin.bias
    P = ofield.FieldParams({k: v.detach().cpu().to(dtype).clone().requires_grad_(True)
                            for k, v in t.items()})
    table = enc.params.detach().cpu().to(dtype).clone().requires_grad_(True)
    fld = ofield.SDFField(lt, table, P, vol_range=enc.vol_range,
                          active_levels=model.neural_sdf.active_levels, dtype=dtype)
    return fld, P


def error_vs_fp64(dev, n_rays=128, n_points=65536):
    """Output error of both precisions against the FP64 oracle, for the step's model at
    its seeded initialisation.  Field: SDF, 4-tap gradient and Hessian trace at
    n_points random points in [-1,1]^3.  Render: n_rays rays of the workload's cameras
    rendered in train mode at the samples the FP32 GPU path drew (stratification
    off), both precisions and the oracle rendering those samples.  sdf, colour: max abs; gradient: max |g - g_ref|
    / mean |g_ref|; tr H: median and 99th-percentile abs error."""
    from oracle import render as orr
    c, v = camera_rays(1)
    c, v = c[::1024 // n_rays][:n_rays].contiguous(), v[::1024 // n_rays][:n_rays].contiguous()
    x = torch.rand(n_points, 3, generator=torch.Generator().manual_seed(5)) * 2 - 1
    out = {"workload": f"field: {n_points} random points in [-1,1]^3; render: {n_rays} rays "
                       "of one camera, 96 samples each; the 2^22 x 16 x 8 model at its seeded "
                       "initialisation, train mode"}
    ref = field_ref = dists = None
    for dtype in (torch.float32, torch.bfloat16):
        m = build_model(dtype, dev)
        cd, vd = c.to(dev), v.to(dev)
        with torch.no_grad():
            # both precisions render the samples the FP32 path drew, where the
            # oracle renders them too (a precision's own hierarchical samples
            # would move with its SDF and compare different points)
            if dists is None:
                dists = m.render_rays(cd, vd)["dists"][..., 0]
            o = m.render_rays(cd, vd, dists=dists)
        if ref is None:
            fld, P = oracle_of(m, torch.float64)
            with torch.no_grad():
                ref = orr.render_rays(fld, P, c.double(), v.double(), m.s_var.detach().cpu().double(),
                                      1.0, m.neural_sdf.normal_eps, True,
                                      dists=dists.cpu().double()[..., None])
        sm = m.neural_sdf
        with torch.no_grad():  # the centre + taps node the train-step render runs
            s_f, _, taps, delta = sm.forward_with_taps(x.to(dev))
            g_f, h_f = sm.gradients_from_taps(taps, s_f, s_f, True, delta)
        if field_ref is None:
            with torch.no_grad():
                s_r = fld.sdf(x.double())
                g_r, h_r = fld.compute_gradients(x.double(), sm.normal_eps, training=True, sdf=s_r)
            field_ref = (s_r, g_r, h_r[:, 0] * 3)
        s_r, g_r, h_r = field_ref
        dhf = (h_f[:, 0].cpu().double() * 3 - h_r).abs()
        field = dict(sdf=float((s_f.cpu().double() - s_r).abs().max()),
                     gradient=float((g_f.cpu().double() - g_r).norm(dim=-1).max()
                                    / g_r.norm(dim=-1).mean()),
                     hessian_trace_median=float(dhf.median()),
                     hessian_trace_p99=float(dhf.quantile(0.99)),
                     hessian_trace_ref_median=float(h_r.abs().median()))
        inside = ~ref["outside"][:, 0]
        g, g_ref = o["gradients"].cpu().double()[inside], ref["gradients"][inside]
        h = o["hessians"].cpu().double()[inside][..., 0] * 3
        h_ref = ref["hessians"][inside][..., 0] * 3
        dh = (h - h_ref).abs().reshape(-1)
        out["fp32" if dtype == torch.float32 else "bf16"] = dict(field=field, render=dict(
            sdf=float((o["sdfs"].cpu().double()[inside] - ref["sdfs"][inside]).abs().max()),
            rgb=float((o["rgb"].cpu().double() - ref["rgb"]).abs().max()),
            gradient=float((g - g_ref).norm(dim=-1).max() / g_ref.norm(dim=-1).mean()),
            hessian_trace_median=float(dh.median()), hessian_trace_p99=float(dh.quantile(0.99)),
            hessian_trace_ref_median=float(h_ref.abs().median())))
        del m
        torch.cuda.empty_cache()
    return out


def cpu_baseline(n_rays=256, log2_T=22, F=8):
    """The CPU oracle (torch, reference formulation) on a bounded sample of the workload."""
    from oracle import render as orr
    torch.set_num_threads(os.cpu_count() or 1)
    torch.manual_seed(0)
    from neuralangelo_amd.model import Model, default_config
    m = Model(default_config(log2_T=log2_T, F=F), compute_dtype=torch.float32)
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    fld, P = oracle_of(m, torch.float32)
    c, v = camera_rays(1)
    c, v = c[:n_rays], v[:n_rays]
    target = torch.rand(n_rays, 3)
    sv = torch.tensor(3.0, requires_grad=True)

    def step():
        o = orr.render_rays(fld, P, c, v, sv, 1.0, 1.0 / 2048, True,
                            rands=torch.rand(n_rays, 64))
        loss, _ = orr.train_loss(o, target)
        loss.backward()

    step()  # warm-up
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    return dict(value=round(n_rays / dt, 2), unit="rays/s", cores=torch.get_num_threads(),
                kind="port", cpu=cpu_model(),
                sample=f"{n_rays} rays x 96 samples, fwd+bwd of the float32 torch oracle, "
                       f"same 2^{log2_T}x16x{F} model, median of 3 after 1 warm-up "
                       f"({dt:.2f} s/step)")


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and its logical core count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({os.cpu_count()} logical cores visible)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 comparison run")
    ap.add_argument("--no-mc", action="store_true", help="skip the mesh-extraction leg")
    ap.add_argument("--no-torch", action="store_true",
                    help="skip the stock-torch FP32 baseline (the plain FP32 GPU path)")
    ap.add_argument("--no-render", action="store_true", help="skip the eval-render leg")
    ap.add_argument("--no-packed", action="store_true", help="skip the ragged packed-batch leg")
    ap.add_argument("--no-err", action="store_true", help="skip the FP64-oracle error leg")
    ap.add_argument("--mc-res", type=int, default=1024)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one device copy of MIB MiB (a known-size dispatch that "
                         "calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of "
                                   "it (the path would shard rays across GPUs and all-reduce "
                                   "the table gradient)"}))
        sys.exit(2)
    dev = torch.device("cuda", 0)
    from neuralangelo_amd import _lib
    _lib.load()
    if args.copy_probe > 0:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty(n, device=dev)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    model, dt_bf16, loss_bf16, taps_ms, n_taps, captured = time_train(
        torch.bfloat16, dev, args.rays, args.steps, args.warmup)
    roof = taps_bwd_roofline(model, args.rays * 96, taps_ms, n_taps, 2, captured)
    del captured
    hash_fwd = encoder_roofline(model, dev)
    del model
    torch.cuda.empty_cache()
    extra = {}
    if not args.no_fp32:
        m32, dt_fp32, loss_fp32, _, _, _ = time_train(torch.float32, dev, args.rays, args.steps,
                                                      args.warmup)
        extra["fp32_rays_per_s"] = round(args.rays / dt_fp32, 1)
        extra["fp32_ms_per_step"] = round(dt_fp32 * 1e3, 3)
        extra["bf16_over_fp32"] = round(dt_fp32 / dt_bf16, 3)
        del m32
        torch.cuda.empty_cache()
    if not args.no_torch:
        dt_torch = time_torch_fp32(dev, args.rays, max(1, min(args.steps, 3)), 1)
        extra["torch_fp32_rays_per_s"] = round(args.rays / dt_torch, 1)
        extra["torch_fp32_ms_per_step"] = round(dt_torch * 1e3, 3)
        extra["bf16_over_torch_fp32"] = round(dt_torch / dt_bf16, 3)
    if not args.no_render:
        extra["render"] = {
            "workload": "eval render via Model.inference: one 32x32 image (1024 rays) at "
                        "radius 3, FOV 50 deg, hash grid 2^19 x 16 x 4, 64 + 2x16 samples",
            "fp32": render_leg(dev, torch.float32), "bf16": render_leg(dev, torch.bfloat16)}
    if not args.no_packed:
        extra["packed"] = packed_leg(dev)
    if not args.no_err:
        extra["error_vs_fp64"] = error_vs_fp64(dev)
    if not args.no_mc:
        # a fixed volume: the FP32 network at its seeded initialisation (not the
        # trained state above, whose surface changes from run to run)
        extra["mc"] = mc_leg(build_model(torch.float32, dev), dev, res=args.mc_res,
                             cpu=not args.no_cpu)
    torch.cuda.empty_cache()
    cpu = cpu_baseline() if not args.no_cpu else None
    rec = {
        "metric": "train rays/s (fwd+bwd) BF16 vs FP32 at stated err; hash GB/s; MC voxels/s",
        "value": round(args.rays / dt_bf16, 1),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_bf16 * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded sphere cameras, random targets, random-init weights)",
        "config": {"workload": "train step: hash grid 2^22 x 16 levels x 8 features, SDF MLP "
                               "1x256 + colour MLP 4x256, 8192 rays x (64 + 2x16) samples, "
                               "4-tap numerical gradient, AdamW",
                   "rays": args.rays, "samples_per_ray": 96},
        "roofline": roof,
        "cpu_baseline": cpu,
        "hash_GBps": hash_fwd["achieved"],
        "hash_fwd": hash_fwd,
        "loss": round(loss_bf16, 5),
        **extra,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
