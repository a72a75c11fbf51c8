"""Benchmark: Neuralangelo train step (fwd+bwd+AdamW) on one MI355X.

Workload (the 2^22 x 16 x 8 train-step configuration): hash grid 16 levels,
2^22 entries/level, 8 features, resolutions 2^5..2^11; SDF MLP 1x256 softplus
(geometric init), colour MLP 4x256 ReLU; 8192 rays from 8 cameras on a sphere
of radius 3 (32x32 px, FOV 50 deg) x (64 coarse + 2 x 16 resampled) samples;
4-tap numerical gradient; loss L1*3 + 0.1 eikonal + 5e-4 curvature; AdamW
(lr 1e-3, wd 1e-2) over all parameters.  Synthetic inputs, random targets.

Prints one JSON line.  value = train rays/s of the BF16 path; the FP32 path
is measured beside it.  roofline = the step's dominant kernel, the hash-grid
backward over the centre point and its 4 taps (HBM/atomic-bound scatter),
timed with HIP events on torch's stream around its launch inside the timed
steps; traffic = its HBM bytes per launch from the committed rocprofv3
counter summary.  hash_fwd = the encoder forward kernel (hash GB/s).
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
    for _ in range(warmup):
        loss = train_step(model, opt, c, v, target)
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
    return model, dt, float(loss.item()), launch_ms, len(pairs)


ROOF_CALL = "na_sdf_taps_bwd"
ROOF_KERNEL = "hashgrid_taps_bwd_kernel"
TRAFFIC_FILE = os.path.join(REPO, "profiles", "hashgrid_taps_bwd_train_step_pmc.txt")


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


def taps_bwd_roofline(model, n_points, launch_ms, n_launches, dy_bytes):
    """Roofline of the dominant kernel of the step: the hash-grid backward over the
    centre point and its four taps (one launch per step, timed with HIP events on
    torch's stream inside the timed steps).  Algorithmic bytes per point-eval of the
    encoder backward: L*F*dy_bytes (dL/dy rows in the compute dtype) + 12 (point)
    + L*8*F*4 (fp32 corner gradient RMW, counted once); 5 evals (centre + 4 taps)
    per point."""
    enc = model.neural_sdf.tcnn_encoding
    L, F = enc.n_levels, enc.n_features_per_level
    per_eval = L * F * dy_bytes + 12 + L * 8 * F * 4
    per_point = 5 * per_eval
    achieved = per_point * n_points / (launch_ms * 1e-3) / 1e9
    traffic = measured_traffic()
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=round(traffic) if traffic is not None else None,
                kernel=f"{ROOF_KERNEL} (centre + 4 taps, merged corner scatter)",
                bytes_per_point=per_point, points_per_launch=n_points,
                avg_launch_ms=round(launch_ms, 4), launches_timed=n_launches,
                traffic_source=os.path.relpath(TRAFFIC_FILE, REPO) if traffic is not None else None)


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
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                kernel="hashgrid_fwd_kernel (SDF input)", bytes_per_point=per_point,
                points_per_launch=n_points, avg_launch_ms=round(ms, 4))


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
                      frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
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


def cpu_baseline(n_rays=256, log2_T=22, F=8):
    """The CPU oracle (torch, reference formulation) on a bounded sample of the workload."""
    from neuralangelo_amd.model import Model, default_config
    from oracle import field as ofield
    from oracle import hashgrid as hg
    from oracle import render as orr
    torch.set_num_threads(os.cpu_count() or 1)
    torch.manual_seed(0)
    m = Model(default_config(log2_T=log2_T, F=F), compute_dtype=torch.float32)
    enc = m.neural_sdf.tcnn_encoding
    lt = hg.level_table(16, F, log2_T, 32, enc.per_level_scale)
    sdf, rgb = m.neural_sdf, m.neural_rgb
    t = {}
    l0, l1 = sdf.mlp.linears
    t["W0v"], t["W0g"], t["b0"] = l0.weight_v, l0.weight_g, l0.bias
    t["W1v"], t["W1g"], t["b1"] = l1.weight_v, l1.weight_g, l1.bias
    t["ws"], t["bs"] = sdf.mlp.linear_sdf.weight, sdf.mlp.linear_sdf.bias
    for i, lin in enumerate(rgb.mlp.linears):
        t[f"R{i}v"], t[f"R{i}g"], t[f"R{i}b"] = lin.weight_v, lin.weight_g, lin.bias
    P = ofield.FieldParams({k: v.detach().clone().requires_grad_(True) for k, v in t.items()})
    table = enc.params.detach().clone().requires_grad_(True)
    fld = ofield.SDFField(lt, table, P, dtype=torch.float32)
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
    for _ in range(2):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    return dict(value=round(n_rays / dt, 2), unit="rays/s", cores=torch.get_num_threads(),
                kind="port",
                sample=f"{n_rays} rays x 96 samples, fwd+bwd of the float32 torch oracle, "
                       f"same 2^{log2_T}x16x{F} model, median of 2 after 1 warm-up "
                       f"({dt:.2f} s/step)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 comparison run")
    ap.add_argument("--no-mc", action="store_true", help="skip the mesh-extraction leg")
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
    model, dt_bf16, loss_bf16, taps_ms, n_taps = time_train(torch.bfloat16, dev, args.rays,
                                                           args.steps, args.warmup)
    roof = taps_bwd_roofline(model, args.rays * 96, taps_ms, n_taps, dy_bytes=2)
    hash_fwd = encoder_roofline(model, dev)
    del model
    torch.cuda.empty_cache()
    extra = {}
    if not args.no_fp32:
        m32, dt_fp32, loss_fp32, _, _ = time_train(torch.float32, dev, args.rays, args.steps,
                                                   args.warmup)
        extra["fp32_rays_per_s"] = round(args.rays / dt_fp32, 1)
        extra["fp32_ms_per_step"] = round(dt_fp32 * 1e3, 3)
        extra["bf16_over_fp32"] = round(dt_fp32 / dt_bf16, 3)
    else:
        m32 = build_model(torch.float32, dev)
    if not args.no_mc:
        # a fixed volume: the FP32 network at its seeded initialisation (not the
        # trained state above, whose surface changes from run to run)
        extra["mc"] = mc_leg(build_model(torch.float32, dev), dev, res=args.mc_res,
                             cpu=not args.no_cpu)
    del m32
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
