"""Launch-shape autotuner.

The library's launch tunables (na_tune_set / na_tune_get: the tap scatter's run
of samples per thread, the tap-increment backward's launch form and workgroups
per CU, the persistent packed launches' rays per workgroup and workgroups per CU)
have a best value that depends on the device and on the shape of the work.  `tune()` times every candidate of each on
the workload it serves -- the train step's samples (8192 rays x 96 samples,
ray-ordered, as the renderer hands them over) and the ragged packed batch
(65,536 rays, 0..192 samples) -- with HIP events on the launching stream, and
keeps the fastest.  The winners are stored per device architecture in
`tune_cache.json` beside this file; `_lib.load()` applies the entries for the
architecture the library is built for (gfx950), so a user never tunes.

    python -m neuralangelo_amd.autotune          # re-tune and rewrite the cache
"""
import ctypes
import json
import os

import torch

from . import _lib

CACHE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_cache.json")
ARCH = "gfx950"

CANDIDATES = {
    "taps_bwd.run": [1, 2, 4, 6, 8, 12, 16],
    "taps_increment_bwd.wg_per_cu": [2, 4, 8, 16],
    "packed.rays_per_wg": [64, 128, 256],
    "packed.wg_per_cu": [1, 2, 4, 8, 16],
    "taps_increment_bwd.fused": [0, 1],
}


def get(name):
    return _lib.call("na_tune_get", name.encode())


def set(name, value):  # noqa: A001  (mirrors na_tune_set)
    return _lib.call("na_tune_set", name.encode(), int(value))


def load_cache(path=CACHE_PATH):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def apply_cache(lib, path=CACHE_PATH, arch=ARCH):
    """Set every tunable the cache holds for `arch` (called by _lib.load)."""
    entries = load_cache(path).get(arch, {})
    for name, rec in entries.items():
        value = rec["value"] if isinstance(rec, dict) else rec
        if lib.na_tune_set(name.encode(), int(value)) < 0:
            raise RuntimeError(f"tune cache {path}: {name} = {value} rejected: "
                               f"{lib.na_last_error().decode()}")


def _time(fn, reps=5, warmup=2):
    """Median milliseconds of fn() with HIP events on torch's current stream."""
    for _ in range(warmup):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = []
    for _ in range(reps):
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    ms.sort()
    return ms[len(ms) // 2]


def _sweep(name, fn):
    """Time fn under every candidate of `name`; leave the best set; return the record."""
    old = get(name)
    timings = {}
    try:
        for v in CANDIDATES[name]:
            set(name, v)
            timings[v] = _time(fn)
    finally:
        set(name, old)
    best = min(timings, key=timings.get)
    set(name, best)
    return dict(value=best, timings_ms={str(k): round(t, 4) for k, t in timings.items()})


def _sphere_rays(n, dev, radius=3.0, seed=0):
    g = torch.Generator().manual_seed(seed)
    c = torch.randn(n, 3, generator=g)
    c = c / c.norm(dim=-1, keepdim=True) * radius
    v = -c / radius + 0.25 * torch.randn(n, 3, generator=g)
    return c.to(dev), (v / v.norm(dim=-1, keepdim=True)).to(dev)


@torch.no_grad()
def _step_points(model, n_rays, dev):
    """The train step's sample points: the model's own sampling of n_rays rays,
    ray-ordered (each ray's samples consecutive)."""
    c, v = _sphere_rays(n_rays, dev)
    d = model.render_rays(c, v)["dists"][..., 0]
    return (c[:, None, :] + v[:, None, :] * d[..., None]).reshape(-1, 3).contiguous()


def _ragged(n_rays, dev, seed=11):
    """Ragged per-ray sample counts 0..192 (a union of spheres) as CSR offsets."""
    g = torch.Generator().manual_seed(seed)
    k = int(torch.randint(1, 9, (1,), generator=g))
    sc = torch.rand(k, 3, generator=g) * 1.2 - 0.6
    sr = torch.rand(k, generator=g) * 0.4 + 0.1
    c, v = _sphere_rays(n_rays, "cpu", seed=seed)
    oc = c[:, None, :] - sc[None]
    b = (oc * v[:, None, :]).sum(-1)
    disc = b * b - ((oc * oc).sum(-1) - sr[None] ** 2)
    n_c = ((2.0 * disc.clamp(min=0.0).sqrt()).sum(-1) * 160.0).floor().clamp(max=160).long()
    counts = torch.where(n_c >= 2, n_c + 32, n_c)
    off = torch.zeros(n_rays + 1, dtype=torch.int32)
    off[1:] = counts.cumsum(0).to(torch.int32)
    return off.to(dev)


def tune(dev=None, write=True, path=CACHE_PATH):
    """Tune every launch tunable on its workload; returns (and by default writes)
    the cache entries for this architecture."""
    from .model import Model, default_config
    dev = dev or torch.device("cuda", torch.cuda.current_device())
    _lib.load()
    arch = torch.cuda.get_device_properties(dev).gcnArchName.split(":")[0]
    torch.manual_seed(0)
    m = Model(default_config(log2_T=22, F=8), compute_dtype=torch.bfloat16).to(dev).train()
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    enc = m.neural_sdf.tcnn_encoding
    out = {}

    # tap scatter on the train step's own inputs: three BF16 train steps (the
    # geometric initialisation leaves the encoder's input gradient zero in the
    # first), the last step's scatter arguments captured and replayed
    from . import functional as Fn
    from . import misc
    from .optim import AdamW
    opt = AdamW(m.parameters(), lr=1e-3, weight_decay=1e-2)
    opt.set_lowp_shadow(enc.params, enc)
    c, v = _sphere_rays(8192, dev)
    target = torch.rand(8192, 3, generator=torch.Generator().manual_seed(1)).to(dev)
    for k in range(3):
        Fn.SCATTER_CAPTURE = [] if k == 2 else None
        out_r = m.render_rays(c, v, stratified=True)
        eik, curv = misc.sdf_losses(out_r["gradients"], out_r["hessians"], out_r["outside"])
        loss = torch.nn.functional.l1_loss(out_r["rgb"], target) * 3 + 0.1 * eik + 5e-4 * curv
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
    cap, Fn.SCATTER_CAPTURE = Fn.SCATTER_CAPTURE, None
    del out_r, loss, opt
    pts, eps, dXc, dXt, dXq, ld, inc = cap[-1]
    P = pts.shape[0]
    grad = torch.zeros(enc.n_params, device=dev)
    vmin, vmax = enc.vol_range
    g = torch.Generator(device=dev).manual_seed(1)

    def scatter():
        _lib.call("na_sdf_taps_bwd", ctypes.byref(enc.desc), _lib.ptr(pts), P, float(vmin),
                  float(vmax), float(eps), _lib.ptr(dXc), _lib.ptr(dXt),
                  _lib.ptr(dXq) if dXq is not None else None, _lib.dtype_code(dXc.dtype), ld,
                  _lib.ptr(grad), int(enc.active_levels), int(inc), _lib.stream_of(pts))
    rec = _sweep("taps_bwd.run", scatter)
    rec["workload"] = (f"na_sdf_taps_bwd on a BF16 train step's own inputs ({P} samples of "
                       "8192 rays, the step's dL/dy rows)")
    out["taps_bwd.run"] = rec
    del cap, dXc, dXt, dXq, grad

    # tap-increment backward at the train step's shape
    P32 = (P + 31) // 32 * 32
    hid, ld_d = 256, 272
    D = (torch.randn(4 * P32, ld_d, device=dev, generator=g) * 1e-3).to(torch.bfloat16)
    W = (torch.randn(hid, ld_d, device=dev, generator=g) * 0.05).to(torch.bfloat16)
    Zc = torch.randn(P32, hid, device=dev, generator=g) * 0.05
    ws = torch.randn(hid, device=dev, generator=g) * 0.05
    dds = torch.randn(P32, 4, device=dev, generator=g)
    dpre = torch.empty(4 * P32, hid, dtype=torch.bfloat16, device=dev)
    centre = torch.zeros(P32, hid, dtype=torch.bfloat16, device=dev)
    q = torch.empty(P32, hid, dtype=torch.bfloat16, device=dev)
    db0 = torch.zeros(hid, device=dev)
    dws = torch.zeros(hid, device=dev)
    work = torch.empty(4 * P32, hid, device=dev)

    def taps_bwd():
        _lib.call("na_taps_increment_bwd", _lib.ptr(D), ld_d, 4 * P32, ld_d, _lib.ptr(W), ld_d,
                  hid, _lib.ptr(Zc), hid, _lib.ptr(ws), 100.0, _lib.ptr(dds), _lib.ptr(dpre), hid, _lib.ptr(centre), hid, _lib.ptr(q), hid, _lib.ptr(db0),
                  _lib.ptr(dws), _lib.ptr(work), 0, _lib.stream_of(D))
    rec = _sweep("taps_increment_bwd.fused", taps_bwd)
    rec["workload"] = f"na_taps_increment_bwd, {4 * P32} tap rows x 256 columns"
    out["taps_increment_bwd.fused"] = rec
    old_fused = get("taps_increment_bwd.fused")
    set("taps_increment_bwd.fused", 0)  # the grid of the split elementwise pass
    rec = _sweep("taps_increment_bwd.wg_per_cu", taps_bwd)
    set("taps_increment_bwd.fused", old_fused)
    rec["workload"] = f"na_taps_increment_bwd (split launch), {4 * P32} tap rows x 256 columns"
    out["taps_increment_bwd.wg_per_cu"] = rec
    del D, W, Zc, dds, dpre, centre, q, work

    # the persistent packed NeuS + compositing launches on the ragged batch: samples
    # along each ray 0.01 apart, the SDF crossing zero along the ray
    off = _ragged(65536, dev)
    R, S = off.numel() - 1, int(off[-1].item())
    ray_idx = torch.empty(S, dtype=torch.int32, device=dev)
    _lib.call("na_packed_ray_index", _lib.ptr(off), R, _lib.ptr(ray_idx), _lib.stream_of(off))
    pos = (torch.arange(S, device=dev) - off.long()[ray_idx.long()]).float()
    dists = 1.0 + 0.01 * pos
    sdfs = 0.3 - 0.007 * pos
    grads = torch.nn.functional.normalize(torch.randn(S, 3, device=dev, generator=g), dim=-1)
    rays = torch.nn.functional.normalize(torch.randn(R, 3, device=dev, generator=g), dim=-1)
    s_var = torch.full((1,), 3.0, device=dev)
    rgbs = torch.rand(S, 3, device=dev, generator=g)
    alpha, wts = torch.empty(S, device=dev), torch.empty(S, device=dev)
    rgb = torch.empty(R, 3, device=dev)
    dout = torch.randn(R, 3, device=dev, generator=g)
    dsdf, dgrad = torch.empty(S, device=dev), torch.empty(S, 3, device=dev)
    drgbs, ds_var = torch.empty(S, 3, device=dev), torch.zeros(1, device=dev)
    queues = torch.zeros(2, dtype=torch.int32, device=dev)

    def packed():
        s = _lib.stream_of(alpha)
        queues.zero_()
        _lib.call("na_neus_composite_packed_fwd", _lib.ptr(rays), _lib.ptr(sdfs), _lib.ptr(grads),
                  _lib.ptr(dists), None, None, 1000.0, _lib.ptr(s_var), 1.0, 0.1, _lib.ptr(off), R,
                  _lib.ptr(rgbs), 3, _lib.ptr(alpha), _lib.ptr(wts), _lib.ptr(rgb),
                  _lib.ptr(queues[:1]), s)
        _lib.call("na_neus_composite_packed_bwd", _lib.ptr(rays), _lib.ptr(sdfs), _lib.ptr(grads),
                  _lib.ptr(dists), None, None, 1000.0, _lib.ptr(s_var), 1.0, 0.1, _lib.ptr(off), R,
                  _lib.ptr(rgbs), 3, _lib.ptr(alpha), _lib.ptr(wts), _lib.ptr(dout), None, None,
                  _lib.ptr(dsdf), _lib.ptr(dgrad), _lib.ptr(drgbs), _lib.ptr(ds_var),
                  _lib.ptr(queues[1:]), s)
    for name in ("packed.rays_per_wg", "packed.wg_per_cu"):
        rec = _sweep(name, packed)
        rec["workload"] = (f"na_neus_composite_packed_fwd + _bwd (persistent), {R} rays, {S} "
                           "samples (0..192 per ray)")
        out[name] = rec
    del m
    torch.cuda.empty_cache()
    if write:
        cache = load_cache(path)
        cache[arch] = out
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
            f.write("\n")
    return out


if __name__ == "__main__":
    res = tune()
    print(json.dumps(res, indent=1))
