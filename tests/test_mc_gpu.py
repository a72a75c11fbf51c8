"""Marching cubes on the HIP path vs the C oracle: bit-exact vertices and faces."""
import numpy as np
import pytest
import torch

from neuralangelo_amd import mcubes
from neuralangelo_amd import mesh as gmesh
from oracle import mesh as omesh

pytestmark = pytest.mark.gpu


def _check(vol, iso=0.0):
    V, F = mcubes.marching_cubes_gpu(torch.from_numpy(vol).cuda(), iso)
    Vr, Fr = omesh.marching_cubes(vol, iso)
    V, F = V.cpu().numpy(), F.cpu().numpy()
    assert V.shape == Vr.shape and F.shape == Fr.shape, (V.shape, Vr.shape, F.shape, Fr.shape)
    assert np.array_equal(F, Fr)
    assert np.array_equal(V, Vr)  # bit-exact float64
    return V, F


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 5, 7), (17, 19, 23), (65, 64, 66), (129, 33, 2),
                                   (2, 40, 3)])
def test_random_volumes_bit_exact(shape):
    rng = np.random.default_rng(sum(shape))
    print("mc random", shape, flush=True)
    _check(rng.normal(size=shape).astype(np.float32))


def test_sphere_bit_exact():
    n = 97
    g = np.linspace(-1, 1, n, dtype=np.float32)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    vol = (X * X + Y * Y + Z * Z - np.float32(0.36)).astype(np.float32)
    V, F = _check(vol)
    assert F.shape[0] > 1000


def test_ties_and_nonzero_iso():
    rng = np.random.default_rng(7)
    vol = rng.integers(-2, 3, size=(31, 29, 27)).astype(np.float32)  # exact ties with iso
    _check(vol, 0.0)
    _check(vol, 1.0)
    _check(rng.normal(size=(20, 21, 22)).astype(np.float32), 0.37)


def test_empty_full_and_degenerate():
    for vol in (np.ones((9, 9, 9), np.float32), -np.ones((9, 9, 9), np.float32),
                np.zeros((1, 4, 4), np.float32), np.zeros((4, 4, 0), np.float32)):
        V, F = mcubes.marching_cubes_gpu(torch.from_numpy(vol).cuda(), 0.0)
        assert V.shape == (0, 3) and F.shape == (0, 3)


def test_large_block_bit_exact():
    """A full 129^3 lattice block (block_res 128) of a noisy sphere."""
    n = 129
    g = np.linspace(-1, 1, n, dtype=np.float32)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    rng = np.random.default_rng(11)
    vol = (np.sqrt(X * X + Y * Y + Z * Z) - np.float32(0.5)
           + np.float32(0.02) * rng.normal(size=X.shape)).astype(np.float32)
    _check(vol)


def test_numpy_dropin_semantics():
    rng = np.random.default_rng(3)
    vol = rng.normal(size=(12, 13, 14)).astype(np.float32)
    V, F = mcubes.marching_cubes(vol, 0.0)
    assert isinstance(V, np.ndarray) and V.dtype == np.float64 and F.dtype == np.int64
    Vr, Fr = omesh.marching_cubes(vol, 0.0)
    assert np.array_equal(V, Vr) and np.array_equal(F, Fr)
    # float64 input is converted to float32 the same way on both sides
    V2, F2 = mcubes.marching_cubes(vol.astype(np.float64), 0.0)
    assert np.array_equal(V2, Vr) and np.array_equal(F2, Fr)


@pytest.mark.parametrize("shape,block_res", [((40, 37, 33), 16), ((33, 33, 33), 8),
                                             ((65, 20, 9), 64), ((17, 18, 19), 1)])
def test_lattice_blocks_bit_exact(shape, block_res):
    """All blocks in one launch chain == the per-block loop, with and without the world map."""
    rng = np.random.default_rng(block_res)
    vol = rng.normal(size=shape).astype(np.float32)
    print("mc lattice", shape, block_res, flush=True)
    V, F = mcubes.lattice_marching_cubes(torch.from_numpy(vol).cuda(), block_res)
    Vr, Fr = omesh.lattice_marching_cubes(vol, block_res)
    assert np.array_equal(F.cpu().numpy(), Fr) and np.array_equal(V.cpu().numpy(), Vr)
    intv = 2.0 / 50
    axes = [torch.arange(-1.0, -1.0 + n * intv - intv / 2, intv)[:n] for n in shape]
    V, F = mcubes.lattice_marching_cubes(torch.from_numpy(vol).cuda(), block_res,
                                         axes=[a.cuda() for a in axes], intv=intv)
    Vr, Fr = omesh.lattice_marching_cubes(vol, block_res, 0.0, [a.numpy() for a in axes], intv)
    assert np.array_equal(F.cpu().numpy(), Fr) and np.array_equal(V.cpu().numpy(), Vr)


def _sphere_sdf(xyz):
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    return -(x * x + y * y + z * z - 0.4)[..., None]


@pytest.mark.parametrize("block_res", [16, 23])
def test_extract_mesh_blocks_match_oracle(block_res):
    bounds = ((-1.0, 1.0), (-1.0, 1.0), (-1.0, 1.0))
    intv = 2.0 / 64
    V, F = gmesh.extract_mesh(_sphere_sdf, bounds, intv, block_res=block_res,
                              device=torch.device("cuda", 0))
    Vr, Fr = omesh.extract_mesh(_sphere_sdf, bounds, intv, block_res=block_res)
    assert V.shape == Vr.shape and F.shape == Fr.shape
    assert np.array_equal(F, Fr)
    assert np.array_equal(V, Vr)
    assert np.all(np.linalg.norm(V, axis=-1) < 1.0)


def _zero_sphere_sdf(xyz):
    """Exact zeros at lattice points: with intv = 1/32 every coordinate is k/32, whose
    square and the sum of three squares are exact in float32, so the sphere of radius^2
    854/1024 passes exactly through the 528 lattice points with a^2 + b^2 + c^2 = 854
    (marching cubes repeats vertices there)."""
    return -((xyz * xyz).sum(-1, keepdim=True) - 854.0 / 1024.0)


@pytest.mark.parametrize("block_res", [16, 64])
def test_weld_blocks_and_processed_extraction_match_oracle(block_res):
    """trimesh's construction-time merge of every block (weld_blocks) against the
    oracle's sequential restatement block by block, and the processed extract_mesh
    (welding, sphere filter, dropping vertices left without faces) against the
    oracle's per-block loop, on a field with exact zeros at lattice points."""
    dev = torch.device("cuda", 0)
    res, intv = 64, 2.0 / 64
    bounds = ((-1.0, 1.0),) * 3
    axes = gmesh.lattice_axes(bounds, intv)
    vol = gmesh.lattice_sdf(_zero_sphere_sdf, axes, dev)
    assert int((vol == 0).sum()) > 20
    V, F, bv, bf = mcubes.lattice_marching_cubes(vol, block_res, 0.0, axes=[a.to(dev) for a in axes],
                                                 intv=intv, block_ranges=True)
    W, G, keep, bw = gmesh.weld_blocks(V, F, bv)
    Vn, Fn, bvn, bfn = V.cpu().numpy(), F.cpu().numpy(), bv.cpu().numpy(), bf.cpu().numpy()
    Ws, Gs, off = [], [], 0
    for b in range(len(bvn) - 1):
        Vb = Vn[bvn[b]:bvn[b + 1]]
        Fb = Fn[bfn[b]:bfn[b + 1]] - bvn[b]
        Vp, Fp = omesh.trimesh_process(Vb, Fb)
        Ws.append(Vp)
        Gs.append(Fp + off)
        off += len(Vp)
    Wr, Gr = np.concatenate(Ws), np.concatenate(Gs)
    assert len(Wr) < len(Vn)  # vertices were merged
    assert np.array_equal(W.cpu().numpy(), Wr) and np.array_equal(G.cpu().numpy(), Gr)
    assert np.array_equal(Vn[keep.cpu().numpy()], Wr)
    assert bw[-1].item() == len(Wr)
    Vg, Fg = gmesh.extract_mesh(_zero_sphere_sdf, bounds, intv, block_res=block_res, device=dev)
    Vr, Fr = omesh.extract_mesh(_zero_sphere_sdf, bounds, intv, block_res=block_res)
    assert np.array_equal(Fg, Fr) and np.array_equal(Vg, Vr)


def _spheres_sdf(xyz):
    """Negated implicit function of a union of small spheres (several components per
    block): min over spheres of |x - c|^2 - r^2, written with products and sums only so
    that the CPU and the GPU evaluate the same float32 values."""
    cs = [(0.3, 0.2, -0.1), (-0.35, 0.1, 0.2), (0.05, -0.4, 0.3), (0.0, 0.0, 0.0),
          (0.62, -0.5, 0.1)]
    rs = [0.18, 0.12, 0.09, 0.05, 0.2]
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    out = None
    for (cx, cy, cz), r in zip(cs, rs):
        dx, dy, dz = x - cx, y - cy, z - cz
        f = dx * dx + dy * dy + dz * dz - r * r
        out = f if out is None else torch.minimum(out, f)
    return -out[..., None]


@pytest.mark.parametrize("shape,block_res", [((65, 64, 66), 16), ((40, 33, 50), 23)])
def test_block_offsets_match_per_block_counts(shape, block_res):
    rng = np.random.default_rng(7)
    vol = rng.normal(size=shape).astype(np.float32)
    V, F, bv, bf = mcubes.lattice_marching_cubes(torch.from_numpy(vol).cuda(), block_res,
                                                 block_ranges=True)
    _, _, cv, cf = omesh.lattice_marching_cubes(vol, block_res, counts=True)
    # the reference's LatticeGrid counts ceil(points / block_res) blocks per axis, so a
    # trailing block may hold a single point (no cells); the kernel's plan has only
    # blocks with cells.  Both orders are x-major.
    nb = [int(np.ceil(d / block_res)) for d in shape]
    starts = [np.arange(k) * block_res for k in nb]
    bx, by, bz = np.meshgrid(*starts, indexing="ij")
    has_cells = ((bx < shape[0] - 1) & (by < shape[1] - 1) & (bz < shape[2] - 1)).reshape(-1)
    assert np.array_equal(bv.cpu().numpy(), np.concatenate([[0], np.cumsum(cv[has_cells])]))
    assert np.array_equal(bf.cpu().numpy(), np.concatenate([[0], np.cumsum(cf[has_cells])]))
    assert bv[-1].item() == V.shape[0] and bf[-1].item() == F.shape[0]


@pytest.mark.parametrize("block_res", [16, 40])
def test_extract_mesh_texture_and_largest_component_match_oracle(block_res):
    """extract_mesh with per-vertex colours and the per-block largest-component filter
    against the oracle's per-block pipeline (mesh.py:119-133).  The colour function is a
    fixed function of the vertex position here, so colours compare exactly."""
    bounds = ((-1.0, 1.0), (-1.0, 1.0), (-1.0, 1.0))
    intv = 2.0 / 80

    def texture(V):
        v = V.cpu().numpy() if torch.is_tensor(V) else np.asarray(V)
        return (np.abs(np.sin(v * 7.0)) * 255).astype(np.uint8)

    V, F, C = gmesh.extract_mesh(_spheres_sdf, bounds, intv, block_res=block_res,
                                 texture_func=texture, filter_lcc=True,
                                 device=torch.device("cuda", 0))
    Vr, Fr, Cr = omesh.extract_mesh(_spheres_sdf, bounds, intv, block_res=block_res,
                                    texture_func=texture, filter_lcc=True)
    print("lcc mesh", V.shape, Vr.shape, flush=True)
    assert V.shape == Vr.shape and F.shape == Fr.shape
    assert np.array_equal(F, Fr)
    assert np.array_equal(V, Vr)
    assert np.array_equal(C, Cr)
    # the filter removed components: fewer faces than without it
    V2, F2 = gmesh.extract_mesh(_spheres_sdf, bounds, intv, block_res=block_res,
                                device=torch.device("cuda", 0))
    assert F2.shape[0] > F.shape[0]


def test_extract_texture_formula():
    """mesh.py:53-65: colour net at the vertex, normal = unit SDF gradient, view
    direction = -normal, rgb * 255 truncated to uint8."""
    from neuralangelo_amd.model import Model, default_config
    from tests.helpers import randomize_model
    torch.manual_seed(0)
    m = Model(default_config(log2_T=15))
    randomize_model(m, 3)
    m = m.cuda().eval()
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    m.neural_sdf.set_normal_epsilon()
    g = torch.Generator().manual_seed(1)
    x = (torch.rand(5000, 3, generator=g) * 1.6 - 0.8).cuda()
    C = gmesh.extract_texture(x.double().cpu().numpy(), m.neural_rgb, m.neural_sdf)
    with torch.no_grad():
        sdf, feat = m.neural_sdf.forward(x)
        grad, _ = m.neural_sdf.compute_gradients(x, training=False, sdf=sdf)
        n = torch.nn.functional.normalize(grad, dim=-1)
        ref = (m.neural_rgb.forward(x, n, -n, feat, None).cpu().numpy() * 255).astype(np.uint8)
    assert C.shape == (5000, 3) and C.dtype == np.uint8
    assert np.abs(C.astype(int) - ref.astype(int)).max() <= 1


def test_full_lattice_1024_network_sdf():
    """Mesh extraction at the 1024^3 configuration: the FP32 2^22 x 16 x 8 network SDF
    (seeded initialisation: a sphere of radius 0.5 under a small hash-table field) over
    the [-1,1)^3 lattice in 128^3 blocks; vertices and faces of every block bit-exact
    against the oracle's block loop on the same stored volume (world coordinates)."""
    from neuralangelo_amd.model import Model, default_config
    dev = torch.device("cuda", 0)
    torch.manual_seed(0)
    m = Model(default_config(log2_T=22, F=8), compute_dtype=torch.float32).to(dev)
    m.neural_sdf.active_levels = m.neural_sdf.anneal_levels = 16
    res, block_res = 1024, 128
    intv = 2.0 / res
    axes = gmesh.lattice_axes(((-1.0, 1.0),) * 3, intv)
    assert [len(a) for a in axes] == [res] * 3
    with torch.no_grad():
        vol = gmesh.lattice_sdf(lambda x: -m.neural_sdf.sdf_nograd(x), axes, dev,
                                max_points=1 << 23)
    del m
    V, F = mcubes.lattice_marching_cubes(vol, block_res, 0.0, axes=[a.to(dev) for a in axes],
                                         intv=intv)
    V, F = V.cpu().numpy(), F.cpu().numpy()
    host = vol.cpu().numpy()
    del vol
    torch.cuda.empty_cache()
    print(f"1024^3 lattice: {V.shape[0]} vertices, {F.shape[0]} faces", flush=True)
    Vr, Fr = omesh.lattice_marching_cubes(host, block_res, 0.0, [a.numpy() for a in axes], intv,
                                          threads=16)
    assert V.shape == Vr.shape and F.shape == Fr.shape
    assert np.array_equal(F, Fr)
    assert np.array_equal(V, Vr)
    assert F.shape[0] > 1_000_000
