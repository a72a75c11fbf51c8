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
