"""Argument checks of the sorted (deterministic) table-gradient entry points; every
case here is rejected on the host before any kernel runs, so it needs no GPU."""
import ctypes

import pytest

from neuralangelo_amd import _lib
from neuralangelo_amd.encoding import Encoding


@pytest.fixture(scope="module")
def enc():
    return Encoding(3, dict(otype="HashGrid", n_levels=8, n_features_per_level=4,
                            log2_hashmap_size=12, base_resolution=8, per_level_scale=1.6))


def test_workspace_size(enc):
    lib = _lib.load()
    for n, evals in ((1, 1), (1000, 1), (1000, 5), (123457, 5)):
        size = lib.na_hashgrid_sorted_workspace(ctypes.byref(enc.desc), n, evals)
        if size < 0 and "device" in _lib.last_error():
            continue  # the radix sort sizes its storage for the device it will run on
        # keys and ids, twice, for n * evals * 8 contributions, plus the sort's storage
        assert size >= 4 * 4 * n * evals * 8 and size % 256 == 0, (n, evals, size)
    assert lib.na_hashgrid_sorted_workspace(ctypes.byref(enc.desc), 0, 5) >= 0
    assert lib.na_hashgrid_sorted_workspace(ctypes.byref(enc.desc), 100, 3) < 0
    assert "evals" in _lib.last_error()
    assert lib.na_hashgrid_sorted_workspace(ctypes.byref(enc.desc), 1 << 26, 5) < 0
    assert "too large" in _lib.last_error()


def test_entry_points_reject_bad_arguments(enc):
    lib = _lib.load()
    d = ctypes.byref(enc.desc)
    fake = ctypes.c_void_p(1 << 40)  # never dereferenced: every call below fails its checks
    need = 1 << 20
    # missing workspace
    st = lib.na_hashgrid_bwd_sorted(d, fake, 100, fake, _lib.NA_F32, fake, 8, None, need, None)
    assert st < 0 and "workspace" in _lib.last_error()
    # misaligned workspace
    st = lib.na_sdf_taps_bwd_sorted(d, fake, 100, -2.0, 2.0, 1e-3, fake, fake, None, _lib.NA_F32,
                                    32, fake, 8, 0, ctypes.c_void_p((1 << 40) + 16), need, None)
    assert st < 0 and "aligned" in _lib.last_error()
    # the shared row belongs to the increment layout only
    st = lib.na_sdf_taps_bwd_sorted(d, fake, 100, -2.0, 2.0, 1e-3, fake, fake, fake, _lib.NA_F32,
                                    32, fake, 8, 0, fake, need, None)
    assert st < 0 and "dy_shared" in _lib.last_error()
    # volume range and leading dimension
    st = lib.na_sdf_input_bwd_sorted(d, fake, 100, 1.0, 1.0, fake, _lib.NA_F32, 32, fake, 8, fake,
                                     need, None)
    assert st < 0 and "vmax" in _lib.last_error()
    st = lib.na_sdf_input_bwd_sorted(d, fake, 100, -1.0, 1.0, fake, _lib.NA_F32, 16, fake, 8, fake,
                                     need, None)
    assert st < 0 and "ld" in _lib.last_error()
    # nothing to do: n = 0 or no active level
    assert lib.na_hashgrid_bwd_sorted(d, None, 0, None, _lib.NA_F32, None, 8, None, 0, None) == 0
    assert lib.na_hashgrid_bwd_sorted(d, fake, 100, fake, _lib.NA_F32, fake, 0, None, 0, None) == 0
