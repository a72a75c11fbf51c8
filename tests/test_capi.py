"""C-ABI library: loads, exports every declared symbol, host-side logic matches the oracle.

No GPU work is launched here (CPU-only suite)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from neuralangelo_amd import _lib
from neuralangelo_amd.encoding import growth_rate, reference_resolutions
from oracle import hashgrid as hg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(na_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 5
    for s in syms:
        assert hasattr(lib, s), s


def test_python_binds_every_declared_symbol():
    assert sorted(declared_symbols()) == sorted(_lib.exported_symbols())


def test_version_and_error_string():
    lib = _lib.load()
    assert lib.na_version() >= 1
    desc = _lib.GridDesc()
    st = lib.na_hashgrid_layout(16, 3, 22, 32, 1.3, ctypes.byref(desc))
    assert st == _lib.NA_ERR_ARGUMENT if hasattr(_lib, "NA_ERR_ARGUMENT") else st < 0
    assert "n_features" in _lib.last_error()


def test_adamw_multi_and_wgrad_workspace_argument_checks():
    """Argument checks of the multi-tensor AdamW and the weight-gradient workspace
    query: they return before any GPU call."""
    lib = _lib.load()
    arr = (_lib.AdamwTensor * 1)()
    st = lib.na_adamw_step_multi(arr, _lib.NA_ADAMW_MULTI_MAX + 1, 1e-3, 0.9, 0.999, 1e-8, 1e-2,
                                 1, None)
    assert st == _lib.NA_ERR_ARGUMENT and "NA_ADAMW_MULTI_MAX" in _lib.last_error()
    st = lib.na_adamw_step_multi(arr, 1, 1e-3, 0.9, 0.999, 1e-8, 1e-2, 0, None)
    assert st == _lib.NA_ERR_ARGUMENT and "step" in _lib.last_error()
    arr[0].n = -1
    st = lib.na_adamw_step_multi(arr, 1, 1e-3, 0.9, 0.999, 1e-8, 1e-2, 1, None)
    assert st == _lib.NA_ERR_ARGUMENT and "size" in _lib.last_error()
    assert lib.na_adamw_step_multi(None, 0, 1e-3, 0.9, 0.999, 1e-8, 1e-2, 1, None) == 0
    assert lib.na_gemm_wgrad_workspace(None) == _lib.NA_ERR_ARGUMENT
    a = _lib.GemmArgs()
    a.M, a.N, a.K = 256, 256, 4096
    a.trans_a, a.trans_b, a.in_dtype = 1, 1, _lib.NA_F32  # fp32 operands: not the bf16 path
    a.epilogue, a.c_dtype = _lib.NA_EPI_ATOMIC, _lib.NA_F32
    assert lib.na_gemm_wgrad_workspace(ctypes.byref(a)) == 0


@pytest.mark.parametrize("log2_T,F", [(22, 8), (19, 4), (12, 2), (16, 1)])
def test_layout_matches_oracle(log2_T, F):
    g = growth_rate(5, 11, 16)
    desc = _lib.GridDesc()
    n_params = _lib.load().na_hashgrid_layout(16, F, log2_T, 32, g, ctypes.byref(desc))
    lt = hg.level_table(16, F, log2_T, 32, g)
    assert n_params == lt["n_params"]
    for lv in range(16):
        assert np.float32(desc.scale[lv]) == lt["scale"][lv]
        assert desc.resolution[lv] == lt["res"][lv]
        assert desc.offset[lv] == lt["offset"][lv]
    assert desc.offset[16] == lt["offset"][16]


def test_reference_resolutions_for_epsilon():
    # modules.py:51-54 resolutions (used only for the numerical-gradient epsilon)
    g = growth_rate(5, 11, 16)
    r = reference_resolutions(5, 16, g)
    assert r == [33, 43, 56, 74, 98, 128, 169, 223, 295, 389, 512, 676, 892, 1177, 1553, 2048]


def test_wgrad_static_argument_checks_and_limit():
    """na_gemm_wgrad_static returns before any GPU call for a NULL or non-weight-gradient
    product; its size limit is the header's constant."""
    lib = _lib.load()
    assert lib.na_gemm_wgrad_static(None, None) == _lib.NA_ERR_ARGUMENT
    a = _lib.GemmArgs()
    a.M, a.N, a.K = 256, 256, 4096
    a.trans_a, a.trans_b, a.in_dtype = 1, 1, _lib.NA_BF16
    a.epilogue, a.c_dtype = _lib.NA_EPI_STORE, _lib.NA_F32  # store epilogue: not this path
    assert lib.na_gemm_wgrad_static(ctypes.byref(a), None) == _lib.NA_ERR_ARGUMENT
    assert "weight-gradient" in _lib.last_error()
    text = open(os.path.join(REPO, "include", "neuralangelo_amd.h")).read()
    m = re.search(r"#define NA_WGRAD_STATIC_MAX_FLOATS (\d+)", text)
    assert m and int(m.group(1)) == _lib.NA_WGRAD_STATIC_MAX_FLOATS
