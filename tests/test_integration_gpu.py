"""The ctypes binding shown in the integration guide runs as written and agrees with the modules."""
import os
import re

import pytest
import torch

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_integration_snippet_runs(cuda, monkeypatch):
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    code = re.search(r"```python\n(.*?)```", text, re.S).group(1)
    monkeypatch.chdir(REPO)
    ns = {}
    exec(compile(code, "INTEGRATION.md", "exec"), ns)  # noqa: S102
    torch.cuda.synchronize()
    from neuralangelo_amd import mcubes
    from neuralangelo_amd.encoding import Encoding
    enc = Encoding(3, dict(otype="HashGrid", n_levels=16, n_features_per_level=8,
                           log2_hashmap_size=22, base_resolution=32,
                           per_level_scale=2 ** 0.4)).cuda()
    assert enc.n_params == ns["n_params"]
    with torch.no_grad():
        enc.params.copy_(ns["table"])
        ref = enc(ns["x"])
    assert torch.equal(ref, ns["out"])
    V, F = mcubes.marching_cubes_gpu(ns["vol"], 0.0)
    assert torch.equal(V, ns["V"]) and torch.equal(F, ns["F"])
