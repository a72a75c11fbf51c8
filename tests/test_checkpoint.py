"""Reference-format checkpoints (imaginaire/trainers/base.py:575-649): key names and
shapes of the network, wrapper prefixes, skipped background keys, and a round trip."""
import pytest
import torch

from neuralangelo_amd.checkpoint import (load_reference_checkpoint, network_state,
                                         reference_checkpoint)
from neuralangelo_amd.model import Model, default_config
from tests.helpers import randomize_model


def _model(seed, log2_T=14):
    torch.manual_seed(seed)
    m = Model(default_config(log2_T=log2_T, F=4))
    randomize_model(m, seed)
    return m


def test_network_keys_follow_the_reference_modules():
    """mlp.py:20-52 (SDF MLP [3+L*F, 256, 256] + linear_sdf), nerf_util.py:158-196
    (colour MLP 278 -> 4 x 256 -> 3), weight_norm's weight_g [out,1] / weight_v [out,in],
    tcnn's flat params, s_var."""
    m = _model(0)
    sd = m.state_dict()
    L, F = 16, 4
    want = {"s_var": (), "neural_sdf.mlp.linears.0.weight_v": (256, 3 + L * F),
            "neural_sdf.mlp.linears.0.weight_g": (256, 1), "neural_sdf.mlp.linears.0.bias": (256,),
            "neural_sdf.mlp.linears.1.weight_v": (256, 256), "neural_sdf.mlp.linear_sdf.weight": (1, 256),
            "neural_sdf.mlp.linear_sdf.bias": (1,), "neural_rgb.mlp.linears.0.weight_v": (256, 278),
            "neural_rgb.mlp.linears.3.weight_v": (256, 256), "neural_rgb.mlp.linears.4.weight_v": (3, 256),
            "neural_rgb.mlp.linears.4.bias": (3,)}
    for k, shape in want.items():
        assert tuple(sd[k].shape) == shape, k
    enc = m.neural_sdf.tcnn_encoding
    assert tuple(sd["neural_sdf.tcnn_encoding.params"].shape) == (enc.n_params,)


@pytest.mark.parametrize("wrappers", [1, 2])
def test_round_trip_through_a_reference_file(tmp_path, wrappers):
    src, dst = _model(1), _model(2)
    ck = reference_checkpoint(src, iteration=12345, wrappers=wrappers)
    # a reference run with the (default) background NeRF also holds its weights
    ck["model"]["module." * wrappers + "background_nerf.mlp_feat.linears.0.weight"] = torch.zeros(3, 3)
    ck["optim"] = {"state": {}, "param_groups": []}
    path = tmp_path / "ckpt.pt"
    torch.save(ck, path)
    it, skipped = load_reference_checkpoint(dst, str(path))
    assert it == 12345 and len(skipped) == 1 and "background_nerf" in skipped[0]
    for k, v in src.state_dict().items():
        assert torch.equal(dst.state_dict()[k], v), k


def test_mismatched_table_raises():
    small, big = _model(3, log2_T=12), _model(4, log2_T=14)
    with pytest.raises(RuntimeError):
        load_reference_checkpoint(big, reference_checkpoint(small))


def test_network_state_strips_prefixes_only():
    net, skipped = network_state({"model": {"module.module.s_var": torch.tensor(1.0),
                                            "module.appear_embed.weight": torch.zeros(2)}})
    assert list(net) == ["s_var"] and skipped == ["module.appear_embed.weight"]


def test_background_weights_load_when_the_model_has_them(tmp_path):
    torch.manual_seed(5)
    src = Model(default_config(log2_T=12, F=4, background=True))
    torch.manual_seed(6)
    dst = Model(default_config(log2_T=12, F=4, background=True))
    path = tmp_path / "bg.pt"
    torch.save(reference_checkpoint(src, iteration=7), path)
    it, skipped = load_reference_checkpoint(dst, str(path))
    assert it == 7 and skipped == []
    k = "background_nerf.mlp_feat.linears.4.weight"
    assert tuple(dst.state_dict()[k].shape) == (256, 340)
    for name, v in src.state_dict().items():
        assert torch.equal(dst.state_dict()[name], v), name


def test_ema_wrapped_checkpoint():
    """An EMA-wrapped reference checkpoint (model_average.py:53-72): live weights under
    module.module.*, the averaged copy under module._averaged_model.*, and the
    module.num_updates_tracked counter.  The live weights load by default, the averaged
    copy with averaged=True; the counter is skipped either way."""
    live, avg, dst = _model(5), _model(6), _model(7)
    state = {"model": {}, "iteration": 1234}
    for k, v in live.state_dict().items():
        state["model"]["module.module." + k] = v.clone()
    for k, v in avg.state_dict().items():
        state["model"]["module._averaged_model." + k] = v.clone()
    state["model"]["module.num_updates_tracked"] = torch.tensor(99)
    it, skipped = load_reference_checkpoint(dst, state)
    assert it == 1234 and "module.num_updates_tracked" in skipped
    for k, v in dst.state_dict().items():
        assert torch.equal(v, live.state_dict()[k]), k
    _, skipped = load_reference_checkpoint(dst, state, averaged=True)
    for k, v in dst.state_dict().items():
        assert torch.equal(v, avg.state_dict()[k]), k
    assert all("_averaged_model." not in k for k in skipped if k != "module.num_updates_tracked")
