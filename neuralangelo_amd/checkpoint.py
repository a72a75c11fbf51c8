"""Checkpoint interchange with the reference trainer (imaginaire/trainers/base.py:575-649).

A reference checkpoint is ``torch.save``-d as ``{"model": state_dict, "optim": ...,
"sched": ..., "epoch": int, "iteration": int}``; the model keys carry one ``module.``
prefix per wrapper around the network (DDP or the dummy WrappedModel,
imaginaire/trainers/utils/get_trainer.py:57-88, plus one more with the EMA model
average).  The network keys themselves follow the modules: ``s_var``,
``neural_sdf.tcnn_encoding.params`` (tcnn's flat table layout),
``neural_sdf.mlp.linears.<i>.weight_g / weight_v / bias`` (weight-normed,
projects/neuralangelo/utils/mlp.py:20-52), ``neural_sdf.mlp.linear_sdf.weight / bias``
and ``neural_rgb.mlp.linears.<i>.*`` (projects/nerf/utils/nerf_util.py:158-196) -- the
same names and shapes as this package's Model.  Files are read with
``torch.load(weights_only=True)``, which executes nothing from the file.
"""
import torch

_WRAPPER_PREFIXES = ("module.",)
_UNSUPPORTED = ("background_nerf.", "appear_embed.")
# the EMA wrapper (imaginaire/models/utils/model_average.py:53-72) keeps the averaged
# copy under "_averaged_model." and a step counter buffer next to the live network
_EMA_PREFIX = "_averaged_model."
_EMA_COUNTER = "num_updates_tracked"


def _strip(key):
    changed = True
    while changed:
        changed = False
        for p in _WRAPPER_PREFIXES:
            if key.startswith(p):
                key = key[len(p):]
                changed = True
    return key


def network_state(state, known=(), averaged=False):
    """Reference checkpoint dict (or its ``model`` entry) -> network keys without the
    wrapper prefixes, and the list of keys this package does not model (background
    NeRF weights when the model has no background, appearance embedding, the EMA
    copy and its counter).  known: network keys of the target model, which are always
    kept.  averaged: take the network from the EMA copy instead of the live weights."""
    if "model" in state and isinstance(state["model"], dict):
        state = state["model"]
    known = set(known)
    out, skipped = {}, []
    for k, v in state.items():
        name = _strip(k)
        ema = name.startswith(_EMA_PREFIX)
        if ema:
            name = _strip(name[len(_EMA_PREFIX):])
        if name == _EMA_COUNTER or ema != averaged:
            skipped.append(k)
            continue
        if name not in known and any(name.startswith(u) for u in _UNSUPPORTED):
            skipped.append(k)
            continue
        out[name] = v
    return out, skipped


def load_reference_checkpoint(model, path_or_state, strict=True, averaged=False):
    """Load a reference checkpoint (path or already-loaded dict) into ``model``.
    Returns (iteration or None, skipped keys).  strict: every network key of the model
    must be present, with its shape (a table of another size raises).  averaged: load
    the EMA copy of an EMA-wrapped checkpoint instead of the live weights."""
    state = path_or_state
    if not isinstance(state, dict):
        state = torch.load(path_or_state, map_location="cpu", weights_only=True)
    iteration = state.get("iteration") if isinstance(state.get("iteration"), int) else None
    own = model.state_dict()
    net, skipped = network_state(state, known=own.keys(), averaged=averaged)
    missing = [k for k in own if k not in net]
    unexpected = [k for k in net if k not in own]
    if strict and (missing or unexpected):
        raise RuntimeError(f"checkpoint does not match the model: missing {missing}, "
                           f"unexpected {unexpected}")
    for k, v in net.items():
        if k in own and tuple(v.shape) != tuple(own[k].shape):
            raise RuntimeError(f"{k}: checkpoint shape {tuple(v.shape)} != model shape "
                               f"{tuple(own[k].shape)}")
    # load_state_dict copies in place; the encoder's BF16 table copy is keyed on the
    # parameter's version counter and is rebuilt on next use
    model.load_state_dict({k: v.to(own[k].dtype) for k, v in net.items() if k in own},
                          strict=strict)
    return iteration, skipped


def reference_checkpoint(model, iteration=0, epoch=0, wrappers=1):
    """``model`` as a reference-format checkpoint dict (``module.`` prefix per wrapper),
    loadable by the reference trainer's ``checkpointer.load(..., load_opt=False,
    load_sch=False)`` as used by extract_mesh.py."""
    pre = "module." * wrappers
    return {"model": {pre + k: v.detach().cpu() for k, v in model.state_dict().items()},
            "epoch": int(epoch), "iteration": int(iteration)}
