"""The shipped closure weights as a file (CPU): they load without pickle,
hold one finite value per network parameter with the padded channels
untouched, and their metadata records the `closure_run train` defaults, so
that the documented command (the defaults and `--seed 1`) names their
recipe."""
from __future__ import annotations

import json

import numpy as np

import oracle.closure_oracle as co
from psyclonebench_amd import closure_run as cr

TAPS = 25


def _weights():
    z = np.load(cr.DEFAULT_WEIGHTS, allow_pickle=False)
    return z["params"], json.loads(str(z["meta"]))


def test_shipped_weights_are_finite_with_padded_channels_zero():
    params, _ = _weights()
    sizes = [TAPS * ci * co_ + co_ for ci, co_ in co.LAYERS]
    assert params.shape == (sum(sizes),)
    assert np.all(np.isfinite(params))
    off = np.concatenate([[0], np.cumsum(sizes)])
    # layer 0: 4 of its 16 input channels carry data
    ci0, co0 = co.LAYERS[0]
    w0 = params[off[0]:off[0] + TAPS * ci0 * co0].reshape(TAPS, co0, ci0)
    assert not np.any(w0[:, :, 4:]) and np.any(w0[:, :, :4])
    # output layer: 3 of its 16 output channels are used
    ci3, co3 = co.LAYERS[3]
    w3 = params[off[3]:off[3] + TAPS * ci3 * co3].reshape(TAPS, co3, ci3)
    b3 = params[off[3] + TAPS * ci3 * co3:off[4]]
    assert not np.any(w3[:, 3:, :]) and not np.any(b3[3:]) and np.any(w3[:, :3, :])


def test_shipped_weights_record_the_train_defaults():
    _, meta = _weights()
    d = cr.build_parser().parse_args(["train"])
    for key, arg in (("n_fine", "n"), ("factor", "factor"), ("bathy", "bathy"),
                     ("coarse_steps_per_seed", "steps"), ("lr", "lr"), ("lr_min", "lr_min"),
                     ("rollout", "rollout"), ("guard_rel", "guard_rel"), ("clip", "clip"),
                     ("max_spin", "max_spin"), ("ema", "ema"), ("ema_every", "ema_every"),
                     ("ema_start", "ema_start"), ("reset_spike", "reset_spike"),
                     ("val_bathy", "val_bathy"), ("val_every", "val_every"),
                     ("val_steps", "val_steps")):
        assert meta[key] == getattr(d, arg), key
    assert tuple(meta["loss_weights"]) == tuple(d.loss_weights)
    norm = meta["norm"]
    assert norm["guard_max_out"] == d.guard_max_out
    assert norm["ssh_cells"] == d.ssh_cells
    assert norm["gate_eps"] == d.gate_eps
