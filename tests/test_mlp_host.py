"""
Host-side checks of the config producer (qmb_hip.mlp) that need no GPU: parameter and buffer
names match the reference module's (qmb/mlp.py:61-110), so state dicts move between them; the
ordering buffers are inverse permutations; the network kernels reject bad arguments before any
device work; the forward refuses CPU tensors (the product path has no CPU fallback).
"""

from __future__ import annotations

import ctypes

import pytest
import torch

from qmb_hip import _lib
from qmb_hip import mlp


def test_state_dict_layout() -> None:
    net = mlp.WaveFunctionElectronUpDown(double_sites=8, physical_dim=2, is_complex=True, spin_up=2, spin_down=1, hidden_size=(5, 3),
                                         ordering=[1, 3, 0, 2])
    sd = net.state_dict()
    assert "amplitude.0.model.0.bias" in sd and "amplitude.0.model.0.weight" not in sd  # zero-input first layer
    assert sd["amplitude.2.model.0.weight"].shape == (5, 4)
    assert sd["amplitude.3.model.2.weight"].shape == (3, 5)
    assert sd["amplitude.3.model.4.weight"].shape == (4, 3)
    assert sd["phase.model.0.weight"].shape == (5, 8) and sd["phase.model.4.weight"].shape == (1, 3)
    assert sd["dummy_param"].numel() == 0
    assert torch.equal(sd["ordering"][sd["ordering_reversed"]], torch.arange(4))
    again = mlp.WaveFunctionElectronUpDown(double_sites=8, physical_dim=2, is_complex=True, spin_up=2, spin_down=1, hidden_size=(5, 3),
                                           ordering=[1, 3, 0, 2])
    again.load_state_dict(sd)
    with pytest.raises(ValueError):
        mlp.WaveFunctionElectronUpDown(double_sites=8, physical_dim=2, is_complex=True, spin_up=2, spin_down=1, hidden_size=(5,),
                                       ordering=[0, 0, 1, 2])
    with pytest.raises(RuntimeError):
        net(torch.zeros((3, 1), dtype=torch.uint8))


def test_network_kernels_reject_bad_arguments() -> None:
    lib = _lib.library()
    buf = ctypes.create_string_buffer(4096)
    p = ctypes.cast(buf, ctypes.c_void_p)
    # sites out of range, electron counts that do not fit, mismatched gradient pointers, bad dtype
    assert lib.qmb_network_amplitude(1, 4, 0, 0, 0, p, 1, p, p, p, None, None, None) != 0
    assert lib.qmb_network_amplitude(1, 4, 4, 5, 0, p, 1, p, p, p, None, None, None) != 0
    assert lib.qmb_network_amplitude(1, 4, 4, 2, 2, p, 1, p, p, p, p, None, None) != 0
    assert b"together" in lib.qmb_last_error()
    assert lib.qmb_network_amplitude(2, 4, 4, 2, 2, p, 1, p, p, p, None, None, None) != 0
    assert lib.qmb_network_inputs(0, 4, 9, p, 2, p, p, None) != 0  # 9 sites do not fit 2 bytes
    assert lib.qmb_sampler_children(0, 4, 4, 4, 2, 2, p, 1, p, p, p, None, 0, p, p, None) != 0  # site == n_sites
    assert lib.qmb_sampler_children(0, 4, 1, 40, 2, 2, p, 1, p, p, p, None, 0, p, p, None) != 0  # 80 bits in one word
    assert lib.qmb_sampler_select(0, 9, p, 2, 1, p, 1, p, p, p, p, p, None) != 0  # more than 4 children per parent
    assert lib.qmb_sampler_prefix_configs(4, 8, p, 1, p, 1, p, None) != 0  # 16 bits in one byte
    # nothing to do is not an error
    assert lib.qmb_sampler_children(0, 0, 0, 4, 2, 2, None, 1, None, None, None, None, 0, None, None, None) == 0
