"""
Packing of small unsigned ints into bytes: the configuration format of the operators.

Same contract as qmb/bitspack.py:8-94: ``pack_int(t, size)`` packs the last dimension of a uint8
tensor holding `size`-bit values (size in 1, 2, 4, 8) into bytes, element e of a group landing at
bit offset (e % (8 // size)) * size of byte e // (8 // size) (LSB first, zero padding at the end);
``unpack_int(t, size, last_dim)`` inverts it.  For size 1 this is the config layout: site s is bit
s % 8 of byte s // 8.
"""

from __future__ import annotations

import torch


def pack_int(tensor: torch.Tensor, size: int) -> torch.Tensor:
    """uint8 [..., n] of `size`-bit values -> uint8 [..., ceil(n * size / 8)]."""
    if tensor.dtype != torch.uint8:
        raise TypeError("pack_int expects a uint8 tensor")
    if size not in (1, 2, 4, 8):
        raise ValueError("size must be 1, 2, 4 or 8")
    per_byte = 8 // size
    n = tensor.shape[-1]
    pad = (-n) % per_byte
    if pad:
        tensor = torch.nn.functional.pad(tensor, (0, pad))
    grouped = tensor.reshape(*tensor.shape[:-1], -1, per_byte).to(torch.int32)
    shifts = torch.arange(0, 8, size, device=tensor.device, dtype=torch.int32)
    packed = (grouped << shifts).sum(dim=-1)
    return packed.to(torch.uint8)


def unpack_int(tensor: torch.Tensor, size: int, last_dim: int) -> torch.Tensor:
    """Inverse of pack_int: uint8 [..., m] -> uint8 [..., last_dim]."""
    if tensor.dtype != torch.uint8:
        raise TypeError("unpack_int expects a uint8 tensor")
    if size not in (1, 2, 4, 8):
        raise ValueError("size must be 1, 2, 4 or 8")
    shifts = torch.arange(0, 8, size, device=tensor.device, dtype=torch.int32)
    mask = (1 << size) - 1
    values = (tensor.to(torch.int32).unsqueeze(-1) >> shifts) & mask
    return values.reshape(*tensor.shape[:-1], -1)[..., :last_dim].to(torch.uint8)
