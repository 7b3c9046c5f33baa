"""
Host-side argument rules of the prepared objects and drivers (no GPU needed):

  - a core index built under a max_occupied bound that some row exceeds is an error, never a
    silent result (check_core_overflow raises with the bound in the message; PreparedTargets.cores
    calls it whenever the bound came from the caller);
  - vmc_cycle refuses relative_count together with the Rayleigh-quotient step, which has no use
    for an extended target set.
"""

from __future__ import annotations

import pytest
import torch

from qmb_hip import prepared as pp
from qmb_hip import vmc


def test_core_overflow_raises() -> None:
    pp.check_core_overflow(0, 14)
    with pytest.raises(RuntimeError, match="max_occupied=12"):
        pp.check_core_overflow(3, 12)


def test_vmc_cycle_rejects_relative_count_with_rayleigh() -> None:
    class NoNetwork(torch.nn.Module):
        def generate_unique(self, count: int) -> None:  # pragma: no cover - never reached
            raise AssertionError("vmc_cycle must reject the arguments before sampling")

    net = NoNetwork()
    opt = torch.optim.SGD([torch.nn.Parameter(torch.zeros(1))], lr=0.1)
    with pytest.raises(ValueError, match="relative_count"):
        vmc.vmc_cycle(object(), net, opt, sampling_count=10, local_step=1, relative_count=20, rayleigh=True)
