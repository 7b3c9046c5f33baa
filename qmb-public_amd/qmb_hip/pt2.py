"""
Epstein-Nesbet second-order perturbation step on the device (qmb/pert.py:37-63).

From a variational state psi on configs:
    E0 = Re(<psi|H|psi> / <psi|psi>)
    other = find_relative(configs, psi, count, configs)        (count = len(configs) in the reference)
    E2 = sum_a |<a|H|psi>|^2 / <psi|psi> / (E0 - H_aa)          over a in other
The three operator calls are the library's: E0 through the prepared operator on the config set
(structured path), <a|H|psi> through the structured path on configs joined with `other`
(prepared.apply_across), H_aa through diagonal_term.  The term table is
compiled once per Hamiltonian and shared by the three.  The reference's loop that would double
`count` runs its body once (it breaks unconditionally), so one find_relative call is made here too.
"""

from __future__ import annotations

import dataclasses

import torch

from . import prepared as pp
from .krylov import WithinOperator, _hamiltonian_of


@dataclasses.dataclass
class Perturbation:
    """energy0: the variational energy; energy2: the second-order correction (what the reference
    logs as "Correct energy"); other: the external configs the correction runs over."""

    energy0: float
    energy2: float
    other: torch.Tensor

    @property
    def energy(self) -> float:
        """E0 + E2, the second-order estimate of the energy."""
        return self.energy0 + self.energy2


def _apply_foreign(model: object, configs: torch.Tensor, psi: torch.Tensor, targets_configs: torch.Tensor) -> torch.Tensor:
    """<a|H|psi> for the rows a of targets_configs (outside the set of configs)."""
    ham = _hamiltonian_of(model)
    if ham is None or not configs.is_cuda:
        return model.apply_within(configs, psi, targets_configs)  # type: ignore[attr-defined]
    nq = configs.size(1)
    terms = ham.prepared_terms(configs.device, nq)
    vec = torch.view_as_real(psi.resolve_conj().contiguous())
    out = pp.apply_across(terms, configs.contiguous(), vec, targets_configs, pp.MODE_FORWARD)
    if out is None:
        # find_relative returns distinct rows
        targets = pp.PreparedTargets(targets_configs, nq, distinct=True)
        out = pp.apply(terms, targets, configs.contiguous(), vec, pp.MODE_FORWARD, pairs=False)
    return torch.view_as_complex(out)


def perturbation(model: object, configs: torch.Tensor, psi: torch.Tensor, count: int | None = None) -> Perturbation:
    """qmb/pert.py:37-63 for (configs, psi); `count` defaults to len(configs), as there."""
    if psi.dtype != torch.complex128 or psi.shape != (configs.size(0),):
        raise RuntimeError("psi must be complex128 with one entry per config")
    configs = configs.contiguous()
    norm2 = (psi.conj() @ psi).real
    energy0 = float(((psi.conj() @ WithinOperator(model, configs)(psi)) / norm2).real)
    number = configs.size(0) if count is None else int(count)
    other = model.find_relative(configs, psi, number, configs)  # type: ignore[attr-defined]
    hpsi = _apply_foreign(model, configs, psi, other)
    energy2_num = (hpsi.conj() * hpsi).real / norm2
    energy2_den = energy0 - model.diagonal_term(other).real  # type: ignore[attr-defined]
    energy2 = float((energy2_num / energy2_den).sum())
    return Perturbation(energy0, energy2, other)
