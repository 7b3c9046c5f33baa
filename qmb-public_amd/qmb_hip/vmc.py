"""
Variational Monte Carlo on the device (qmb/vmc.py:80-125): the energy loss of a network state on a
sampled config set, its gradient, and one optimisation cycle.

The reference's cycle samples configs with network.generate_unique (vmc.py:83), extends them with
find_relative up to relative_count rows (vmc.py:87-92), and runs `local_step` optimizer steps of the
closure (vmc.py:104-116):
    psi_src = network(configs_src)
    no_grad: h_psi = apply_within(configs_dst, network(configs_dst), configs_src)
    energy = Re(psi_src^H h_psi / psi_src^H psi_src.detach());  energy.backward()
Two losses are offered:
  - `closure_energy` is that closure as written.  Its gradient flows through the conjugated psi_src
    only; for a Hermitian H and configs_dst = configs_src it is half the gradient of the Rayleigh
    quotient (the same direction).
  - `energy_backward` is the Rayleigh quotient Re <psi|H|psi> / <psi|psi> on one set, differentiated
    in full: the backward of H psi is the adjoint kernel (autograd.ApplyWithin), so no property of H
    is assumed and H is never materialised.
`energy_backward` pulls the gradient back through the network in row chunks: the loss is formed on
psi detached from the network, and its gradient with respect to psi is sent back one chunk of
configs at a time.  That is exact (the loss depends on the parameters only through psi) and bounds
the activation memory by the chunk: the electron up/down network keeps sites x width floats per
config and layer, 78 KB per config and layer at 38 sites and width 512.
"""

from __future__ import annotations

import dataclasses
import typing

import torch

from . import autograd as ag
from . import prepared as pp
from .krylov import _hamiltonian_of


def _terms(model: object, configs: torch.Tensor) -> pp.PreparedTerms:
    ham = _hamiltonian_of(model)
    if ham is None or not configs.is_cuda:
        raise RuntimeError("the energy loss needs a qmb_hip Hamiltonian (or a model holding one) and configs on a HIP device")
    return ham.prepared_terms(configs.device, configs.size(1))


class EnergySet:
    """One sampled config set with its prepared index, built once and used by every local step of a
    cycle (the reference re-sorts configs_j inside each apply_within call).

    configs: distinct uint8 [n, n_qubytes] rows (generate_unique returns distinct rows).
    max_occupied: a bound on the occupied sites of every row (saves one pass when known)."""

    def __init__(self, model: object, configs: torch.Tensor, max_occupied: int | None = None) -> None:
        if configs.dtype != torch.uint8 or configs.dim() != 2:
            raise RuntimeError("configs must be a uint8 [n, n_qubytes] tensor.")
        self.configs = configs.contiguous()
        self.terms = _terms(model, self.configs)
        self.targets = pp.PreparedTargets(self.configs, self.configs.size(1), distinct=True, max_occupied=max_occupied)

    def energy(self, psi: torch.Tensor) -> torch.Tensor:
        """Re <psi|H|psi> / <psi|psi>, differentiable in psi (complex128 [n]) through the adjoint."""
        return ag.energy(psi, self.configs, self.terms, self.targets)


def network_forward(network: torch.nn.Module, configs: torch.Tensor, chunk: int) -> torch.Tensor:
    """psi = network(configs) without gradients, complex128, one chunk of rows at a time."""
    n = configs.size(0)
    step = max(1, int(chunk))
    with torch.no_grad():
        parts = [network(configs[lo:min(n, lo + step)]) for lo in range(0, n, step)]
    return (torch.cat(parts) if len(parts) != 1 else parts[0]).to(torch.complex128)


def network_backward(network: torch.nn.Module, configs: torch.Tensor, grad_psi: torch.Tensor, chunk: int) -> None:
    """Accumulate into the parameters' .grad the pull-back of grad_psi (the gradient of a real loss
    with respect to psi = network(configs), in PyTorch's convention for complex tensors), one chunk
    of rows at a time."""
    n = configs.size(0)
    step = max(1, int(chunk))
    for lo in range(0, n, step):
        hi = min(n, lo + step)
        psi = network(configs[lo:hi])
        psi.backward(grad_psi[lo:hi])


def energy_backward(network: torch.nn.Module, energy_set: EnergySet, chunk: int = 65536) -> float:
    """The Rayleigh quotient of network(configs) on the set; adds its gradient to the network's
    parameters (call optimizer.zero_grad() first, as the reference's closure does) and returns it."""
    psi = network_forward(network, energy_set.configs, chunk).requires_grad_(True)
    energy = energy_set.energy(psi)
    energy.backward()
    if psi.grad is None:
        raise RuntimeError("the energy did not reach psi: no gradient to pull back through the network")
    network_backward(network, energy_set.configs, psi.grad, chunk)
    return float(energy.detach())


def closure_energy(model: object, network: torch.nn.Module, configs_src: torch.Tensor, configs_dst: torch.Tensor) -> torch.Tensor:
    """The reference's loss (qmb/vmc.py:104-116): psi_src keeps its graph, psi_dst and H psi_dst
    do not.  Returns the real energy tensor; the caller runs .backward()."""
    psi_src = network(configs_src)
    with torch.no_grad():
        psi_dst = network(configs_dst)
        hamiltonian_psi_dst = model.apply_within(configs_dst, psi_dst, configs_src)  # type: ignore[attr-defined]
    num = psi_src.conj() @ hamiltonian_psi_dst
    den = psi_src.conj() @ psi_src.detach()
    return (num / den).real


@dataclasses.dataclass
class CycleResult:
    """configs: the cycle's sampled set; energies: the energy of every local step."""

    configs: torch.Tensor
    energies: list[float]


def vmc_cycle(model: object, network: torch.nn.Module, optimizer: torch.optim.Optimizer, *, sampling_count: int,
              local_step: int, relative_count: int = 0, rayleigh: bool = True, chunk: int = 65536,
              seed: int | None = None, callback: typing.Callable[[int, float], None] | None = None) -> CycleResult:
    """One cycle of qmb/vmc.py:80-125: sample, (optionally) extend with find_relative, then
    `local_step` optimizer steps.  rayleigh=True steps on the full Rayleigh quotient over the sampled
    set (energy_backward, the adjoint path); False on the reference's closure over (configs_src,
    configs_dst).  relative_count applies to the closure only: with rayleigh=True a nonzero
    relative_count raises ValueError.  Logging, tensorboard and checkpoints stay with the caller
    (`callback(step, energy)`)."""
    if rayleigh and relative_count > 0:
        raise ValueError("relative_count extends the closure's target set (rayleigh=False); the Rayleigh-quotient step "
                         "(rayleigh=True) works on the sampled set alone")
    generate = getattr(network, "generate_unique")
    with torch.no_grad():  # psi_i feeds find_relative only; the losses evaluate the network again
        configs_i, psi_i, _, _ = generate(sampling_count, seed=seed) if seed is not None else generate(sampling_count)
    energies: list[float] = []
    if rayleigh:
        energy_set = EnergySet(model, configs_i)
        for i in range(local_step):
            optimizer.zero_grad()
            e = energy_backward(network, energy_set, chunk)
            optimizer.step()
            energies.append(e)
            if callback is not None:
                callback(i, e)
        return CycleResult(configs_i, energies)
    configs_src = configs_i
    configs_dst = configs_i
    if relative_count > len(configs_i):
        extra = model.find_relative(configs_i, psi_i, relative_count - len(configs_i))  # type: ignore[attr-defined]
        configs_dst = torch.cat([configs_i, extra])

    def closure() -> torch.Tensor:
        optimizer.zero_grad()
        energy = closure_energy(model, network, configs_src, configs_dst)
        energy.backward()  # type: ignore[no-untyped-call]
        return energy

    for i in range(local_step):
        e = float(optimizer.step(closure).detach())  # type: ignore[union-attr,arg-type]
        energies.append(e)
        if callback is not None:
            callback(i, e)
    return CycleResult(configs_i, energies)
