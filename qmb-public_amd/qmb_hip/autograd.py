"""
Differentiable Hamiltonian application.

apply_within is linear in psi_i, psi_j = A psi_i with A[k, i] = sum_t sign * coef_t over terms t
mapping configs_i[i] to configs_j[k].  For a real loss L, PyTorch's gradient convention for complex
tensors gives dL/dpsi_i = A^H dL/dpsi_j, which the adjoint kernel computes as a gather over the same
connection enumeration (no stored connections, no atomics on other configs' rows).  The reference
never differentiates this path (qmb/vmc.py:104-116 calls apply_within under no_grad and relies on
H being Hermitian); with this Function the energy <psi|H|psi> / <psi|psi> trains a network without
materialising H.
"""

from __future__ import annotations

import torch

from . import prepared as pp


class ApplyWithin(torch.autograd.Function):
    """psi_j = H|_{configs_j <- configs_i} psi_i with a hand-written adjoint; prepared objects are reused."""

    @staticmethod
    def forward(ctx: torch.autograd.function.FunctionCtx, psi_i: torch.Tensor, configs_i: torch.Tensor,  # type: ignore[override]
                terms: pp.PreparedTerms, targets: pp.PreparedTargets, tile: int = 0, groups: int = 0,
                symmetric: bool | None = None) -> torch.Tensor:
        if psi_i.dtype != torch.complex128:
            raise RuntimeError("psi must be complex128.")
        if symmetric is None:
            symmetric = pp.can_use_symmetric(terms, targets, configs_i)
        out = pp.apply(terms, targets, configs_i, torch.view_as_real(psi_i.resolve_conj().contiguous()), pp.MODE_FORWARD,
                       tile=tile, groups=groups, symmetric=symmetric)
        ctx.terms = terms
        ctx.targets = targets
        ctx.configs_i = configs_i
        ctx.launch = (tile, groups, symmetric)
        return torch.view_as_complex(out)

    @staticmethod
    def backward(ctx: torch.autograd.function.FunctionCtx, grad_j: torch.Tensor) -> tuple[object, ...]:  # type: ignore[override]
        g = torch.view_as_real(grad_j.resolve_conj().contiguous())
        tile, groups, symmetric = ctx.launch
        grad_i = pp.apply(ctx.terms, ctx.targets, ctx.configs_i, g, pp.MODE_ADJOINT, tile=tile, groups=groups, symmetric=symmetric)
        return torch.view_as_complex(grad_i), None, None, None, None, None, None


def apply_within(psi_i: torch.Tensor, configs_i: torch.Tensor, terms: pp.PreparedTerms, targets: pp.PreparedTargets,
                 tile: int = 0, groups: int = 0, symmetric: bool | None = None) -> torch.Tensor:
    """Differentiable psi_j (complex128 [n_j]) on the configs indexed by `targets`.

    symmetric=None takes the half-lookup path whenever the operator is Hermitian and `targets` indexes
    configs_i itself (distinct rows); False forces the general scatter / gather kernels.
    """
    return ApplyWithin.apply(psi_i, configs_i, terms, targets, tile, groups, symmetric)


def energy(psi: torch.Tensor, configs: torch.Tensor, terms: pp.PreparedTerms, targets: pp.PreparedTargets) -> torch.Tensor:
    """Re <psi|H|psi> / <psi|psi> on one config set (`targets` indexes `configs`); differentiable in psi."""
    h_psi = apply_within(psi, configs, terms, targets)
    num = torch.vdot(psi, h_psi)
    den = torch.vdot(psi, psi)
    return (num / den).real
