"""
Drop-in replacement of qmb.hamiltonian.Hamiltonian (qmb/hamiltonian.py:12-206).

Same constructor (term dict or ready (site, kind, coef) tuple, ``kind="fermi" | "bose2"``), same
attributes (site, kind, coef, particle_cut), same methods and return conventions; the native work
runs in libqmb_hip.so through the ops registered by qmb_hip.ops under the reference's op names.
Differences, all compatible with the reference's callers:
  - modules are built ahead of time (no per-shape JIT); ``_load_module`` returns the op namespace;
  - the order of terms with equal |coef| is fixed (stable sort) instead of implementation-defined;
  - apply_within is differentiable with respect to psi_i (the reference op is not);
  - CPU tensors are rejected by the dispatcher: this build computes on the GPU only.
"""

from __future__ import annotations

import typing

import torch

from . import ops
from .prepared import PreparedTerms
from .synthetic import terms_from_dict


class _Declaration:
    """Stands in for the reference's pybind 'declaration' module, which exposes prepare()."""

    @staticmethod
    def prepare(hamiltonian: dict[tuple[tuple[int, int], ...], complex]) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        return terms_from_dict(hamiltonian)


class Hamiltonian:
    """A second-quantized operator sum_t coef_t O_t held as a term table (site int16[T, 4],
    kind uint8[T, 4], coef float64[T, 2]), with the four operations on bit-packed configurations:
    apply_within, find_relative, diagonal_term and single_relative (the class of
    qmb/hamiltonian.py:12-206, same constructor, attributes and methods)."""

    _hamiltonian_module: dict[tuple[str, int, int], object] = {}

    @classmethod
    def _load_module(cls, device_type: str = "declaration", n_qubytes: int = 0, particle_cut: int = 0) -> object:
        """qmb/hamiltonian.py:37-69: the 'declaration' module, or the op namespace for (nq, pc)."""
        key = (device_type, n_qubytes, particle_cut)
        if key not in cls._hamiltonian_module:
            match device_type:
                case "declaration":
                    cls._hamiltonian_module[key] = _Declaration()
                case "cpu" | "cuda":
                    cls._hamiltonian_module[key] = ops.load(n_qubytes, particle_cut)
                case _:
                    raise ValueError("Unsupported device type")
        return cls._hamiltonian_module[key]

    @classmethod
    def _prepare(cls, hamiltonian: dict[tuple[tuple[int, int], ...], complex]) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        return getattr(cls._load_module(), "prepare")(hamiltonian)

    def __init__(self, hamiltonian: dict[tuple[tuple[int, int], ...], complex] | tuple[torch.Tensor, torch.Tensor, torch.Tensor], *,
                 kind: str) -> None:
        self.site: torch.Tensor
        self.kind: torch.Tensor
        self.coef: torch.Tensor
        if isinstance(hamiltonian, dict):
            self.site, self.kind, self.coef = self._prepare(hamiltonian)
            self._sort_site_kind_coef()
        else:
            self.site, self.kind, self.coef = hamiltonian
        self.particle_cut: int
        match kind:
            case "fermi":
                self.particle_cut = 1
            case "bose2":
                self.particle_cut = 2
            case _:
                raise ValueError(f"Unknown kind: {kind}")

    def _sort_site_kind_coef(self) -> None:
        """qmb/hamiltonian.py:93-97: terms by |coef| descending (stable here)."""
        order = self.coef.norm(dim=1).argsort(descending=True, stable=True)
        self.site = self.site[order]
        self.kind = self.kind[order]
        self.coef = self.coef[order]

    def _prepare_data(self, device: torch.device) -> None:
        self.site = self.site.to(device=device).contiguous()
        self.kind = self.kind.to(device=device).contiguous()
        self.coef = self.coef.to(device=device).contiguous()

    def prepared_terms(self, device: torch.device, n_qubytes: int) -> PreparedTerms:
        """The compiled term table (buckets, structured pair/single/diagonal tables) for configs of
        n_qubytes bytes on `device`, built on first use and kept while site/kind/coef stay the same
        tensors with the same contents (their in-place versions are part of the key).  Repeated
        applications on one config set (Krylov loops) start from it."""
        self._prepare_data(device)
        # versions too: an in-place change of the table (ham.coef.mul_(2)) must rebuild it
        key = (str(device), n_qubytes, self.site.data_ptr(), self.kind.data_ptr(), self.coef.data_ptr(), self.site.size(0),
               self.site._version, self.kind._version, self.coef._version)  # pylint: disable=protected-access
        cached = getattr(self, "_prepared", None)
        if cached is None or cached[0] != key:
            cached = (key, PreparedTerms(self.site, self.kind, self.coef, n_qubytes, self.particle_cut))
            self._prepared = cached  # pylint: disable=attribute-defined-outside-init
        return cached[1]

    def apply_within(self, configs_i: torch.Tensor, psi_i: torch.Tensor, configs_j: torch.Tensor) -> torch.Tensor:
        """
        Applies the Hamiltonian to the given vector (qmb/hamiltonian.py:104-131).

        configs_i: uint8 [batch_size_i, n_qubytes]; psi_i: complex128 [batch_size_i];
        configs_j: uint8 [batch_size_j, n_qubytes].  Returns complex128 [batch_size_j] in configs_j's
        row order.
        """
        self._prepare_data(configs_i.device)
        _apply_within: typing.Callable[..., torch.Tensor]
        _apply_within = getattr(self._load_module(configs_i.device.type, configs_i.size(1), self.particle_cut), "apply_within")
        psi_j = torch.view_as_complex(_apply_within(configs_i, torch.view_as_real(psi_i), configs_j, self.site, self.kind, self.coef))
        return psi_j

    def find_relative(self, configs_i: torch.Tensor, psi_i: torch.Tensor, count_selected: int,
                      configs_exclude: torch.Tensor | None = None) -> torch.Tensor:
        """
        Configurations reached from configs_i, outside configs_exclude (default configs_i), from the
        count_selected largest contributions |coef * psi|^2; distinct and sorted by bytes
        (qmb/hamiltonian.py:133-166).
        """
        if configs_exclude is None:
            configs_exclude = configs_i
        self._prepare_data(configs_i.device)
        _find_relative: typing.Callable[..., torch.Tensor]
        _find_relative = getattr(self._load_module(configs_i.device.type, configs_i.size(1), self.particle_cut), "find_relative")
        configs_j = _find_relative(configs_i, torch.view_as_real(psi_i), count_selected, self.site, self.kind, self.coef, configs_exclude)
        return configs_j

    def diagonal_term(self, configs: torch.Tensor) -> torch.Tensor:
        """Diagonal matrix elements for the given configurations (qmb/hamiltonian.py:168-186)."""
        self._prepare_data(configs.device)
        _diagonal_term: typing.Callable[..., torch.Tensor]
        _diagonal_term = getattr(self._load_module(configs.device.type, configs.size(1), self.particle_cut), "diagonal_term")
        psi_result = torch.view_as_complex(_diagonal_term(configs, self.site, self.kind, self.coef))
        return psi_result

    def single_relative(self, configs: torch.Tensor) -> torch.Tensor:
        """One weighted-random neighbour per configuration outside the set (qmb/hamiltonian.py:188-206)."""
        self._prepare_data(configs.device)
        _single_relative: typing.Callable[..., torch.Tensor]
        _single_relative = getattr(self._load_module(configs.device.type, configs.size(1), self.particle_cut), "single_relative")
        configs_result = _single_relative(configs, self.site, self.kind, self.coef)
        return configs_result
