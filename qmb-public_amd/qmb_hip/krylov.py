"""
Lanczos drivers on the device: the Krylov loops of qmb/haar.py:21-183 (_DynamicLanczos) and
qmb/rldiag.py:19-52 (lanczos_energy), with the same fields, modes, yields and stopping rule.

What changes is where the work runs and what is rebuilt per step:
  - H restricted to the basis is a WithinOperator: the term table is compiled once per Hamiltonian
    (Hamiltonian.prepared_terms) and the basis indexed once per basis (PreparedTargets, plus the
    core indexes of the structured path); every step is then one application through the pair,
    single and diagonal kernels (or the enumeration kernel for tables without pair excitations),
    never a per-step sort or table rebuild;
  - the vector arithmetic between two applications (alpha, the three-term update, the norm and the
    normalisation) is one fused library call, qmb_lanczos_step: three HBM passes and one 16-byte
    read-back per step;
  - the Krylov vectors stay in device memory (HBM holds tens of 16 MB-per-million-configs
    vectors); keep_on_device=False moves all but the last two to host memory, as the reference
    does (qmb/haar.py:142);
  - batch_count_apply_within (the reference's memory cap: the input rows applied in that many
    slices, qmb/haar.py:144-166) is accepted and validated; the prepared application holds no
    scratch proportional to the input rows, so the basis is applied in one pass and the sum is the
    same up to rounding order.
The tridiagonal problem is solved on the host with SciPy's stebz, lowest eigenpair only, as in the
reference (qmb/haar.py:168-183).
"""

from __future__ import annotations

import ctypes
import dataclasses
import logging
import typing

import numpy
import scipy.linalg
import torch

from . import _lib
from . import prepared as pp
from .hamiltonian import Hamiltonian
from .ops import _ptr, _stream, _workspace


def _hamiltonian_of(model: object) -> Hamiltonian | None:
    """The Hamiltonian behind a model: the object itself, or its `hamiltonian` attribute (the
    reference's models delegate apply_within/find_relative to one)."""
    if isinstance(model, Hamiltonian):
        return model
    inner = getattr(model, "hamiltonian", None)
    return inner if isinstance(inner, Hamiltonian) else None


class WithinOperator:
    """v -> H v on one config set (configs_i = configs_j = the basis), through the prepared tables.

    A model that does not expose a qmb_hip Hamiltonian is applied through its own apply_within."""

    def __init__(self, model: object, configs: torch.Tensor) -> None:
        if configs.dtype != torch.uint8 or configs.dim() != 2:
            raise RuntimeError("configs must be a uint8 [n, n_qubytes] tensor.")
        self.model = model
        self.configs = configs.contiguous()
        self.n = self.configs.size(0)
        ham = _hamiltonian_of(model)
        self.terms: pp.PreparedTerms | None = None
        self.targets: pp.PreparedTargets | None = None
        if ham is not None and self.configs.is_cuda:
            self.terms = ham.prepared_terms(self.configs.device, self.configs.size(1))
            self.targets = pp.PreparedTargets(self.configs, self.configs.size(1))

    def __call__(self, v: torch.Tensor) -> torch.Tensor:
        if self.terms is None or self.targets is None:
            return self.model.apply_within(self.configs, v, self.configs)  # type: ignore[attr-defined]
        vec = torch.view_as_real(v.resolve_conj().contiguous())
        out = pp.apply(self.terms, self.targets, self.configs, vec, pp.MODE_FORWARD)
        return torch.view_as_complex(out)


class LanczosStep:
    """The fused vector step of one device: alpha = Re(w^H v); w -= alpha v + beta v_prev;
    |w|; v_next = w / |w| (qmb_lanczos_step)."""

    def __init__(self, device: torch.device) -> None:
        lib = _lib.library()
        size = ctypes.c_size_t(0)
        _lib.check(lib.qmb_lanczos_workspace_size(ctypes.byref(size)), "qmb_lanczos_workspace_size")
        self.device = device
        self.workspace = _workspace(size.value, device)
        self.scalars = torch.zeros(2, dtype=torch.float64, device=device)

    def __call__(self, v: torch.Tensor, v_prev: torch.Tensor | None, beta: float, w: torch.Tensor) -> tuple[float, float, torch.Tensor]:
        """Updates w in place; returns (alpha, |w|, v_next).  Synchronises once for the scalars."""
        n = v.numel()
        for t in (v, w) + ((v_prev,) if v_prev is not None else ()):
            if t.dtype != torch.complex128 or t.shape != (n,) or not t.is_contiguous() or t.device != self.device:
                raise RuntimeError("Lanczos vectors must be contiguous complex128 [n] tensors on the step's device")
        v_next = torch.empty_like(v)
        with torch.cuda.device(self.device):
            _lib.check(_lib.library().qmb_lanczos_step(n, _ptr(v), _ptr(v_prev), float(beta), _ptr(w), _ptr(v_next), _ptr(self.scalars),
                                                       _ptr(self.workspace), self.workspace.numel(), _stream(self.device)),
                       "qmb_lanczos_step")
        alpha, norm = self.scalars.tolist()
        return alpha, norm, v_next


def lowest_ritz_pair(alpha: typing.Sequence[torch.Tensor], beta: typing.Sequence[torch.Tensor],
                     vectors: typing.Sequence[torch.Tensor], device: torch.device) -> tuple[torch.Tensor, torch.Tensor]:
    """Lowest eigenpair of the Lanczos tridiagonal T (diagonal alpha, off-diagonal beta), lifted to
    the basis: (theta, sum_k y_k vectors[k]) on `device`, theta the smallest eigenvalue of T and y
    its unit eigenvector.  With a single alpha T is 1 x 1, so the pair is (alpha_0, vectors[0]).
    The eigenproblem goes to LAPACK's bisection driver (stebz), which computes only the one
    eigenvalue asked for, in O(len(alpha)) memory."""
    if len(beta) == 0:
        return alpha[0], vectors[0]
    diagonal = numpy.fromiter((float(a) for a in alpha), dtype=numpy.float64, count=len(alpha))
    off_diagonal = numpy.fromiter((float(b) for b in beta), dtype=numpy.float64, count=len(beta))
    theta, y = scipy.linalg.eigh_tridiagonal(diagonal, off_diagonal, select="i", select_range=(0, 0), lapack_driver="stebz")
    ritz = torch.zeros(vectors[0].shape, dtype=vectors[0].dtype, device=device)
    for weight, vector in zip(y[:, 0].tolist(), vectors):
        ritz.add_(vector.to(device=device), alpha=weight)
    return torch.tensor(float(theta[0]), dtype=torch.float64), ritz


def lanczos(operator: typing.Callable[[torch.Tensor], torch.Tensor], psi: torch.Tensor, threshold: float,
            keep_on_device: bool = True) -> typing.Iterator[tuple[list[torch.Tensor], list[torch.Tensor], list[torch.Tensor]]]:
    """The Lanczos recurrence of qmb/haar.py:107-142 from the start vector psi: yields (alpha, beta, v)
    after every application; stops when the residual norm falls below threshold.  alpha and beta
    hold float64 scalar tensors on the host, v the Krylov vectors."""
    if psi.dtype != torch.complex128 or psi.dim() != 1:
        raise RuntimeError("psi must be a complex128 [n] tensor.")
    step = LanczosStep(psi.device)
    v: list[torch.Tensor] = [(psi / torch.linalg.norm(psi)).contiguous()]  # pylint: disable=not-callable
    alpha: list[torch.Tensor] = []
    beta: list[torch.Tensor] = []
    w = operator(v[-1]).contiguous()
    a, norm_w, v_next = step(v[-1], None, 0.0, w)
    alpha.append(torch.tensor(a, dtype=torch.float64))
    yield (alpha, beta, v)
    while True:
        if norm_w < threshold:
            break
        beta.append(torch.tensor(norm_w, dtype=torch.float64))
        v.append(v_next)
        w = operator(v[-1]).contiguous()
        a, norm_w, v_next = step(v[-1], v[-2], norm_w, w)
        alpha.append(torch.tensor(a, dtype=torch.float64))
        yield (alpha, beta, v)
        if not keep_on_device:
            v[-2] = v[-2].cpu()  # the recurrence only reads the last two vectors


@dataclasses.dataclass
class DynamicLanczos:
    """Lanczos on a basis that can grow between runs (the Lanczos loop of qmb/haar.py:21-183, same
    fields and yields).  Every run of the recurrence starts from the stored start vector `psi`,
    zero-padded onto the current basis, and `run()` yields (energy, basis, Ritz vector) tuples:

      count_extend == 0   one run on the fixed basis; a tuple after each of its first step + 1
                          applications;
      first_extend        `step` growth rounds first -- each seeds find_relative with the
                          count_extend largest-|amplitude| rows of a running amplitude vector
                          (first the start vector, then H applied from the previous basis onto
                          the grown one) -- then as count_extend == 0;
      single_extend       as count_extend == 0, one growth seeded by the whole basis with
                          amplitudes sqrt(sum_k |v_k|^2) over the run's Krylov vectors, then as
                          count_extend == 0 again;
      otherwise           for s = 1 .. step + 1: a run of s applications, one tuple from its last
                          state, then (except after the last) a growth seeded by the whole basis
                          with the run's last Krylov vector.

    A growth appends find_relative(seeds, amplitudes, count_extend, exclude = basis) to the basis."""

    # pylint: disable=too-many-instance-attributes

    model: typing.Any
    configs: torch.Tensor
    psi: torch.Tensor
    step: int
    threshold: float
    count_extend: int
    batch_count_apply_within: int
    single_extend: bool
    first_extend: bool
    keep_on_device: bool = True
    _operator: WithinOperator | None = dataclasses.field(default=None, init=False, repr=False)

    def __post_init__(self) -> None:
        if self.batch_count_apply_within < 1:
            raise ValueError("batch_count_apply_within must be at least 1")
        if self.psi.dtype != torch.complex128 or self.psi.shape != (self.configs.size(0),):
            raise RuntimeError("psi must be complex128 with one entry per config")

    def _within(self) -> WithinOperator:
        """H on the current basis; a new basis gets a new operator (and new indexes)."""
        if self._operator is None or self._operator.configs is not self.configs:
            self.configs = self.configs.contiguous()
            self._operator = WithinOperator(self.model, self.configs)
        return self._operator

    def _grow(self, amplitudes: torch.Tensor, seeds: torch.Tensor | None = None) -> None:
        """Appends find_relative(seeds (default: the basis), amplitudes, count_extend, basis) to the
        basis and zeros for the new rows to the start vector.  find_relative ranks |coef * amplitude|^2,
        so a real amplitude vector is taken as complex."""
        origin = self.configs if seeds is None else seeds
        if not amplitudes.is_complex():
            amplitudes = amplitudes.to(torch.complex128)
        before = self.configs.size(0)
        found = self.model.find_relative(origin, amplitudes, self.count_extend, self.configs)
        self.configs = torch.cat((self.configs, found))
        self.psi = torch.cat((self.psi, self.psi.new_zeros(self.configs.size(0) - before)))
        logging.info("Lanczos basis grown from %d to %d configurations", before, self.configs.size(0))

    def _states(self, applications: int) -> typing.Iterator[tuple[list[torch.Tensor], list[torch.Tensor], list[torch.Tensor]]]:
        """The recurrence's (alpha, beta, v) after each of its first `applications` applications
        (fewer if it stops on the threshold), from the start vector on the current basis."""
        if applications < 1:
            return
        for done, state in enumerate(lanczos(self._within(), self.psi, self.threshold, self.keep_on_device), start=1):
            yield state
            if done == applications:
                return

    def _ritz(self, state: tuple[list[torch.Tensor], list[torch.Tensor], list[torch.Tensor]]) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        alpha, beta, v = state
        energy, vector = lowest_ritz_pair(alpha, beta, v, self.configs.device)
        return energy, self.configs, vector

    def _grow_ahead(self) -> None:
        """The first_extend rounds: grow from the largest amplitudes, then carry the amplitudes onto
        the grown basis through H."""
        running = self.psi
        for _ in range(self.step):
            weight = torch.view_as_real(running).square().sum(dim=-1)
            top = torch.argsort(weight, descending=True)[:self.count_extend]
            previous = self.configs
            self._grow(running[top], previous[top])
            running = self.model.apply_within(previous, running, self.configs)

    def run(self) -> typing.Iterator[tuple[torch.Tensor, torch.Tensor, torch.Tensor]]:
        """Yields (energy, basis, Ritz vector) as described in the class docstring."""
        full = self.step + 1
        if self.count_extend == 0 or self.first_extend or self.single_extend:
            if self.count_extend != 0 and self.first_extend:
                self._grow_ahead()
            vectors: list[torch.Tensor] = []
            for state in self._states(full):
                vectors = state[2]
                yield self._ritz(state)
            if self.count_extend == 0 or self.first_extend:
                return
            # single_extend: one growth from the root-sum-square of every Krylov vector of the run
            spread = torch.zeros(self.configs.size(0), dtype=torch.float64, device=self.configs.device)
            for vector in vectors:
                spread.add_(vector.to(device=spread.device).abs().square())
            self._grow(spread.sqrt_())
            for state in self._states(full):
                yield self._ritz(state)
            return
        for applications in range(1, full + 1):
            last = None
            for last in self._states(applications):
                pass
            assert last is not None
            yield self._ritz(last)
            if applications < full:
                self._grow(last[2][-1])


def lanczos_energy(model: object, configs: torch.Tensor, step: int, threshold: float,
                   generator: torch.Generator | None = None) -> tuple[float, torch.Tensor]:
    """Lowest Ritz value and vector after at most step + 1 applications from a random start
    (qmb/rldiag.py:19-52)."""
    vector = torch.randn([configs.size(0)], dtype=torch.complex128, device=configs.device, generator=generator)
    alpha: list[torch.Tensor] = []
    beta: list[torch.Tensor] = []
    v: list[torch.Tensor] = []
    for i, (alpha, beta, v) in enumerate(lanczos(WithinOperator(model, configs), vector, threshold)):
        if i == step + 1:
            break
    if len(beta) == 0:
        return float(alpha[0]), v[0]
    energy, result = lowest_ritz_pair(alpha, beta, v, configs.device)
    return float(energy), result
