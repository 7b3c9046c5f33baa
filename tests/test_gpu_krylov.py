"""
GPU checks of the Lanczos drivers (qmb_hip.krylov, qmb/haar.py:21-183, qmb/rldiag.py:19-52):

  - the fused vector step (qmb_lanczos_step) against float64 numpy arithmetic of the same
    formulas, with and without the previous vector, on empty, ragged and large vectors; and
    bit-for-bit equal on a repeated call (fixed-order reductions);
  - the recurrence through the prepared operator: its alpha/beta agree with the recurrence on
    the dense matrix assembled column by column from the oracle's apply_within, over the first
    steps (before orthogonality loss makes the two sequences drift apart);
  - the lowest Ritz value equals numpy.linalg.eigh's lowest eigenvalue of that dense matrix once the
    Krylov space spans the sector, and the Ritz vector's residual is small (molecular sector on the
    structured path, Hubbard sector on the enumeration path, a model without a qmb_hip Hamiltonian
    through its own apply_within);
  - the basis-growing modes: dynamic extension from one config reaches the full sector (the set
    of configs connected to it by the Hamiltonian, from the dense matrix) and every energy is
    variational (>= the sector's lowest eigenvalue, Cauchy interlacing); first_extend and
    single_extend grow the basis where the reference does, to the connected set the dense matrix
    gives; rldiag's lanczos_energy gives the same lowest eigenvalue.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import Hamiltonian
from qmb_hip import krylov as kr
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu


def _dev(x: np.ndarray, device: torch.device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _dense(pc: int, cfg: np.ndarray, s: np.ndarray, k: np.ndarray, c: np.ndarray) -> np.ndarray:
    """H restricted to cfg, column j = oracle apply_within of the j-th unit vector."""
    n = len(cfg)
    h = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[j] = 1.0
        h[:, j] = orc.apply_within(pc, cfg, e, cfg, s, k, c)[0]
    return h


def _molecular_case() -> tuple[tuple[np.ndarray, np.ndarray, np.ndarray], np.ndarray, int, str]:
    return sorted_terms(*sy.molecular_terms(4, 3)), sy.sector_configs(4, 2, 2, 36, 0), 1, "fermi"


def _hubbard_case() -> tuple[tuple[np.ndarray, np.ndarray, np.ndarray], np.ndarray, int, str]:
    s, k, c = sy.terms_from_dict(sy.hubbard_dict(2, 2))
    return sorted_terms(s.numpy(), k.numpy(), c.numpy()), sy.sector_configs(4, 2, 2, 36, 0), 1, "fermi"


CASES = [("molecular-8so-sector", _molecular_case), ("hubbard-2x2-sector", _hubbard_case)]


def _hamiltonian(terms: tuple[np.ndarray, np.ndarray, np.ndarray], kind: str, device: torch.device) -> Hamiltonian:
    s, k, c = terms
    return Hamiltonian((_dev(s, device), _dev(k, device), _dev(c, device)), kind=kind)


def _as_set(cfg: np.ndarray) -> set[bytes]:
    return {bytes(r) for r in np.asarray(cfg)}


@pytest.mark.parametrize("n,with_prev", [(0, False), (1, True), (1000, False), (4097, True), (300_001, True)])
def test_lanczos_step_matches_numpy(cuda_device: torch.device, n: int, with_prev: bool) -> None:
    rng = np.random.default_rng(n + 7)
    v, w0, vp = (rng.normal(size=n) + 1j * rng.normal(size=n) for _ in range(3))
    beta = 0.75
    step = kr.LanczosStep(cuda_device)
    w = _dev(w0, cuda_device)
    alpha, norm, v_next = step(_dev(v, cuda_device), _dev(vp, cuda_device) if with_prev else None, beta, w)
    want_alpha = float(np.sum(w0.real * v.real + w0.imag * v.imag))
    want_w = w0 - want_alpha * v - (beta * vp if with_prev else 0.0)
    want_norm = float(np.sqrt(np.sum(np.abs(want_w) ** 2)))
    scale = float(np.sum(np.abs(w0) * np.abs(v))) + 1e-300
    assert abs(alpha - want_alpha) <= 1e-13 * scale, (alpha, want_alpha)
    assert abs(norm - want_norm) <= 1e-13 * (want_norm + 1e-300)
    mag = np.abs(w0) + np.abs(v) * abs(want_alpha) + (beta * np.abs(vp) if with_prev else 0.0)
    assert np.all(np.abs(w.cpu().numpy() - want_w) <= 1e-13 * mag + 1e-300)
    if n:
        np.testing.assert_allclose(v_next.cpu().numpy(), want_w / want_norm, rtol=0, atol=1e-13 * np.abs(want_w / want_norm).max())
    # a repeated step on the same inputs is bitwise the same (fixed reduction order)
    w2 = _dev(w0, cuda_device)
    again = step(_dev(v, cuda_device), _dev(vp, cuda_device) if with_prev else None, beta, w2)
    assert (again[0], again[1]) == (alpha, norm)
    assert torch.equal(w2, w) and torch.equal(again[2], v_next)


def test_lanczos_step_rejects_bad_vectors(cuda_device: torch.device) -> None:
    step = kr.LanczosStep(cuda_device)
    v = torch.zeros(8, dtype=torch.complex128, device=cuda_device)
    with pytest.raises(RuntimeError):
        step(v, None, 0.0, torch.zeros(9, dtype=torch.complex128, device=cuda_device))
    with pytest.raises(RuntimeError):
        step(v, None, 0.0, torch.zeros(8, dtype=torch.complex64, device=cuda_device))
    with pytest.raises(RuntimeError):
        step(v, None, 0.0, v)  # w aliases v


def _numpy_lanczos(h: np.ndarray, psi: np.ndarray, n_steps: int) -> tuple[list[float], list[float]]:
    """The dense recurrence of qmb/haar.py:123-142 without stopping rule, for n_steps applications."""
    v = [psi / np.linalg.norm(psi)]
    w = h @ v[-1]
    alpha = [float(np.vdot(w, v[-1]).real)]
    beta: list[float] = []
    w = w - alpha[-1] * v[-1]
    for _ in range(n_steps - 1):
        beta.append(float(np.linalg.norm(w)))
        v.append(w / beta[-1])
        w = h @ v[-1]
        alpha.append(float(np.vdot(w, v[-1]).real))
        w = w - alpha[-1] * v[-1] - beta[-1] * v[-2]
    return alpha, beta


@pytest.mark.parametrize("name,case_fn", CASES, ids=[c[0] for c in CASES])
def test_lanczos_matches_dense(cuda_device: torch.device, name: str, case_fn: object) -> None:
    terms, cfg, pc, kind = case_fn()
    h = _dense(pc, cfg, *terms)
    assert np.allclose(h, h.conj().T, atol=1e-12), name
    e0 = float(np.linalg.eigvalsh(h)[0])
    ham = _hamiltonian(terms, kind, cuda_device)
    cfg_t = _dev(cfg, cuda_device)
    psi = sy.random_psi(len(cfg), 31)
    op = kr.WithinOperator(ham, cfg_t)
    assert op.terms is not None
    if name.startswith("molecular"):
        assert op.terms.pairs_desc is not None  # the structured path
    else:
        assert op.terms.pairs_desc is None  # the enumeration path
    want_a, want_b = _numpy_lanczos(h, psi, 8)
    hnorm = np.abs(h).sum(axis=1).max()
    for i, (alpha, beta, v) in enumerate(kr.lanczos(op, _dev(psi, cuda_device), 1e-12)):
        assert len(alpha) == i + 1 and len(beta) == i and len(v) == i + 1
        if i == 7:
            break
    np.testing.assert_allclose([float(a) for a in alpha], want_a, rtol=0, atol=1e-10 * hnorm)
    np.testing.assert_allclose([float(b) for b in beta], want_b, rtol=0, atol=1e-10 * hnorm)
    # the full driver in its plain mode: lowest Ritz value -> lowest eigenvalue
    driver = kr.DynamicLanczos(ham, cfg_t, _dev(psi, cuda_device), step=80, threshold=1e-12, count_extend=0, batch_count_apply_within=1,
                               single_extend=False, first_extend=False)
    results = list(driver.run())
    assert 1 < len(results) <= 81
    energy, configs, vec = results[-1]
    assert configs is cfg_t or torch.equal(configs, cfg_t)
    assert abs(float(energy) - e0) <= 1e-9 * hnorm, (name, float(energy), e0)
    x = vec.cpu().numpy()
    x = x / np.linalg.norm(x)
    assert np.linalg.norm(h @ x - float(energy) * x) <= 1e-5 * hnorm
    for e, _, _ in results:
        assert float(e) >= e0 - 1e-10 * hnorm  # Ritz values of H are bounded below by its lowest eigenvalue
    # vectors offloaded to the host give the same recurrence
    offload = kr.DynamicLanczos(ham, cfg_t, _dev(psi, cuda_device), step=80, threshold=1e-12, count_extend=0, batch_count_apply_within=4,
                                single_extend=False, first_extend=False, keep_on_device=False)
    assert abs(float(list(offload.run())[-1][0]) - float(energy)) <= 1e-12 * hnorm


class _ForeignModel:
    """A model exposing only the reference's ModelProto methods, backed by a dense matrix on the GPU."""

    def __init__(self, h: np.ndarray, device: torch.device) -> None:
        self.h = _dev(h, device)

    def apply_within(self, configs_i: torch.Tensor, psi_i: torch.Tensor, configs_j: torch.Tensor) -> torch.Tensor:
        assert configs_i.size(0) == configs_j.size(0) == self.h.size(0)
        return self.h @ psi_i


def test_lanczos_through_model_apply_within(cuda_device: torch.device) -> None:
    terms, cfg, pc, _ = _molecular_case()
    h = _dense(pc, cfg, *terms)
    model = _ForeignModel(h, cuda_device)
    cfg_t = _dev(cfg, cuda_device)
    op = kr.WithinOperator(model, cfg_t)
    assert op.terms is None
    gen = torch.Generator(device=cuda_device)
    gen.manual_seed(5)
    energy, vec = kr.lanczos_energy(model, cfg_t, 60, 1e-12, generator=gen)
    assert abs(energy - float(np.linalg.eigvalsh(h)[0])) <= 1e-9 * np.abs(h).sum(axis=1).max()
    assert vec.shape == (len(cfg),)


def _connected(h: np.ndarray, rows: np.ndarray) -> np.ndarray:
    """Indices reached from `rows` by one application of H (nonzero matrix elements)."""
    return np.nonzero(np.abs(h[:, rows]).sum(axis=1) > 0)[0]


def test_dynamic_extension_reaches_sector(cuda_device: torch.device) -> None:
    terms, cfg, pc, kind = _molecular_case()
    h = _dense(pc, cfg, *terms)
    e0 = float(np.linalg.eigvalsh(h)[0])
    hnorm = np.abs(h).sum(axis=1).max()
    ham = _hamiltonian(terms, kind, cuda_device)
    start = int(np.argmin(np.diag(h).real))  # the lowest diagonal element: the Hartree-Fock-like config
    driver = kr.DynamicLanczos(ham, _dev(cfg[start:start + 1], cuda_device), torch.ones(1, dtype=torch.complex128, device=cuda_device),
                               step=40, threshold=1e-12, count_extend=100_000, batch_count_apply_within=1, single_extend=False,
                               first_extend=False)
    seen = {start}
    index = {bytes(r): i for i, r in enumerate(cfg)}
    previous: np.ndarray | None = None
    energies = []
    for energy, configs, psi in driver.run():
        c = configs.cpu().numpy()
        assert len(_as_set(c)) == len(c)  # distinct rows
        if previous is not None:
            np.testing.assert_array_equal(c[:len(previous)], previous)  # the basis only grows, in place
        rows = np.array([index[bytes(r)] for r in c])
        assert set(rows.tolist()) >= seen
        seen = set(rows.tolist())
        # the extension can only add configs connected to the current basis
        if previous is not None:
            prev_rows = np.array([index[bytes(r)] for r in previous])
            assert set(rows.tolist()) <= set(prev_rows.tolist()) | set(_connected(h, prev_rows).tolist())
        assert psi.shape == (len(c),)
        energies.append(float(energy))
        assert float(energy) >= e0 - 1e-10 * hnorm
        previous = c
    assert seen == set(range(len(cfg)))  # the whole sector is connected to the start
    assert abs(energies[-1] - e0) <= 1e-9 * hnorm


def test_first_and_single_extend(cuda_device: torch.device) -> None:
    terms, cfg, pc, kind = _molecular_case()
    h = _dense(pc, cfg, *terms)
    e0 = float(np.linalg.eigvalsh(h)[0])
    hnorm = np.abs(h).sum(axis=1).max()
    ham = _hamiltonian(terms, kind, cuda_device)
    index = {bytes(r): i for i, r in enumerate(cfg)}
    start = int(np.argmin(np.diag(h).real))
    one = _dev(cfg[start:start + 1], cuda_device)
    psi1 = torch.ones(1, dtype=torch.complex128, device=cuda_device)
    # first_extend: `step` extensions before iterating, so the basis is the set within `step` hops
    for steps in (1, 2):
        driver = kr.DynamicLanczos(ham, one, psi1, step=steps, threshold=1e-12, count_extend=100_000, batch_count_apply_within=1,
                                   single_extend=False, first_extend=True)
        out = list(driver.run())
        assert len(out) <= steps + 1
        reach = {start}
        for _ in range(steps):
            reach |= set(_connected(h, np.array(sorted(reach))).tolist())
        for energy, configs, _ in out:
            assert {index[bytes(r)] for r in configs.cpu().numpy()} == reach
            assert float(energy) >= e0 - 1e-10 * hnorm
    # single_extend: one run on the start basis, one extension from every Krylov vector, one more run
    driver = kr.DynamicLanczos(ham, one, psi1, step=3, threshold=1e-12, count_extend=100_000, batch_count_apply_within=1,
                               single_extend=True, first_extend=False)
    out = list(driver.run())
    sizes = [len(c) for _, c, _ in out]
    assert sizes[0] == 1 and sizes[-1] == len({start} | set(_connected(h, np.array([start])).tolist()))
    assert all(float(e) >= e0 - 1e-10 * hnorm for e, _, _ in out)
