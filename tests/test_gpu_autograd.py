"""
GPU checks of the prepared objects, the symmetric (Hermitian) application, the connection counts and
the differentiable apply_within:

  - forward / adjoint through prepared objects (general and symmetric kernels, both strategies)
    against the oracle, within 1e-12 relative to each entry's sum of |contributions|;
  - M_gen / M_hit from the count kernel equal the oracle's counts exactly;
  - autograd backward equals the dense conjugate transpose (non-Hermitian table);
  - the energy <psi|H|psi>/<psi|psi> gradient against central finite differences (step 1e-6) on
    random psi entries, and against PyTorch autograd through a dense matrix built by the oracle;
  - the same central differences at the differentiation workload's full size: the open 64-site
    Hubbard chain (128 spin orbitals, 16-byte configs, t = 1, U = 4, half filling), 4M configs,
    64 random psi entries, real and imaginary parts.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import autograd as ag
from qmb_hip import prepared as pp
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu


def _dev_terms(terms: tuple, device: torch.device) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    return tuple(torch.from_numpy(np.ascontiguousarray(x)).to(device) for x in terms)  # type: ignore[return-value]


CASES = [
    ("molecular-20so", 1, lambda: sorted_terms(*sy.molecular_terms(10, 1001)), lambda: sy.sector_configs(10, 5, 5, 3000, 2)),
    ("hubbard-chain-80so", 1, lambda: tuple(x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(40, 1, 1.0, 4.0))),
     lambda: sy.hubbard_chain_configs(40, 3000, 3, n_seed_configs=512)),
    ("ising-bose2", 2, lambda: tuple(x.numpy() for x in sy.terms_from_dict(sy.ising_dict(4, 3, x=0.7, zh=1.0, yh=0.3))),
     lambda: sy.sort_unique_configs(np.random.default_rng(4).integers(0, 256, size=(700, 2), dtype=np.uint8) & np.array([255, 15], np.uint8))),
]


@pytest.mark.parametrize("strategy", [pp.STRATEGY_TILES, pp.STRATEGY_DIRECT])
@pytest.mark.parametrize("name,pc,terms_fn,cfg_fn", CASES, ids=[c[0] for c in CASES])
def test_prepared_symmetric_and_general(cuda_device: torch.device, name: str, pc: int, terms_fn: object, cfg_fn: object,
                                        strategy: int) -> None:
    s, k, c = terms_fn()
    cfg = cfg_fn()
    psi = sy.random_psi(len(cfg), 5)
    want, asum, _ = orc.apply_within(pc, cfg, psi, cfg, s, k, c)
    st, kt, ct = _dev_terms((s, k, c), cuda_device)
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, cfg.shape[1], pc)
    targets = pp.PreparedTargets(cfg_t, cfg.shape[1])
    assert terms.hermitian and pp.can_use_symmetric(terms, targets, cfg_t)
    vec = torch.view_as_real(torch.from_numpy(psi).to(cuda_device)).contiguous()
    for symmetric in (pp.SYMMETRIC_OFF, pp.SYMMETRIC_GATHER, pp.SYMMETRIC_HALF):
        for mode in (pp.MODE_FORWARD, pp.MODE_ADJOINT):
            out = pp.apply(terms, targets, cfg_t, vec, mode, strategy=strategy, symmetric=symmetric).cpu().numpy()
            got = out[:, 0] + 1j * out[:, 1]
            # Hermitian: the adjoint map equals the forward map
            err = np.abs(got - want)
            assert np.all(err <= 1e-12 * asum + 1e-300), (name, symmetric, mode, err.max())


@pytest.mark.parametrize("name,pc,terms_fn,cfg_fn", CASES, ids=[c[0] for c in CASES])
def test_counts_match_oracle(cuda_device: torch.device, name: str, pc: int, terms_fn: object, cfg_fn: object) -> None:
    s, k, c = terms_fn()
    cfg = cfg_fn()[:500]
    cfg_j = cfg_fn()
    st, kt, ct = _dev_terms((s, k, c), cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, cfg.shape[1], pc, check_hermitian=False)
    targets = pp.PreparedTargets(torch.from_numpy(cfg_j).to(cuda_device), cfg.shape[1])
    for strategy in (pp.STRATEGY_TILES, pp.STRATEGY_DIRECT):
        got = pp.count(terms, targets, torch.from_numpy(cfg).to(cuda_device), strategy=strategy)
        assert got == orc.count_connections(pc, cfg, cfg_j, s, k), (name, strategy)


def test_symmetric_disabled_for_non_hermitian_or_foreign_sets(cuda_device: torch.device) -> None:
    s, k, c = (x.numpy() for x in sy.random_terms(12, 300, 7))
    cfg = sy.sort_unique_configs(np.random.default_rng(8).integers(0, 256, size=(400, 2), dtype=np.uint8) & np.array([255, 15], np.uint8))
    st, kt, ct = _dev_terms((s, k, c), cuda_device)
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, 2, 1)
    targets = pp.PreparedTargets(cfg_t, 2)
    assert not terms.hermitian and not pp.can_use_symmetric(terms, targets, cfg_t)
    dup = torch.cat([cfg_t, cfg_t[:3]])
    herm = pp.PreparedTerms(*_dev_terms(sorted_terms(*sy.molecular_terms(4, 11)), cuda_device), 2, 1)
    dup_targets = pp.PreparedTargets(dup, 2)
    assert dup_targets.duplicates() == 3 and not pp.can_use_symmetric(herm, dup_targets, dup)
    assert not pp.can_use_symmetric(herm, targets, cfg_t.clone())


def test_autograd_backward_is_dense_conjugate_transpose(cuda_device: torch.device) -> None:
    s, k, c = (x.numpy() for x in sy.random_terms(10, 200, 61))
    cfg = sy.sort_unique_configs(np.random.default_rng(62).integers(0, 256, size=(300, 2), dtype=np.uint8) & np.array([255, 3], np.uint8))
    n = len(cfg)
    dense = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[i] = 1.0
        dense[:, i], _, _ = orc.apply_within(1, cfg, e, cfg, s, k, c)
    st, kt, ct = _dev_terms((s, k, c), cuda_device)
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, 2, 1)
    targets = pp.PreparedTargets(cfg_t, 2)
    psi = torch.from_numpy(sy.random_psi(n, 1)).to(cuda_device).requires_grad_(True)
    g = sy.random_psi(n, 2)
    y = ag.apply_within(psi, cfg_t, terms, targets)
    np.testing.assert_allclose(y.detach().cpu().numpy(), dense @ psi.detach().cpu().numpy(), atol=1e-12)
    y.backward(torch.from_numpy(g).to(cuda_device))
    np.testing.assert_allclose(psi.grad.cpu().numpy(), dense.conj().T @ g, atol=1e-12)


def _energy_np(dense: np.ndarray, psi: np.ndarray) -> float:
    return float((np.vdot(psi, dense @ psi) / np.vdot(psi, psi)).real)


def test_energy_gradient_central_differences(cuda_device: torch.device) -> None:
    """Hubbard chain (4 sites, half filling): d E / d psi against central differences, step 1e-6."""
    s, k, c = (x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(4, 1, 1.0, 4.0)))
    cfg = sy.sector_configs(4, 2, 2, 36, 0)
    n = len(cfg)
    dense = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[i] = 1.0
        dense[:, i], _, _ = orc.apply_within(1, cfg, e, cfg, s, k, c)
    st, kt, ct = _dev_terms((s, k, c), cuda_device)
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, 1, 1)
    targets = pp.PreparedTargets(cfg_t, 1)
    psi0 = sy.random_psi(n, 9)
    psi = torch.from_numpy(psi0).to(cuda_device).requires_grad_(True)
    e = ag.energy(psi, cfg_t, terms, targets)
    assert abs(e.item() - _energy_np(dense, psi0)) < 1e-12
    e.backward()
    grad = psi.grad.cpu().numpy()
    h = 1e-6
    rng = np.random.default_rng(10)
    for idx in rng.choice(n, size=12, replace=False):
        for part, unit in (("re", 1.0), ("im", 1j)):
            p_plus, p_minus = psi0.copy(), psi0.copy()
            p_plus[idx] += h * unit
            p_minus[idx] -= h * unit
            fd = (_energy_np(dense, p_plus) - _energy_np(dense, p_minus)) / (2 * h)
            # PyTorch's complex gradient of a real loss is dL/dRe + i dL/dIm
            an = grad[idx].real if part == "re" else grad[idx].imag
            assert abs(fd - an) < 1e-7 * max(1.0, abs(fd)), (idx, part, fd, an)


def test_energy_gradient_central_differences_hubbard128(cuda_device: torch.device) -> None:
    """dE/dpsi from the adjoint kernel against central differences of the energy (forward kernel),
    step 1e-6, on 64 random entries of a 4M-config psi over the 128-spin-orbital chain."""
    s, k, c = (x.contiguous().to(cuda_device) for x in sy.terms_from_dict(sy.hubbard_dict(64, 1, 1.0, 4.0)))
    cfg_t = torch.from_numpy(sy.hubbard_chain_configs(64, 4_000_000, 3000)).to(cuda_device)
    n = cfg_t.size(0)
    assert n == 4_000_000 and cfg_t.size(1) == 16
    terms = pp.PreparedTerms(s, k, c, 16, 1)
    targets = pp.PreparedTargets(cfg_t, 16, distinct=True, max_occupied=64)
    psi0 = torch.from_numpy(sy.random_psi(n, 3001)).to(cuda_device)
    psi = psi0.clone().requires_grad_(True)
    ag.energy(psi, cfg_t, terms, targets).backward()
    grad = psi.grad.cpu().numpy()

    def energy(p: torch.Tensor) -> float:
        with torch.no_grad():
            return float(ag.energy(p, cfg_t, terms, targets))

    h = 1e-6
    # The gradient entries are ~2e-4 here and each energy (sums over 4M rows) carries a few 1e-14 of
    # rounding, so a difference quotient is good to ~2e-8 absolute; a wrong sign, conjugation or
    # dropped term moves an entry by O(|entry|).
    fds, ans = [], []
    for idx in np.random.default_rng(3002).choice(n, size=64, replace=False):
        for part, unit in (("re", 1.0), ("im", 1j)):
            p_plus, p_minus = psi0.clone(), psi0.clone()
            p_plus[idx] += h * unit
            p_minus[idx] -= h * unit
            fd = (energy(p_plus) - energy(p_minus)) / (2 * h)
            an = grad[idx].real if part == "re" else grad[idx].imag
            assert abs(fd - an) <= 1e-4 * abs(an) + 1e-7, (int(idx), part, fd, an)
            fds.append(fd)
            ans.append(an)
    assert np.linalg.norm(np.array(fds) - np.array(ans)) <= 1e-3 * np.linalg.norm(ans)


def test_energy_gradient_matches_dense_autograd(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(4, 3))
    cfg = sy.sector_configs(4, 2, 2, 36, 0)
    n = len(cfg)
    dense = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[i] = 1.0
        dense[:, i], _, _ = orc.apply_within(1, cfg, e, cfg, s, k, c)
    psi0 = sy.random_psi(n, 11)
    ref = torch.from_numpy(psi0).requires_grad_(True)
    h_t = torch.from_numpy(dense)
    e_ref = (torch.vdot(ref, h_t @ ref) / torch.vdot(ref, ref)).real
    e_ref.backward()
    st, kt, ct = _dev_terms((s, k, c), cuda_device)
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, 1, 1)
    targets = pp.PreparedTargets(cfg_t, 1)
    psi = torch.from_numpy(psi0).to(cuda_device).requires_grad_(True)
    e = ag.energy(psi, cfg_t, terms, targets)
    e.backward()
    assert abs(e.item() - e_ref.item()) < 1e-12
    np.testing.assert_allclose(psi.grad.cpu().numpy(), ref.grad.numpy(), atol=1e-12)


def test_drop_in_op_is_differentiable(cuda_device: torch.device) -> None:
    """torch.ops...apply_within has an autograd formula (the adjoint op)."""
    from qmb_hip import Hamiltonian  # pylint: disable=import-outside-toplevel

    ham = Hamiltonian(sy.hubbard_dict(3, 1, 1.0, 2.0), kind="fermi")
    cfg = sy.sector_configs(3, 1, 2, 9, 0)
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    psi = torch.from_numpy(sy.random_psi(9, 1)).to(cuda_device).requires_grad_(True)
    g = torch.from_numpy(sy.random_psi(9, 2)).to(cuda_device)
    ham.apply_within(cfg_t, psi, cfg_t).backward(g)
    # Hermitian table: A^H g = A g
    want = ham.apply_within(cfg_t, g, cfg_t)
    torch.testing.assert_close(psi.grad, want, atol=1e-12, rtol=0)
