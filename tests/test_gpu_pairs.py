"""
GPU checks of the pair-excitation path (merged pair coefficients + core-group gather kernel + the
residual table on the enumeration kernel) against the oracle and against the enumeration-only path:

  - forward through the pair path equals oracle apply_within within 1e-12 relative to each entry's
    sum of |contributions|, for molecular tables (Hermitian, pc = 1), random non-Hermitian tables
    over two config words (pc = 1 and 2) and every symmetric setting;
  - the adjoint through the pair path equals the enumeration-only adjoint (itself checked against
    the dense conjugate transpose elsewhere);
  - the split is exact: residual terms + merged pair table reproduce the counts of the full table;
  - the pair path is refused for foreign config sets and chosen automatically otherwise.
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


def _pair_heavy_table(sites: np.ndarray, n_terms: int, seed: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Random non-Hermitian table on `sites`: pair excitations in every operator order, hops,
    number-dressed hops and diagonal terms."""
    rng = np.random.default_rng(seed)
    site = np.zeros((n_terms, 4), dtype=np.int16)
    kind = np.full((n_terms, 4), 2, dtype=np.uint8)
    for t in range(n_terms):
        u = rng.random()
        if u < 0.6:
            site[t] = rng.choice(sites, size=4, replace=False)
            kind[t] = rng.permutation([0, 0, 1, 1])
        elif u < 0.8:
            site[t, :2] = rng.choice(sites, size=2, replace=False)
            kind[t, :2] = [1, 0]
        elif u < 0.9:
            p, k, r = rng.choice(sites, size=3, replace=False)
            site[t] = [p, k, k, r]
            kind[t] = [1, 1, 0, 0]
        else:
            p = rng.choice(sites)
            site[t, :2] = [p, p]
            kind[t, :2] = [1, 0]
    coef = rng.normal(size=(n_terms, 2))
    return site, kind, coef


def _fixed_weight_configs(sites: np.ndarray, weight: int, nq: int) -> np.ndarray:
    from itertools import combinations  # pylint: disable=import-outside-toplevel

    rows = []
    for comb in combinations(sites.tolist(), weight):
        bits = np.zeros(8 * nq, dtype=np.uint8)
        bits[list(comb)] = 1
        rows.append(np.packbits(bits, bitorder="little"))
    return sy.sort_unique_configs(np.array(rows, dtype=np.uint8))


def _two_word_case(pc: int) -> tuple[tuple[np.ndarray, np.ndarray, np.ndarray], np.ndarray]:
    sites = np.array([0, 1, 2, 3, 5, 8, 13, 30, 63, 64, 65, 66, 70, 71], dtype=np.int64)  # both words of nq = 9
    return _pair_heavy_table(sites, 900, 40 + pc), _fixed_weight_configs(sites, 5, 9)


CASES = [
    ("molecular-20so", 1, lambda: (sorted_terms(*sy.molecular_terms(10, 1001)), sy.sector_configs(10, 5, 5, 3000, 2))),
    ("molecular-8so-full-sector", 1, lambda: (sorted_terms(*sy.molecular_terms(4, 3)), sy.sector_configs(4, 2, 2, 36, 0))),
    ("random-two-words-fermi", 1, lambda: _two_word_case(1)),
    ("random-two-words-bose2", 2, lambda: _two_word_case(2)),
]


def _dev(x: np.ndarray, device: torch.device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


@pytest.mark.parametrize("name,pc,case_fn", CASES, ids=[c[0] for c in CASES])
def test_pair_path_forward_matches_oracle(cuda_device: torch.device, name: str, pc: int, case_fn: object) -> None:
    (s, k, c), cfg = case_fn()
    nq = cfg.shape[1]
    psi = sy.random_psi(len(cfg), 17)
    want, asum, _ = orc.apply_within(pc, cfg, psi, cfg, s, k, c)
    terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), nq, pc)
    assert terms.pairs_desc is not None and terms.n_pair_terms > 0
    cfg_t = _dev(cfg, cuda_device)
    targets = pp.PreparedTargets(cfg_t, nq)
    assert pp.can_use_pairs(terms, targets, cfg_t)
    vec = torch.view_as_real(_dev(psi, cuda_device)).contiguous()
    settings = [pp.SYMMETRIC_OFF] + ([pp.SYMMETRIC_GATHER, pp.SYMMETRIC_HALF] if terms.hermitian else [])
    for symmetric in settings:
        for strategy in (pp.STRATEGY_TILES, pp.STRATEGY_DIRECT):
            out = pp.apply(terms, targets, cfg_t, vec, pp.MODE_FORWARD, strategy=strategy, symmetric=symmetric, pairs=True)
            got = out[:, 0].cpu().numpy() + 1j * out[:, 1].cpu().numpy()
            err = np.abs(got - want)
            assert np.all(err <= 1e-12 * asum + 1e-300), (name, symmetric, strategy, err.max())


@pytest.mark.parametrize("name,pc,case_fn", CASES, ids=[c[0] for c in CASES])
def test_pair_path_adjoint_matches_enumeration(cuda_device: torch.device, name: str, pc: int, case_fn: object) -> None:
    (s, k, c), cfg = case_fn()
    nq = cfg.shape[1]
    g = sy.random_psi(len(cfg), 23)
    terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), nq, pc)
    cfg_t = _dev(cfg, cuda_device)
    targets = pp.PreparedTargets(cfg_t, nq)
    vec = torch.view_as_real(_dev(g, cuda_device)).contiguous()
    ref = pp.apply(terms, targets, cfg_t, vec, pp.MODE_ADJOINT, symmetric=False, pairs=False).cpu().numpy()
    # the scale of each entry: the adjoint of |A| applied to |g|
    absmax = np.abs(c).sum(axis=1).max()
    scale = np.abs(ref).sum(axis=1) + absmax * np.abs(g).sum()
    for symmetric in ([pp.SYMMETRIC_OFF, pp.SYMMETRIC_GATHER] if terms.hermitian else [pp.SYMMETRIC_OFF]):
        got = pp.apply(terms, targets, cfg_t, vec, pp.MODE_ADJOINT, symmetric=symmetric, pairs=True).cpu().numpy()
        err = np.abs(got - ref).max(axis=1)
        assert np.all(err <= 1e-12 * scale), (name, symmetric, err.max())


def test_pair_split_is_exact(cuda_device: torch.device) -> None:
    """Counts of the residual table plus the pair-excitation term count give back the full table."""
    s, k, c = sorted_terms(*sy.molecular_terms(10, 1001))
    terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), 3, 1)
    full_valid, full_diag, _ = terms.counts()
    off = int(terms.residual_desc.counts) - terms.residual_buffer.data_ptr()
    res = terms.residual_buffer[off:off + 32].view(torch.int64).tolist()
    assert res[0] + terms.n_pair_terms == full_valid
    assert res[1] == full_diag
    # number of pair-excitation terms from the host classifier
    n_pairs = 0
    for row_s, row_k in zip(s, k):
        distinct = len(set(row_s.tolist())) == 4 and np.all(row_k < 2)
        n_pairs += int(distinct and sorted(row_k.tolist()) == [0, 0, 1, 1])
    assert n_pairs == terms.n_pair_terms


def test_pair_path_selection(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(4, 3))
    cfg = sy.sector_configs(4, 2, 2, 36, 0)
    cfg_t = _dev(cfg, cuda_device)
    terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), 1, 1)
    targets = pp.PreparedTargets(cfg_t, 1)
    assert pp.can_use_pairs(terms, targets, cfg_t)
    assert not pp.can_use_pairs(terms, targets, cfg_t.clone())
    vec = torch.view_as_real(_dev(sy.random_psi(36, 3), cuda_device)).contiguous()
    with pytest.raises(RuntimeError):
        pp.apply(terms, targets, cfg_t.clone(), vec, pp.MODE_FORWARD, pairs=True)
    no_pairs = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), 1, 1, pairs=False)
    assert no_pairs.pairs_desc is None and not pp.can_use_pairs(no_pairs, targets, cfg_t)
    # energy gradient through autograd takes the pair path and matches the enumeration-only terms
    psi = _dev(sy.random_psi(36, 5), cuda_device)
    grads = []
    for t in (terms, no_pairs):
        x = psi.clone().requires_grad_(True)
        ag.energy(x, cfg_t, t, targets).backward()
        grads.append(x.grad.cpu().numpy())
    np.testing.assert_allclose(grads[0], grads[1], atol=1e-12)
