"""
GPU parity of find_relative, diagonal_term and single_relative, and of the drop-in Hamiltonian
class, against the CPU oracle.

find_relative bar (reference semantics, qmb/_hamiltonian_cpu.cpp:341-546): every config owning a
contribution strictly heavier than the K-th largest contribution weight is returned, none whose
heaviest contribution is lighter is returned, output is distinct and sorted by bytes.  Weights are
computed without FMA contraction on both sides, so away from ties the sets are identical.
diagonal_term and single_relative (with injected keys) are bit-exact.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from brute import contribution_weights
from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import Hamiltonian, ops
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu


def _to(x: np.ndarray, device: torch.device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _psi_t(psi: np.ndarray, device: torch.device) -> torch.Tensor:
    return torch.view_as_real(torch.from_numpy(psi)).contiguous().to(device)


def _rows(a: np.ndarray) -> set[bytes]:
    return {bytes(r) for r in a}


def _weights(pc: int, cfg: np.ndarray, psi: np.ndarray, exclude: np.ndarray, s: np.ndarray, k: np.ndarray,
             c: np.ndarray) -> dict[bytes, float]:
    """Max contribution weight per reachable non-excluded config (brute force over terms x configs)."""
    best: dict[bytes, float] = {}
    for w, key in contribution_weights(pc, cfg, psi, exclude, s, k, c):
        if w > best.get(key, -1.0):
            best[key] = w
    return best


def _assert_same_selection(name: str, pc: int, cfg: np.ndarray, psi: np.ndarray, exclude: np.ndarray, s: np.ndarray,
                           k: np.ndarray, c: np.ndarray, got: np.ndarray, want: np.ndarray, kth: float) -> None:
    """Distinct rows, strictly increasing by bytes, equal to the oracle's up to the tie class at the K-th weight."""
    keys = [bytes(r) for r in got]
    assert all(a < b for a, b in zip(keys, keys[1:])), f"{name}: rows not strictly increasing by bytes"
    diff = _rows(got) ^ _rows(want)
    if not diff:
        return
    best = _weights(pc, cfg, psi, exclude, s, k, c)
    for key in diff:
        assert best[key] == kth, (name, key, best[key], kth)


FR_CASES = [
    ("molecular-8so", 1, lambda: sorted_terms(*sy.molecular_terms(4, 11)), lambda: sy.sector_configs(4, 2, 2, 12, 3), 20),
    ("molecular-20so", 1, lambda: sorted_terms(*sy.molecular_terms(10, 1001)), lambda: sy.sector_configs(10, 5, 5, 300, 4), 2000),
    ("molecular-20so-small-k", 1, lambda: sorted_terms(*sy.molecular_terms(10, 1001)), lambda: sy.sector_configs(10, 5, 5, 300, 4), 7),
    ("hubbard-4x2-ties", 1, lambda: tuple(x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(4, 2, 1.0, 4.0))),
     lambda: sy.sector_configs(8, 4, 4, 200, 5), 150),
    ("ising-bose2", 2, lambda: tuple(x.numpy() for x in sy.terms_from_dict(sy.ising_dict(4, 2, x=0.8, zh=1.0, zv=0.4))),
     lambda: sy.sort_unique_configs(np.random.default_rng(6).integers(0, 256, size=(60, 1), dtype=np.uint8)), 40),
    ("k-exceeds-candidates", 1, lambda: sorted_terms(*sy.molecular_terms(4, 11)), lambda: sy.sector_configs(4, 2, 2, 5, 3), 100000),
    # at least 4096 rows and K >= 256: the sampled threshold runs first (and is checked on the device)
    ("molecular-20so-sampled", 1, lambda: sorted_terms(*sy.molecular_terms(10, 1001)), lambda: sy.sector_configs(10, 5, 5, 8192, 31),
     4000),
    ("hubbard-4x2-sampled-ties", 1, lambda: tuple(x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(4, 2, 1.0, 4.0))),
     lambda: sy.sector_configs(8, 4, 4, 4500, 32), 300),
    ("hubbard-4x2-sampled-all", 1, lambda: tuple(x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(4, 2, 1.0, 4.0))),
     lambda: sy.sector_configs(8, 4, 4, 4500, 33), 1_000_000),
]


@pytest.mark.parametrize("name,pc,terms_fn,cfg_fn,k_sel", FR_CASES, ids=[c[0] for c in FR_CASES])
def test_find_relative_matches_oracle(cuda_device: torch.device, name: str, pc: int, terms_fn: object, cfg_fn: object,
                                      k_sel: int) -> None:
    s, k, c = terms_fn()
    cfg = cfg_fn()
    psi = sy.random_psi(len(cfg), 9)
    want, kth = orc.find_relative(pc, cfg, psi, k_sel, s, k, c)
    ns = ops.load(cfg.shape[1], pc)
    got = ns.find_relative(_to(cfg, cuda_device), _psi_t(psi, cuda_device), k_sel, _to(s, cuda_device), _to(k, cuda_device),
                           _to(c, cuda_device), _to(cfg, cuda_device)).cpu().numpy()
    _assert_same_selection(name, pc, cfg, psi, cfg, s, k, c, got, want, kth)


def test_find_relative_distinct_exclude_set(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(10, 1001))
    cfg = sy.sector_configs(10, 5, 5, 200, 14)
    exclude = sy.sector_configs(10, 5, 5, 5000, 15)
    psi = sy.random_psi(len(cfg), 16)
    want, kth = orc.find_relative(1, cfg, psi, 3000, s, k, c, exclude)
    got = ops.load(3, 1).find_relative(_to(cfg, cuda_device), _psi_t(psi, cuda_device), 3000, _to(s, cuda_device),
                                       _to(k, cuda_device), _to(c, cuda_device), _to(exclude, cuda_device)).cpu().numpy()
    _assert_same_selection("distinct-exclude", 1, cfg, psi, exclude, s, k, c, got, want, kth)


DIAG_CASES = [
    ("molecular-20so", 1, lambda: sorted_terms(*sy.molecular_terms(10, 1001)), lambda: sy.sector_configs(10, 5, 5, 2000, 21)),
    ("random-ops", 1, lambda: tuple(x.numpy() for x in sy.random_terms(12, 500, 22)),
     lambda: sy.sort_unique_configs(np.random.default_rng(23).integers(0, 256, size=(500, 2), dtype=np.uint8) & np.array([255, 15], np.uint8))),
    ("random-ops-bose2", 2, lambda: tuple(x.numpy() for x in sy.random_terms(12, 500, 22)),
     lambda: sy.sort_unique_configs(np.random.default_rng(23).integers(0, 256, size=(500, 2), dtype=np.uint8) & np.array([255, 15], np.uint8))),
    ("hubbard-chain-80so", 1, lambda: tuple(x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(40, 1, 1.0, 4.0))),
     lambda: sy.hubbard_chain_configs(40, 2000, 24, n_seed_configs=256)),
]


@pytest.mark.parametrize("name,pc,terms_fn,cfg_fn", DIAG_CASES, ids=[c[0] for c in DIAG_CASES])
def test_diagonal_term_bit_exact(cuda_device: torch.device, name: str, pc: int, terms_fn: object, cfg_fn: object) -> None:
    s, k, c = terms_fn()
    cfg = cfg_fn()
    want = orc.diagonal_term(pc, cfg, s, k, c)
    got = ops.load(cfg.shape[1], pc).diagonal_term(_to(cfg, cuda_device), _to(s, cuda_device), _to(k, cuda_device),
                                                   _to(c, cuda_device)).cpu().numpy()
    got = got[:, 0] + 1j * got[:, 1]
    assert np.array_equal(got, want), name


@pytest.mark.parametrize("name,pc,terms_fn,cfg_fn", DIAG_CASES, ids=[c[0] for c in DIAG_CASES])
def test_single_relative_with_keys_bit_exact(cuda_device: torch.device, name: str, pc: int, terms_fn: object,
                                             cfg_fn: object) -> None:
    s, k, c = terms_fn()
    cfg = cfg_fn()
    rng = np.random.default_rng(25)
    keys = rng.random(len(s))
    keys[rng.integers(0, len(s), size=len(s) // 10)] = 0.5  # ties: the lower term index wins
    want = orc.single_relative(pc, cfg, s, k, keys)
    got = ops.load(cfg.shape[1], pc).single_relative_keyed(_to(cfg, cuda_device), _to(s, cuda_device), _to(k, cuda_device),
                                                           _to(c, cuda_device), _to(keys, cuda_device)).cpu().numpy()
    assert np.array_equal(got, want), name


def test_single_relative_sampling_outside_set(cuda_device: torch.device) -> None:
    """Random draws: every returned row is a neighbour outside the set (or zero when none exists)."""
    s, k, c = sorted_terms(*sy.molecular_terms(4, 11))
    cfg = sy.sector_configs(4, 2, 2, 30, 26)
    ns = ops.load(1, 1)
    args = (_to(cfg, cuda_device), _to(s, cuda_device), _to(k, cuda_device), _to(c, cuda_device))
    inset = _rows(cfg)
    for seed in range(5):
        got = ns.single_relative_seeded(*args, seed).cpu().numpy()
        for i in range(len(cfg)):
            neigh = set()
            for t in range(len(s)):
                ok, out, _ = orc.apply_term(cfg[i], s[t], k[t], 1)
                if ok and bytes(out) not in inset:
                    neigh.add(bytes(out))
            if neigh:
                assert bytes(got[i]) in neigh
            else:
                assert not got[i].any()


def test_hamiltonian_class_drop_in(cuda_device: torch.device) -> None:
    d = sy.hubbard_dict(3, 2, 1.0, 4.0)
    ham = Hamiltonian(d, kind="fermi")
    assert ham.particle_cut == 1 and ham.site.dtype == torch.int16 and ham.kind.dtype == torch.uint8
    norms = ham.coef.norm(dim=1)
    assert torch.all(norms[:-1] >= norms[1:])
    cfg = sy.sector_configs(6, 3, 3, 300, 27)
    psi = sy.random_psi(len(cfg), 28)
    s, k, c = ham.site.numpy(), ham.kind.numpy(), ham.coef.numpy()
    want, asum, _ = orc.apply_within(1, cfg, psi, cfg, s, k, c)
    cfg_t = _to(cfg, cuda_device)
    got = ham.apply_within(cfg_t, torch.from_numpy(psi).to(cuda_device), cfg_t).cpu().numpy()
    assert np.all(np.abs(got - want) <= 1e-12 * asum + 1e-300)
    diag = ham.diagonal_term(cfg_t).cpu().numpy()
    assert np.array_equal(diag, orc.diagonal_term(1, cfg, s, k, c))
    fr = ham.find_relative(cfg_t, torch.from_numpy(psi).to(cuda_device), 50)
    assert fr.dtype == torch.uint8 and fr.shape[1] == 2 and len(fr) <= 50
    sr = ham.single_relative(cfg_t)
    assert sr.shape == cfg_t.shape
    with pytest.raises(ValueError):
        Hamiltonian(d, kind="boson")
    # the cached tuple form (qmb/fcidump.py:151-155) is taken as is
    ham2 = Hamiltonian((ham.site.cpu(), ham.kind.cpu(), ham.coef.cpu()), kind="fermi")
    assert torch.equal(ham2.site, ham.site.cpu())
