"""
GPU checks of the perturbation step (qmb_hip.pt2, qmb/pert.py:37-63):

  - with a count large enough to keep every contribution, `other` is exactly the set of configs
    connected to the state outside it, and E0 and E2 equal the Rayleigh quotient and the
    Epstein-Nesbet sum computed from the dense sector matrix (assembled from oracle columns);
  - with the reference's count (len(configs)), E0 and E2 equal the same formulas evaluated with
    the oracle's apply_within and diagonal_term on the selected set, which is distinct, disjoint
    from configs and no larger than the count;
  - a state on the whole sector has no external configs and E2 = 0.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import Hamiltonian
from qmb_hip import pt2
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu


def _ham(s: np.ndarray, k: np.ndarray, c: np.ndarray, device: torch.device) -> Hamiltonian:
    return Hamiltonian(tuple(torch.from_numpy(np.ascontiguousarray(x)).to(device) for x in (s, k, c)), kind="fermi")


def _rows(cfg: np.ndarray) -> list[bytes]:
    return [bytes(r) for r in cfg]


def test_pt2_full_connection_matches_dense(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(4, 3))
    sector = sy.sector_configs(4, 2, 2, 36, 0)
    dense = np.zeros((36, 36), dtype=np.complex128)
    for j in range(36):
        dense[:, j] = orc.apply_within(1, sector[j:j + 1], np.ones(1, dtype=np.complex128), sector, s, k, c)[0]
    pick = np.array([0, 3, 5, 8, 13, 21, 30])
    psi = sy.random_psi(len(pick), 71)
    ham = _ham(s, k, c, cuda_device)
    res = pt2.perturbation(ham, torch.from_numpy(sector[pick]).to(cuda_device), torch.from_numpy(psi).to(cuda_device), count=10**6)
    h_ss = dense[np.ix_(pick, pick)]
    norm2 = np.vdot(psi, psi).real
    e0 = (np.vdot(psi, h_ss @ psi) / norm2).real
    outside = np.setdiff1d(np.arange(36), pick)
    amp = dense[np.ix_(outside, pick)] @ psi
    connected = outside[np.abs(amp) > 0]
    assert set(_rows(res.other.cpu().numpy())) == set(_rows(sector[connected]))
    e2 = float(np.sum(np.abs(dense[np.ix_(connected, pick)] @ psi) ** 2 / norm2 / (e0 - np.diag(dense).real[connected])))
    scale = np.abs(dense).sum(axis=1).max()
    assert abs(res.energy0 - e0) <= 1e-12 * scale
    assert abs(res.energy2 - e2) <= 1e-11 * max(1.0, abs(e2))
    assert res.energy == res.energy0 + res.energy2


def test_pt2_reference_count(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(10, 1001))
    cfg = sy.sector_configs(10, 5, 5, 2000, 7)
    psi = sy.random_psi(len(cfg), 72)
    ham = _ham(s, k, c, cuda_device)
    res = pt2.perturbation(ham, torch.from_numpy(cfg).to(cuda_device), torch.from_numpy(psi).to(cuda_device))
    other = res.other.cpu().numpy()
    assert 0 < len(other) <= len(cfg)
    assert len(set(_rows(other))) == len(other) and not set(_rows(other)) & set(_rows(cfg))
    norm2 = np.vdot(psi, psi).real
    hpsi, asum, _ = orc.apply_within(1, cfg, psi, cfg, s, k, c)
    e0 = (np.vdot(psi, hpsi) / norm2).real
    assert abs(res.energy0 - e0) <= 1e-12 * np.abs(psi).dot(asum) / norm2
    amp, _, _ = orc.apply_within(1, cfg, psi, other, s, k, c)
    diag = orc.diagonal_term(1, other, s, k, c).real
    e2 = float(np.sum(np.abs(amp) ** 2 / norm2 / (e0 - diag)))
    assert abs(res.energy2 - e2) <= 1e-10 * abs(e2)


def test_pt2_whole_sector_has_no_correction(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(4, 3))
    sector = sy.sector_configs(4, 2, 2, 36, 0)
    psi = sy.random_psi(36, 73)
    res = pt2.perturbation(_ham(s, k, c, cuda_device), torch.from_numpy(sector).to(cuda_device), torch.from_numpy(psi).to(cuda_device))
    assert res.other.shape[0] == 0 and res.energy2 == 0.0
