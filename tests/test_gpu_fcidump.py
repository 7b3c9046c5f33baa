"""
GPU end-to-end checks of an FCIDUMP model (qmb_hip.fcidump.Model, qmb/fcidump.py:117-199): the
ingested Hamiltonian drives the HIP ops and the Lanczos driver.

  - H2 in the minimal basis: the Lanczos ground-state energy on the GPU equals the 2x2 CI closed
    form of its integrals;
  - a random 6-orbital FCIDUMP (12 spin orbitals, the 3a+3b sector, 400 configs): apply_within
    through the model equals the oracle within 1e-12 of each entry's sum of |contributions|, the
    structured path is taken, and the Lanczos energy equals the lowest eigenvalue of the dense
    sector matrix assembled from oracle columns.
"""

from __future__ import annotations

import pathlib

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from qmb_hip import fcidump as fd
from qmb_hip import krylov as kr
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu


def _model(tmp_path: pathlib.Path, monkeypatch: pytest.MonkeyPatch, name: str, text: str | None = None,
           integrals: tuple[float, np.ndarray, np.ndarray, int] | None = None) -> fd.Model:
    monkeypatch.setenv("QMB_HIP_CACHE_DIR", str(tmp_path / "cache"))
    models = tmp_path / "models"
    models.mkdir(exist_ok=True)
    if text is not None:
        (models / name).write_text(text)
    else:
        assert integrals is not None
        fd.write_fcidump(models / name, *integrals)
    return fd.Model(fd.ModelConfig(name, models))


def test_h2_ground_state_on_gpu(cuda_device: torch.device, tmp_path: pathlib.Path, monkeypatch: pytest.MonkeyPatch) -> None:
    h11, h22, j11, j22, j12, k12, e_nuc = -1.2528, -0.4756, 0.6746, 0.6975, 0.6636, 0.1813, 1 / 1.4
    text = (" &FCI NORB=2,NELEC=2,MS2=0,\n &END\n"
            f"{j11!r} 1 1 1 1\n{k12!r} 2 1 2 1\n{j12!r} 2 2 1 1\n{j22!r} 2 2 2 2\n{h11!r} 1 1 0 0\n{h22!r} 2 2 0 0\n{e_nuc!r} 0 0 0 0\n")
    model = _model(tmp_path, monkeypatch, "H2", text=text)
    cfg = torch.tensor([[(1 << a) | (1 << b)] for a in (0, 2) for b in (1, 3)], dtype=torch.uint8, device=cuda_device)
    energy, _ = kr.lanczos_energy(model, cfg, 10, 1e-12)
    a, b = 2 * h11 + j11, 2 * h22 + j22
    closed = (a + b) / 2 - np.sqrt(((a - b) / 2) ** 2 + k12**2) + e_nuc
    assert abs(energy - closed) < 1e-12


def test_fcidump_model_on_gpu(cuda_device: torch.device, tmp_path: pathlib.Path, monkeypatch: pytest.MonkeyPatch) -> None:
    n = 6
    h, g, e0 = sy.molecular_integrals(n, 61)
    model = _model(tmp_path, monkeypatch, "mol6.FCIDUMP.gz", integrals=(e0, h, g, 6, 0))
    assert (model.n_qubit, model.n_electron, model.n_spin) == (12, 6, 0)
    cfg = sy.sector_configs(n, 3, 3, 400, 0)
    site, kind, coef = (t.numpy() for t in (model.hamiltonian.site, model.hamiltonian.kind, model.hamiltonian.coef))
    psi = sy.random_psi(len(cfg), 62)
    want, asum, _ = orc.apply_within(1, cfg, psi, cfg, site, kind, coef)
    cfg_t = torch.from_numpy(cfg).to(cuda_device)
    got = model.apply_within(cfg_t, torch.from_numpy(psi).to(cuda_device), cfg_t).cpu().numpy()
    assert np.all(np.abs(got - want) <= 1e-12 * asum + 1e-300)
    op = kr.WithinOperator(model, cfg_t)
    assert op.terms is not None and op.terms.pairs_desc is not None and op.terms.n_residual_terms == 0
    dense = np.zeros((len(cfg), len(cfg)), dtype=np.complex128)
    for j in range(len(cfg)):
        dense[:, j] = orc.apply_within(1, cfg[j:j + 1], np.ones(1, dtype=np.complex128), cfg, site, kind, coef)[0]
    e_min = float(np.linalg.eigvalsh(dense)[0])
    driver = kr.DynamicLanczos(model, cfg_t, torch.from_numpy(psi).to(cuda_device), step=150, threshold=1e-12, count_extend=0,
                               batch_count_apply_within=1, single_extend=False, first_extend=False)
    energy = float(list(driver.run())[-1][0])
    assert abs(energy - e_min) <= 1e-9 * np.abs(dense).sum(axis=1).max()
    assert model.show_config(cfg_t[0]).startswith("[") and len(model.show_config(cfg_t[0])) == n + 2
