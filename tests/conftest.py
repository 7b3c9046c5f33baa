"""Shared pytest configuration: import paths, the `gpu` marker and common fixtures."""

from __future__ import annotations

import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT / "qmb-public_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def cuda_device() -> object:
    import torch  # pylint: disable=import-outside-toplevel

    if not torch.cuda.is_available():
        pytest.skip("no HIP device available")
    return torch.device("cuda:0")


def sorted_terms(site: object, kind: object, coef: object) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Terms ordered by |coef| descending (stable), as the Hamiltonian class stores them."""
    s = np.asarray(site)
    k = np.asarray(kind)
    c = np.asarray(coef)
    order = np.argsort(-np.linalg.norm(c, axis=1), kind="stable")
    return s[order], k[order], c[order]
