"""
Host-side checks of the Lanczos driver (qmb_hip.krylov) that need no GPU:

  - the tridiagonal solve (qmb/haar.py:168-183: SciPy stebz, lowest pair only) returns the lowest
    eigenvalue of the dense tridiagonal matrix and combines the Krylov vectors with its eigenvector;
    with no beta it returns the first alpha and vector unchanged;
  - DynamicLanczos validates its arguments as the reference's arithmetic would fail on them
    (batch_count_apply_within < 1 divides by zero there, a psi of the wrong length);
  - qmb_lanczos_step rejects null, aliased and short-workspace arguments before any device work.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from qmb_hip import _lib
from qmb_hip import krylov as kr


def test_tridiagonal_lowest_pair() -> None:
    rng = np.random.default_rng(3)
    m, n = 9, 40
    a = rng.normal(size=m)
    b = np.abs(rng.normal(size=m - 1)) + 0.1
    t = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    vals, vecs = np.linalg.eigh(t)
    basis = [torch.from_numpy(rng.normal(size=n) + 1j * rng.normal(size=n)) for _ in range(m)]
    energy, psi = kr.lowest_ritz_pair([torch.tensor(x) for x in a], [torch.tensor(x) for x in b], basis, torch.device("cpu"))
    assert abs(float(energy) - vals[0]) <= 1e-12 * np.abs(vals).max()
    want = sum(vecs[k, 0] * basis[k].numpy() for k in range(m))
    # eigenvectors are defined up to sign
    sign = np.sign(np.vdot(want, psi.numpy()).real)
    np.testing.assert_allclose(sign * psi.numpy(), want, atol=1e-10)
    e1, p1 = kr.lowest_ritz_pair([torch.tensor(1.5)], [], basis[:1], torch.device("cpu"))
    assert float(e1) == 1.5 and p1 is basis[0]


def test_driver_validates_arguments() -> None:
    configs = torch.zeros((4, 1), dtype=torch.uint8)
    psi = torch.ones(4, dtype=torch.complex128)
    with pytest.raises(ValueError):
        kr.DynamicLanczos(object(), configs, psi, step=2, threshold=1e-8, count_extend=0, batch_count_apply_within=0, single_extend=False,
                          first_extend=False)
    with pytest.raises(RuntimeError):
        kr.DynamicLanczos(object(), configs, psi[:3], step=2, threshold=1e-8, count_extend=0, batch_count_apply_within=1,
                          single_extend=False, first_extend=False)
    with pytest.raises(RuntimeError):
        kr.WithinOperator(object(), configs.to(torch.int64))


def test_lanczos_step_rejects_bad_arguments() -> None:
    lib = _lib.library()
    size = ctypes.c_size_t(0)
    assert lib.qmb_lanczos_workspace_size(ctypes.byref(size)) == 0
    assert size.value >= 2 * 1024 * 8
    assert lib.qmb_lanczos_workspace_size(None) != 0
    ws = ctypes.create_string_buffer(size.value)
    scal = ctypes.create_string_buffer(16)
    bufs = [ctypes.create_string_buffer(16 * 8) for _ in range(4)]
    v, vp, w, vn = (ctypes.cast(b, ctypes.c_void_p) for b in bufs)
    wsp, sp = ctypes.cast(ws, ctypes.c_void_p), ctypes.cast(scal, ctypes.c_void_p)
    assert lib.qmb_lanczos_step(-1, v, vp, 0.0, w, vn, sp, wsp, size.value, None) != 0
    assert lib.qmb_lanczos_step(8, None, vp, 0.0, w, vn, sp, wsp, size.value, None) != 0
    assert lib.qmb_lanczos_step(8, v, vp, 0.0, w, vn, None, wsp, size.value, None) != 0
    assert lib.qmb_lanczos_step(8, v, vp, 0.0, w, vn, sp, wsp, size.value - 1, None) != 0
    assert b"workspace" in lib.qmb_last_error()
    assert lib.qmb_lanczos_step(8, v, vp, 0.0, w, w, sp, wsp, size.value, None) != 0  # v_next aliases w
    assert lib.qmb_lanczos_step(8, v, vp, 0.0, vp, vn, sp, wsp, size.value, None) != 0  # w aliases v_prev
    assert b"distinct" in lib.qmb_last_error()
