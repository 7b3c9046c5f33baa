"""
qmb_apply_within_noalloc -- apply_within with its scratch in the library's fixed device array
instead of a caller workspace -- called through the C ABI:

  - on the GPU, two calls in a row (the second on a smaller input, so the array holds the first
    call's larger table and index when the second starts) each equal qmb_apply_within with a
    workspace from the caching allocator, and the oracle, to 1e-12 of each entry's sum of
    |contributions|;
  - on the host, a call whose table and index exceed the array is refused with QMB_ERR_WORKSPACE
    before any device work (the sizes are checked first).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import _lib
from qmb_hip import synthetic as sy


def _p(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr() if t.numel() else 0)


def _inputs(case: str, device: torch.device) -> tuple[int, np.ndarray, np.ndarray, tuple[np.ndarray, np.ndarray, np.ndarray], dict]:
    if case == "molecular":
        terms = sorted_terms(*sy.molecular_terms(10, 1001))
        cfg = sy.sector_configs(10, 5, 5, 6000, 1002)
        psi = sy.random_psi(len(cfg), 1003)
    else:
        terms = sorted_terms(*sy.terms_from_dict(sy.hubbard_dict(4, 2, 1.0, 4.0)))
        cfg = sy.sector_configs(8, 4, 4, 900, 1004)
        psi = sy.random_psi(len(cfg), 1005)
    dev = {"cfg": torch.from_numpy(cfg).to(device), "psi": torch.view_as_real(torch.from_numpy(psi)).contiguous().to(device),
           "site": torch.from_numpy(np.ascontiguousarray(terms[0])).to(device),
           "kind": torch.from_numpy(np.ascontiguousarray(terms[1])).to(device),
           "coef": torch.from_numpy(np.ascontiguousarray(terms[2])).to(device)}
    return 1, cfg, psi, terms, dev


@pytest.mark.gpu
def test_noalloc_matches_workspace_entry(cuda_device: torch.device) -> None:
    lib = _lib.library()
    stream = ctypes.c_void_p(torch.cuda.current_stream(cuda_device).cuda_stream)
    for case in ("molecular", "hubbard"):  # the second input is the smaller one
        pc, cfg, psi, terms, d = _inputs(case, cuda_device)
        nq, n, n_terms = cfg.shape[1], len(cfg), len(terms[0])
        size = ctypes.c_size_t(0)
        _lib.check(lib.qmb_apply_within_workspace_size(nq, n, n, n_terms, ctypes.byref(size)), "workspace size")
        ws = torch.empty(max(size.value, 1), dtype=torch.uint8, device=cuda_device)
        want_t = torch.empty((n, 2), dtype=torch.float64, device=cuda_device)
        got_t = torch.full((n, 2), float("nan"), dtype=torch.float64, device=cuda_device)
        args = (nq, pc, _p(d["cfg"]), _p(d["psi"]), n, _p(d["cfg"]), n, _p(d["site"]), _p(d["kind"]), _p(d["coef"]), n_terms)
        _lib.check(lib.qmb_apply_within(*args, _p(want_t), _p(ws), ws.numel(), stream), "qmb_apply_within")
        _lib.check(lib.qmb_apply_within_noalloc(*args, _p(got_t), stream), "qmb_apply_within_noalloc")
        torch.cuda.synchronize()
        got = torch.view_as_complex(got_t).cpu().numpy()
        want = torch.view_as_complex(want_t).cpu().numpy()
        ref, asum, _ = orc.apply_within(pc, cfg, psi, cfg, *terms)
        bound = 1e-12 * asum
        assert np.all(np.abs(got - want) <= bound), case
        assert np.all(np.abs(got - ref) <= bound), case
        assert np.all(got[asum == 0] == 0), case


def test_noalloc_refuses_inputs_above_its_array() -> None:
    lib = _lib.library()
    limit = int(lib.qmb_static_scratch_bytes())
    assert limit == 64 << 20
    fake = ctypes.c_void_p(256)  # never dereferenced: the size check comes first
    status = lib.qmb_apply_within_noalloc(7, 1, fake, fake, 10, fake, 4_000_000, fake, fake, fake, 100, fake, None)
    assert status == _lib.QMB_ERR_WORKSPACE
    assert b"above its fixed" in lib.qmb_last_error()
