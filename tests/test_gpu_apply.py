"""
GPU parity of apply_within, its adjoint and the reference-scheme kernel against the CPU oracle.

Bar: FP64 outputs within 1e-12 of the oracle relative to the sum of |contributions| of each entry
(the only slack is the order of atomic additions); exact zeros where the oracle has no contribution.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import ops
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu

REL = 1e-12


def _cases() -> list[tuple[str, int, tuple[np.ndarray, np.ndarray, np.ndarray], np.ndarray, np.ndarray]]:
    cases = []
    s, k, c = sorted_terms(*sy.molecular_terms(4, 11))
    cfg = sy.sector_configs(4, 2, 2, 36, 12)
    cases.append(("molecular-8so-full-sector", 1, (s, k, c), cfg, cfg))
    s, k, c = sorted_terms(*sy.molecular_terms(10, 1001))
    cfg = sy.sector_configs(10, 5, 5, 3000, 1002)
    cases.append(("molecular-20so-3000", 1, (s, k, c), cfg, cfg))
    cfg_i = sy.sector_configs(10, 5, 5, 1500, 7)
    cfg_j = sy.sector_configs(10, 5, 5, 2500, 8)
    cases.append(("molecular-20so-disjoint-sets", 1, (s, k, c), cfg_i, cfg_j))
    s, k, c = sy.random_terms(12, 400, 21)
    cfg = sy.sort_unique_configs(np.random.default_rng(22).integers(0, 256, size=(600, 2), dtype=np.uint8) & np.array([255, 15], np.uint8))
    cases.append(("random-ops-12-sites", 1, (s.numpy(), k.numpy(), c.numpy()), cfg, cfg))
    cases.append(("random-ops-12-sites-bose2", 2, (s.numpy(), k.numpy(), c.numpy()), cfg, cfg))
    s, k, c = sy.terms_from_dict(sy.hubbard_dict(3, 2, 1.0, 4.0))
    cfg = sy.sector_configs(6, 3, 3, 400, 31)
    cases.append(("hubbard-3x2", 1, (s.numpy(), k.numpy(), c.numpy()), cfg, cfg))
    s, k, c = sy.terms_from_dict(sy.ising_dict(4, 3, x=0.7, zh=1.0, zv=0.5, yh=0.3))
    cfg = sy.sort_unique_configs(np.random.default_rng(32).integers(0, 256, size=(700, 2), dtype=np.uint8) & np.array([255, 15], np.uint8))
    cases.append(("ising-4x3-bose2", 2, (s.numpy(), k.numpy(), c.numpy()), cfg, cfg))
    s, k, c = sy.terms_from_dict(sy.hubbard_dict(40, 1, 1.0, 4.0))  # 80 spin orbitals -> 10 bytes, two words
    cfg = sy.hubbard_chain_configs(40, 3000, 41, n_seed_configs=512)
    cases.append(("hubbard-chain-80so", 1, (s.numpy(), k.numpy(), c.numpy()), cfg, cfg))
    s, k, c = sy.random_terms(200, 300, 51)  # 25 bytes, four words
    rng = np.random.default_rng(52)
    base = rng.integers(0, 256, size=(64, 25), dtype=np.uint8)
    cfg = sy.sort_unique_configs(np.concatenate([base, base ^ (rng.integers(0, 2, size=(64, 25)) << rng.integers(0, 8, size=(64, 25))).astype(np.uint8)]))
    cases.append(("random-ops-200-sites", 1, (s.numpy(), k.numpy(), c.numpy()), cfg, cfg))
    return cases


CASES = _cases()


def _to(x: np.ndarray, device: torch.device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _complex(t: torch.Tensor) -> np.ndarray:
    a = t.cpu().numpy()
    return a[:, 0] + 1j * a[:, 1]


def _assert_close(got: np.ndarray, want: np.ndarray, asum: np.ndarray, what: str) -> None:
    err = np.abs(got - want)
    bound = REL * asum + 1e-300
    bad = err > bound
    assert not bad.any(), f"{what}: {bad.sum()} entries off, worst err {err.max():.3e} vs bound {bound[np.argmax(err)]:.3e}"
    assert np.all(got[asum == 0] == 0), f"{what}: nonzero output where the oracle has no contribution"


@pytest.mark.parametrize("name,pc,terms,cfg_i,cfg_j", CASES, ids=[c[0] for c in CASES])
def test_apply_within_matches_oracle(cuda_device: torch.device, name: str, pc: int, terms: tuple, cfg_i: np.ndarray,
                                     cfg_j: np.ndarray) -> None:
    s, k, c = terms
    psi = sy.random_psi(len(cfg_i), 5)
    want, asum, _ = orc.apply_within(pc, cfg_i, psi, cfg_j, s, k, c)
    ns = ops.load(cfg_i.shape[1], pc)
    psi_t = torch.view_as_real(torch.from_numpy(psi)).contiguous().to(cuda_device)
    args = (_to(cfg_i, cuda_device), psi_t, _to(cfg_j, cuda_device), _to(s, cuda_device), _to(k, cuda_device), _to(c, cuda_device))
    got = _complex(ns.apply_within(*args))
    _assert_close(got, want, asum, name)
    plain = _complex(ns.apply_within_plain(*args))
    _assert_close(plain, want, asum, name + " (plain kernel)")
    if cfg_i.shape == cfg_j.shape and np.array_equal(cfg_i, cfg_j):
        # one tensor for both sets: the structured path when the table has pair excitations
        same = _to(cfg_i, cuda_device)
        got_same = _complex(ns.apply_within(same, psi_t, same, *args[3:]))
        _assert_close(got_same, want, asum, name + " (one config set)")


@pytest.mark.parametrize("name,pc,terms,cfg_i,cfg_j", CASES, ids=[c[0] for c in CASES])
def test_adjoint_identity(cuda_device: torch.device, name: str, pc: int, terms: tuple, cfg_i: np.ndarray, cfg_j: np.ndarray) -> None:
    """<g, A psi> == <A^H g, psi> with A from the oracle and A^H from the adjoint kernel."""
    s, k, c = terms
    psi = sy.random_psi(len(cfg_i), 6)
    g = sy.random_psi(len(cfg_j), 7)
    a_psi, _, _ = orc.apply_within(pc, cfg_i, psi, cfg_j, s, k, c)
    ns = ops.load(cfg_i.shape[1], pc)
    g_t = torch.view_as_real(torch.from_numpy(g)).contiguous().to(cuda_device)
    adj = _complex(ns.apply_within_adjoint(_to(cfg_i, cuda_device), g_t, _to(cfg_j, cuda_device), _to(s, cuda_device),
                                           _to(k, cuda_device), _to(c, cuda_device)))
    lhs = np.vdot(g, a_psi)
    rhs = np.vdot(adj, psi)
    scale = np.sum(np.abs(g) * np.abs(a_psi)) + 1e-300
    assert abs(lhs - rhs) <= 1e-12 * max(scale, abs(lhs)) + 1e-15, (name, lhs, rhs)


def test_adjoint_matches_dense_transpose(cuda_device: torch.device) -> None:
    """Entrywise: the adjoint of a non-Hermitian term table equals conj-transpose applied by the oracle."""
    s, k, c = sy.random_terms(10, 200, 61)
    s, k, c = s.numpy(), k.numpy(), c.numpy()
    cfg = sy.sort_unique_configs(np.random.default_rng(62).integers(0, 256, size=(300, 2), dtype=np.uint8) & np.array([255, 3], np.uint8))
    n = len(cfg)
    # dense A columns from the oracle: A[:, i] = apply(e_i)
    dense = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[i] = 1.0
        dense[:, i], _, _ = orc.apply_within(1, cfg, e, cfg, s, k, c)
    g = sy.random_psi(n, 63)
    want = dense.conj().T @ g
    ns = ops.load(2, 1)
    g_t = torch.view_as_real(torch.from_numpy(g)).contiguous().to(cuda_device)
    got = _complex(ns.apply_within_adjoint(_to(cfg, cuda_device), g_t, _to(cfg, cuda_device), _to(s, cuda_device), _to(k, cuda_device),
                                           _to(c, cuda_device)))
    asum = np.abs(dense.conj().T) @ np.abs(g)
    _assert_close(got, want, asum, "adjoint vs dense")


def test_empty_inputs(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(4, 11))
    ns = ops.load(1, 1)
    cfg = _to(sy.sector_configs(4, 2, 2, 10, 1), cuda_device)
    psi = torch.zeros((10, 2), dtype=torch.float64, device=cuda_device)
    e_cfg = torch.zeros((0, 1), dtype=torch.uint8, device=cuda_device)
    e_psi = torch.zeros((0, 2), dtype=torch.float64, device=cuda_device)
    st, kt, ct = _to(s, cuda_device), _to(k, cuda_device), _to(c, cuda_device)
    assert ns.apply_within(e_cfg, e_psi, cfg, st, kt, ct).abs().sum().item() == 0
    assert ns.apply_within(cfg, psi, e_cfg, st, kt, ct).shape == (0, 2)
    no_terms = (st[:0], kt[:0], ct[:0])
    assert ns.apply_within(cfg, psi + 1, cfg, *no_terms).abs().sum().item() == 0


def test_duplicate_targets_lowest_row_receives(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(4, 11))
    cfg = sy.sector_configs(4, 2, 2, 36, 12)
    psi = sy.random_psi(36, 13)
    want, _, _ = orc.apply_within(1, cfg, psi, cfg, s, k, c)
    dup = np.concatenate([cfg, cfg[:5]])
    ns = ops.load(1, 1)
    psi_t = torch.view_as_real(torch.from_numpy(psi)).contiguous().to(cuda_device)
    got = _complex(ns.apply_within(_to(cfg, cuda_device), psi_t, _to(dup, cuda_device), _to(s, cuda_device), _to(k, cuda_device),
                                   _to(c, cuda_device)))
    np.testing.assert_allclose(got[:36], want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert np.all(got[36:] == 0)


def test_argument_checks(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(4, 11))
    ns = ops.load(1, 1)
    cfg = _to(sy.sector_configs(4, 2, 2, 10, 1), cuda_device)
    psi = torch.zeros((10, 2), dtype=torch.float64, device=cuda_device)
    st, kt, ct = _to(s, cuda_device), _to(k, cuda_device), _to(c, cuda_device)
    with pytest.raises(RuntimeError, match="psi must be float64"):
        ns.apply_within(cfg, psi.float(), cfg, st, kt, ct)
    with pytest.raises(RuntimeError, match="must be contiguous"):
        ns.apply_within(cfg, torch.zeros((2, 10), dtype=torch.float64, device=cuda_device).t(), cfg, st, kt, ct)
    bad = st.clone()
    bad[0, 0] = 9  # site 9 does not exist in a 1-byte config
    kb = kt.clone()
    kb[0, 0] = 1
    with pytest.raises(RuntimeError, match="site outside"):
        ns.apply_within(cfg, psi, cfg, bad, kb, ct)


def test_op_structured_path_cache(cuda_device: torch.device) -> None:
    """apply_within on one config tensor keeps the compiled table per term tensors, rebuilds it after
    an in-place change of the coefficients, and differentiates through the same path."""
    s, k, c = sorted_terms(*sy.molecular_terms(4, 11))
    cfg = sy.sector_configs(4, 2, 2, 36, 14)
    psi = sy.random_psi(36, 15)
    ns = ops.load(1, 1)
    cfg_t = _to(cfg, cuda_device)
    st, kt, ct = _to(s, cuda_device), _to(k, cuda_device), _to(c.copy(), cuda_device)
    psi_t = torch.view_as_real(torch.from_numpy(psi)).contiguous().to(cuda_device)
    ops._PREPARED.clear()  # pylint: disable=protected-access
    first = ns.apply_within(cfg_t, psi_t, cfg_t, st, kt, ct)
    assert len(ops._PREPARED) == 1  # pylint: disable=protected-access
    second = ns.apply_within(cfg_t, psi_t, cfg_t, st, kt, ct)
    assert len(ops._PREPARED) == 1 and torch.equal(first, second)  # pylint: disable=protected-access
    want, asum, _ = orc.apply_within(1, cfg, psi, cfg, s, k, 2 * c)
    ct.mul_(2.0)
    doubled = _complex(ns.apply_within(cfg_t, psi_t, cfg_t, st, kt, ct))
    assert len(ops._PREPARED) == 2  # pylint: disable=protected-access
    _assert_close(doubled, want, 2 * asum, "after an in-place change")
    # the gradient through the op equals the enumeration adjoint on a copy of the set
    x = psi_t.clone().requires_grad_(True)
    g = torch.view_as_real(torch.from_numpy(sy.random_psi(36, 16))).contiguous().to(cuda_device)
    (ns.apply_within(cfg_t, x, cfg_t, st, kt, ct) * g).sum().backward()
    ref = ns.apply_within_adjoint(cfg_t, g, cfg_t.clone(), st, kt, ct)
    torch.testing.assert_close(x.grad, ref, rtol=0, atol=1e-12 * float(ref.abs().max()))


def test_within_operator_sees_in_place_coefficient_change(cuda_device: torch.device) -> None:
    """Hamiltonian.prepared_terms (behind the Krylov WithinOperator) keys its compiled table on the
    term tensors' versions too: after ham.coef.mul_(2) a new operator applies the doubled table."""
    from qmb_hip import Hamiltonian  # pylint: disable=import-outside-toplevel
    from qmb_hip import krylov as kr  # pylint: disable=import-outside-toplevel

    s, k, c = sorted_terms(*sy.molecular_terms(4, 11))
    ham = Hamiltonian((torch.from_numpy(s), torch.from_numpy(k), torch.from_numpy(c.copy())), kind="fermi")
    cfg = sy.sector_configs(4, 2, 2, 36, 14)
    psi = sy.random_psi(36, 15)
    cfg_t = _to(cfg, cuda_device)
    psi_c = torch.from_numpy(psi).to(cuda_device)
    before = kr.WithinOperator(ham, cfg_t)(psi_c).cpu().numpy()
    want, asum, _ = orc.apply_within(1, cfg, psi, cfg, s, k, c)
    _assert_close(before, want, asum, "before the change")
    ham.coef.mul_(2.0)
    after = kr.WithinOperator(ham, cfg_t)(psi_c).cpu().numpy()
    _assert_close(after, 2 * want, 2 * asum, "after ham.coef.mul_(2)")
