"""
GPU parity at each workload's own size, entry by entry:

  - 20 spin orbitals, 5a+5b, 10,000 configs (14,251 terms): the drop-in op (structured path) and
    the enumeration kernel against the reference CPU op's own output stored in
    tests/golden/aw_c1_molecular_20so_10k.npz, to 1e-12 of each entry's sum of |contributions|;
  - 52 spin orbitals, 7a+7b, 669,579 terms: find_relative from a seeded 2,000-row sample of the
    1,000,000 configs with the whole set excluded and K = 1/10 of the sample's contributions
    outside it, against the oracle's selection: the same configs, except ones whose heaviest
    contribution ties the K-th weight (the reference's heap order among equal weights);
  - 64 molecular instances (24..80 spin orbitals, 1e3..5e5 configs): one two-vector batched launch
    against every instance's own structured application, forward and adjoint, on every row to 1e-12
    of the entry's sum of |contributions|; three instances also against the oracle on four rows;
  - 76 spin orbitals, 13a+13b, 3.06M terms, 2,000,000 configs (the training step's set): the
    structured path against the enumeration path on every row, forward and adjoint, to 1e-12 of
    each entry's sum of |contributions|, and six rows against the oracle.

Each entry's sum of |contributions| comes from tests/bounds.py (the enumeration kernel on the
table of |coef| without signs) or, for the stored reference output, from the oracle.
"""

from __future__ import annotations

import pathlib
import sys

import numpy as np
import pytest
import torch

from bounds import abs_bound, abs_terms
from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import batch as qb
from qmb_hip import ops
from qmb_hip import prepared as pp
from qmb_hip import synthetic as sy

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def _say(msg: str) -> None:
    print(f"[configs] {msg}", file=sys.stderr, flush=True)


def _dev(x: np.ndarray, device: torch.device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _real(psi: np.ndarray, device: torch.device) -> torch.Tensor:
    return torch.view_as_real(_dev(np.asarray(psi, dtype=np.complex128), device)).contiguous()


def _complex(t: torch.Tensor) -> np.ndarray:
    return torch.view_as_complex(t.contiguous()).cpu().numpy()


def _oracle_row(cfg: np.ndarray, j: int, vec: np.ndarray, terms: tuple[np.ndarray, ...]) -> tuple[complex, float]:
    """(H vec)_j of a Hermitian table from the oracle's column j of H over the set, and its bound."""
    column, asum, (_, m_hit) = orc.apply_within(1, cfg[j:j + 1], np.ones(1, dtype=np.complex128), cfg, *terms)
    assert m_hit > 1
    return complex(np.vdot(column, vec)), 1e-12 * float(asum @ np.abs(vec))


def test_molecular20_golden_on_gpu(cuda_device: torch.device) -> None:
    with np.load(GOLDEN / "aw_c1_molecular_20so_10k.npz", allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    pc = int(d["particle_cut"])
    s, k, c = d["site"], d["kind"], d["coef"]
    ci, cj = d["configs_i"], d["configs_j"]
    assert ci.shape == (10_000, 3) and len(s) == 14_251
    _, asum, _ = orc.apply_within(pc, ci, d["psi_i"], cj, s, k, c)
    ns = ops.load(3, pc)
    st, kt, ct = _dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device)
    ci_t, cj_t = _dev(ci, cuda_device), _dev(cj, cuda_device)
    psi_t = _real(d["psi_i"], cuda_device)
    results = {"drop-in op": _complex(ns.apply_within(ci_t, psi_t, cj_t, st, kt, ct))}
    terms = pp.PreparedTerms(st, kt, ct, 3, pc, pairs=False)
    targets = pp.PreparedTargets(cj_t, 3)
    results["enumeration kernel"] = _complex(pp.apply(terms, targets, ci_t, psi_t, pp.MODE_FORWARD, symmetric=False))
    for name, got in results.items():
        err = np.abs(got - d["result"])
        assert np.all(err <= 1e-12 * asum), (name, float((err / np.maximum(asum, 1e-300)).max()))
        assert np.all(got[asum == 0] == 0), name


def test_molecular52_find_relative_membership(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(26, 2001, threshold=1e-10))
    cfg = sy.excitation_configs(26, 7, 7, 1_000_000, 2002)
    psi = sy.random_psi(len(cfg), 2003)
    rows = np.sort(np.random.default_rng(2006).choice(len(cfg), size=2000, replace=False))
    sample, sample_psi = cfg[rows], psi[rows]
    st, kt, ct = _dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device)
    cfg_t = _dev(cfg, cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, 7, 1, pairs=False)
    m_gen, m_hit = pp.count(terms, pp.PreparedTargets(cfg_t, 7, distinct=True), _dev(sample, cuda_device))
    assert m_gen == 2000 * 28_701
    k_sel = (m_gen - m_hit) // 10
    _say(f"find_relative: 2000 rows, {m_gen - m_hit} contributions outside the set, K = {k_sel}")
    got = ops.load(7, 1).find_relative(_dev(sample, cuda_device), _real(sample_psi, cuda_device), k_sel, st, kt, ct, cfg_t).cpu().numpy()
    want, kth = orc.find_relative(1, sample, sample_psi, k_sel, s, k, c, cfg)
    _say(f"find_relative: GPU {len(got)} rows, oracle {len(want)} rows, K-th weight {kth:.6e}")
    keys = [r.tobytes() for r in got]
    assert keys == sorted(set(keys))
    diff = sorted(set(keys) ^ {r.tobytes() for r in want})
    if diff:
        diff_rows = np.frombuffer(b"".join(diff), dtype=np.uint8).reshape(-1, 7)
        best = orc.best_weights(1, sample, sample_psi, s, k, c, diff_rows)
        assert np.all(best == kth), (len(diff), float(best.min()), float(best.max()), kth)
    assert len(diff) <= 1000


@pytest.fixture(scope="module")
def batch64(cuda_device: torch.device) -> list[dict[str, object]]:
    return sy.batch_instances(4000, 64, threads=16, progress=lambda d, n: _say(f"batch64: built {d}/{n}") if d % 16 == 0 else None)


def test_batch64_matches_per_instance(cuda_device: torch.device, batch64: list[dict[str, object]]) -> None:
    instances, prepared = batch64, []
    for it in instances:
        cfg_t = _dev(it["configs"], cuda_device)
        nq = cfg_t.size(1)
        terms = pp.PreparedTerms(*(x.contiguous().to(cuda_device) for x in (it["site"], it["kind"], it["coef"])), nq, 1)
        targets = pp.PreparedTargets(cfg_t, nq, distinct=True, max_occupied=sum(it["electrons"]))
        prepared.append((terms, targets))
    _say(f"batch64: {sum(len(it['configs']) for it in instances)} configs prepared")
    b = qb.Batch(prepared, two_vectors=True)
    b.set_inputs([_dev(it["psi"], cuda_device) for it in instances], [_dev(it["grad"], cuda_device) for it in instances])
    b.launch_both()
    for kk, (it, (terms, targets)) in enumerate(zip(instances, prepared)):
        nq = targets.n_qubytes
        t_abs = abs_terms(it["site"].numpy(), it["kind"].numpy(), it["coef"].numpy(), nq, cuda_device)
        for vec, mode, second in ((it["psi"], pp.MODE_FORWARD, False), (it["grad"], pp.MODE_ADJOINT, True)):
            want = _complex(pp.apply(terms, targets, targets.configs, _real(vec, cuda_device), mode))
            got = b.output(kk, second).cpu().numpy()
            bound = 1e-12 * abs_bound(t_abs, targets, targets.configs, vec, mode)
            err = np.abs(got - want)
            assert np.all(err <= bound), (kk, mode, float((err / np.maximum(bound, 1e-300)).max()))
        del t_abs
    _say("batch64: every row of every instance matches")
    for kk in (0, 31, int(np.argmax([len(it["configs"]) for it in instances]))):
        it = instances[kk]
        terms_np = (it["site"].numpy(), it["kind"].numpy(), it["coef"].numpy())
        got_f = b.output(kk).cpu().numpy()
        got_b = b.output(kk, True).cpu().numpy()
        for j in np.random.default_rng(4100 + kk).choice(len(it["configs"]), size=4, replace=False):
            for got, vec in ((got_f, it["psi"]), (got_b, it["grad"])):
                want, bound = _oracle_row(it["configs"], int(j), vec, terms_np)
                assert abs(got[j] - want) <= bound, (kk, int(j), got[j], want, bound)


def test_vmc76_structured_matches_enumeration(cuda_device: torch.device) -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(38, 5001, threshold=1e-8))
    cfg = sy.excitation_configs(38, 13, 13, 2_000_000, 5002)
    assert len(s) == 3_063_651
    st, kt, ct = _dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device)
    cfg_t = _dev(cfg, cuda_device)
    terms = pp.PreparedTerms(st, kt, ct, 10, 1)
    targets = pp.PreparedTargets(cfg_t, 10, distinct=True, max_occupied=26)
    assert pp.can_use_pairs(terms, targets, cfg_t)
    t_abs = abs_terms(s, k, c, 10, cuda_device)
    results = []
    for seed, mode in ((5003, pp.MODE_FORWARD), (5004, pp.MODE_ADJOINT)):
        vec = sy.random_psi(len(cfg), seed)
        structured = _complex(pp.apply(terms, targets, cfg_t, _real(vec, cuda_device), mode))
        enumerated = _complex(pp.apply(terms, targets, cfg_t, _real(vec, cuda_device), mode, pairs=False))
        bound = 1e-12 * abs_bound(t_abs, targets, cfg_t, vec, mode)
        err = np.abs(structured - enumerated)
        assert np.all(err <= bound), (mode, float((err / np.maximum(bound, 1e-300)).max()))
        _say(f"vmc76: mode {mode} every row within bound (max err/bound {float((err / np.maximum(bound, 1e-300)).max()):.3e})")
        results.append((structured, vec))
    for j in np.random.default_rng(5006).choice(len(cfg), size=6, replace=False):
        for got, vec in results:
            want, bound = _oracle_row(cfg, int(j), vec, (s, k, c))
            assert abs(got[j] - want) <= bound, (int(j), got[j], want, bound)
