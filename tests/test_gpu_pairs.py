"""
GPU checks of the structured path (merged pair- and single-excitation coefficients applied by the
core-group gather kernels, the diagonal kernel, and the remaining terms on the enumeration kernel)
against the oracle and against the enumeration-only path:

  - forward through the pair path equals oracle apply_within within 1e-12 relative to each entry's
    sum of |contributions|, for molecular tables (Hermitian, pc = 1), random non-Hermitian tables
    over two config words (pc = 1 and 2) and every symmetric setting;
  - the adjoint through the pair path equals the enumeration-only adjoint (itself checked against
    the dense conjugate transpose elsewhere);
  - the split is exact: every valid term lands in exactly one of the pair table, the single table,
    the diagonal records or the residual enumeration;
  - the pair path is refused for foreign config sets and chosen automatically otherwise.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import _lib
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
        if u < 0.5:  # pair excitation, any operator order
            site[t] = rng.choice(sites, size=4, replace=False)
            kind[t] = rng.permutation([0, 0, 1, 1])
        elif u < 0.65:  # hop
            site[t, :2] = rng.choice(sites, size=2, replace=False)
            kind[t, :2] = [1, 0]
        elif u < 0.75:  # hop dressed by an occupied site: a+_p a+_k a_k a_r
            p, k, r = rng.choice(sites, size=3, replace=False)
            site[t] = [p, k, k, r]
            kind[t] = [1, 1, 0, 0]
        elif u < 0.82:  # hop dressed by an empty site: a+_p a_k a+_k a_r
            p, k, r = rng.choice(sites, size=3, replace=False)
            site[t] = [p, k, k, r]
            kind[t] = [1, 0, 1, 0]
        elif u < 0.9:  # diagonal: n_p or n_p n_q
            p, q = rng.choice(sites, size=2, replace=False)
            if rng.random() < 0.5:
                site[t, :2] = [p, p]
                kind[t, :2] = [1, 0]
            else:
                site[t] = [p, q, q, p]
                kind[t] = [1, 1, 0, 0]
        else:  # neither: a+_p a+_q a_r (changes the particle number)
            site[t, :3] = rng.choice(sites, size=3, replace=False)
            kind[t, :3] = [1, 1, 0]
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


def _compile(site4: np.ndarray, kind4: np.ndarray, pc: int, nq: int) -> tuple[int, int]:
    lib = _lib.library()
    s = np.ascontiguousarray(site4, dtype=np.int16)
    k = np.ascontiguousarray(kind4, dtype=np.uint8)
    rec, status, bucket, diag = ctypes.c_uint64(0), ctypes.c_int32(0), ctypes.c_int64(0), ctypes.c_int32(0)
    _lib.check(lib.qmb_compile_term_host(s.ctypes.data, k.ctypes.data, pc, nq, ctypes.byref(rec), ctypes.byref(status),
                                         ctypes.byref(bucket), ctypes.byref(diag)), "compile")
    return rec.value, status.value


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


@pytest.mark.parametrize("name,pc,case_fn", CASES, ids=[c[0] for c in CASES])
def test_structured_split_is_exact(cuda_device: torch.device, name: str, pc: int, case_fn: object) -> None:
    """Every valid term goes to exactly one of: pair table, single table, diagonal records, residual
    enumeration; the device counts agree with the host classifiers."""
    (s, k, c), cfg = case_fn()
    nq = cfg.shape[1]
    terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), nq, pc)
    full_valid, full_diag, _ = terms.counts()
    assert terms.n_residual_terms + terms.n_pair_terms + terms.n_single_terms + full_diag == full_valid, name
    off = int(terms.pairs_desc.counts) - terms.pairs_buffer.data_ptr()
    assert int(terms.pairs_buffer[off + 8:off + 16].view(torch.int64).item()) == full_diag  # every diagonal term folded
    # real tables switch the pair kernels to 8-byte coefficients, and no merged entry has an imaginary part
    assert terms.pairs_desc.real == int(np.all(np.asarray(c)[:, 1] == 0)), name
    if terms.pairs_desc.real:
        assert int(terms.pairs_buffer[off + 16:off + 24].view(torch.int64).item()) == 0
    lib = _lib.library()
    out = (ctypes.c_int32 * 5)()
    n_pairs = n_singles = 0
    for row_s, row_k in zip(s, k):
        rec, status = _compile(row_s, row_k, pc, nq)
        if status != 0:
            continue
        n_pairs += lib.qmb_pair_excitation_host(rec, pc, out)
        n_singles += lib.qmb_single_excitation_host(rec, pc, out)
    assert (n_pairs, n_singles) == (terms.n_pair_terms, terms.n_single_terms), name
    if name.startswith("molecular"):
        assert terms.n_residual_terms == 0  # one-body, two-body and constant terms are all structured


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


ACROSS = ["disjoint", "subset", "superset", "overlap", "repeated-source-rows", "repeated-missing-rows", "repeated-target-rows"]


@pytest.mark.parametrize("case", ACROSS)
def test_apply_across_matches_oracle(cuda_device: torch.device, case: str) -> None:
    """The structured path between two config sets (one joint set holding both) against the oracle,
    forward and adjoint; a target set with repeated rows is left to the enumeration kernel."""
    s, k, c = sorted_terms(*sy.molecular_terms(10, 1001))
    pool = sy.sector_configs(10, 5, 5, 3000, 51)
    rng = np.random.default_rng(52)
    perm = rng.permutation(len(pool))
    a, b = pool[perm[:1200]], pool[perm[1200:2400]]
    cfg_i, cfg_j = {
        "disjoint": (a, b),
        "subset": (a[:500], a),
        "superset": (a, a[:500]),
        "overlap": (np.concatenate([a[:700], b[:300]]), np.concatenate([a[400:900], b[200:900]])),
        "repeated-source-rows": (np.concatenate([a[:600], a[:100]]), a),  # repeats that configs_j holds: inputs add up
        "repeated-missing-rows": (np.concatenate([a, a[:100]]), b),  # repeats that configs_j lacks: not covered
        "repeated-target-rows": (a, np.concatenate([b, b[:50]])),
    }[case]
    terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), 3, 1)
    psi = sy.random_psi(len(cfg_i), 53)
    g = sy.random_psi(len(cfg_j), 54)
    ci, cj = _dev(cfg_i, cuda_device), _dev(cfg_j, cuda_device)
    fwd = pp.apply_across(terms, ci, torch.view_as_real(_dev(psi, cuda_device)).contiguous(), cj, pp.MODE_FORWARD)
    adj = pp.apply_across(terms, ci, torch.view_as_real(_dev(g, cuda_device)).contiguous(), cj, pp.MODE_ADJOINT)
    if case in ("repeated-target-rows", "repeated-missing-rows"):
        assert fwd is None and adj is None
        return
    want, asum, _ = orc.apply_within(1, cfg_i, psi, cfg_j, s, k, c)
    got = fwd.cpu().numpy()
    assert np.all(np.abs(got[:, 0] + 1j * got[:, 1] - want) <= 1e-12 * asum + 1e-300), case
    # adjoint: <g, A psi> = <A^H g, psi> for every psi, so compare with the dense columns' conjugate sums
    ref = pp.apply(terms, pp.PreparedTargets(cj, 3), ci, torch.view_as_real(_dev(g, cuda_device)).contiguous(), pp.MODE_ADJOINT,
                   pairs=False).cpu().numpy()
    got_a = adj.cpu().numpy()
    scale = np.abs(ref).sum(axis=1) + np.abs(c).sum(axis=1).max() * np.abs(g).sum()
    assert np.all(np.abs(got_a - ref).max(axis=1) <= 1e-12 * scale), case


def test_too_low_max_occupied_raises(cuda_device: torch.device) -> None:
    """A max_occupied bound below a row's occupancy is an error from the core index build, not a
    NaN row or missing pair terms."""
    s, k, c = sorted_terms(*sy.molecular_terms(6, 21))
    cfg = torch.from_numpy(sy.sector_configs(6, 3, 3, 120, 22)).to(cuda_device)
    terms = pp.PreparedTerms(torch.from_numpy(s).to(cuda_device), torch.from_numpy(k).to(cuda_device),
                             torch.from_numpy(c).to(cuda_device), 2, 1)
    vec = torch.zeros((cfg.size(0), 2), dtype=torch.float64, device=cuda_device)
    low = pp.PreparedTargets(cfg, 2, distinct=True, max_occupied=5)
    with pytest.raises(RuntimeError, match="max_occupied=5"):
        pp.apply(terms, low, cfg, vec, pp.MODE_FORWARD)
    exact = pp.PreparedTargets(cfg, 2, distinct=True, max_occupied=6)
    assert torch.isfinite(pp.apply(terms, exact, cfg, vec, pp.MODE_FORWARD)).all()


def _core_index(targets: pp.PreparedTargets, order: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(vals, starts with their flag bit, counts) of a built core index, copied to the host."""
    desc = targets.cores(order)
    buf = targets._cores[order][0]  # pylint: disable=protected-access
    base = buf.data_ptr()
    counts = buf[int(desc.counts) - base:int(desc.counts) - base + 32].view(torch.int64).cpu().numpy()
    entries, groups = int(counts[0]), int(counts[1])
    vals = buf[int(desc.vals) - base:int(desc.vals) - base + 8 * entries].view(torch.int64).cpu().numpy().astype(np.uint64)
    off = int(desc.starts) - base
    starts = buf[off:off + 4 * (groups + 1)].view(torch.int32).cpu().numpy().astype(np.uint32)
    return vals, starts, counts


def _build_cores_with_narrow_hash(targets: pp.PreparedTargets, bits: int, seed: int) -> None:
    """Builds both core indexes of the set from random `bits`-bit hash columns, so that many
    different cores share a key, then puts the set's own columns back (its hash index, built later
    on first use, keeps them)."""
    off = int(targets.keys_desc.cols) - targets.buffer.data_ptr()
    view = targets.buffer[off:off + 8 * 256].view(torch.int64)
    own = view.clone()
    rng = np.random.default_rng(seed)
    view.copy_(torch.from_numpy(rng.integers(0, 1 << bits, size=256, dtype=np.int64)))
    targets.cores(2)
    targets.cores(1)
    view.copy_(own)


MERGE_CASES = [c for c in CASES if c[0] in ("molecular-20so", "random-two-words-fermi")]


@pytest.mark.parametrize("bits", [3, 9, 12, 14])
@pytest.mark.parametrize("name,pc,case_fn", MERGE_CASES, ids=[c[0] for c in MERGE_CASES])
def test_merged_core_groups(cuda_device: torch.device, name: str, pc: int, case_fn: object, bits: int) -> None:
    """Core keys from narrowed hash columns merge many different cores into one group (3 bits: a
    few thousand entries a group; 14 bits: a few).  The build flags exactly the groups holding more
    than one core, and the forward and adjoint through those mixed groups -- every group class:
    2..8 entries, 9..32 packed, larger -- still equal the oracle and the enumeration."""
    (s, k, c), cfg = case_fn()
    nq = cfg.shape[1]
    cfg_t = _dev(cfg, cuda_device)
    terms = pp.PreparedTerms(_dev(s, cuda_device), _dev(k, cuda_device), _dev(c, cuda_device), nq, pc)
    targets = pp.PreparedTargets(cfg_t, nq, distinct=True)
    _build_cores_with_narrow_hash(targets, bits, 7 + bits)
    bitsof = np.unpackbits(cfg, axis=1, bitorder="little")
    sizes_seen = set()
    n_mixed = 0
    for order in (2, 1):
        vals, starts, _ = _core_index(targets, order)
        flag = starts[:-1] >> 31
        st = (starts & 0x7FFFFFFF).astype(np.int64)
        sizes = np.diff(st)
        assert st[0] == 0 and st[-1] == len(vals) and (sizes > 0).all()
        # the core of every entry: its row's sites without the removed ones
        rows = (vals >> np.uint64(16)).astype(np.int64)
        core = bitsof[rows].copy()
        e = np.arange(len(vals))
        core[e, (vals & np.uint64(0xFF)).astype(np.int64)] = 0
        if order == 2:
            core[e, ((vals >> np.uint64(8)) & np.uint64(0xFF)).astype(np.int64)] = 0
        packed = np.packbits(core, axis=1, bitorder="little")
        first = np.repeat(st[:-1], sizes)
        differs = (packed != packed[first]).any(axis=1).astype(np.int64)
        mixed = np.add.reduceat(differs, st[:-1]) > 0
        np.testing.assert_array_equal(flag.astype(bool), mixed)
        n_mixed += int(mixed.sum())
        sizes_seen.update(np.unique(np.minimum(sizes[mixed], 33)).tolist())
    if name.startswith("molecular"):
        # ~40k distinct cores: merged at every width, down to groups of 2..4 entries at 14 bits
        assert n_mixed > 0
        if bits == 14:
            assert {2, 3, 4}.issubset(sizes_seen), sizes_seen
    elif bits <= 9:
        assert n_mixed > 0  # a few hundred distinct cores merge only under narrow keys
    psi = sy.random_psi(len(cfg), 61)
    vec = torch.view_as_real(_dev(psi, cuda_device)).contiguous()
    want, asum, _ = orc.apply_within(pc, cfg, psi, cfg, s, k, c)
    got = pp.apply(terms, targets, cfg_t, vec, pp.MODE_FORWARD, symmetric=pp.SYMMETRIC_OFF, pairs=True).cpu().numpy()
    err = np.abs(got[:, 0] + 1j * got[:, 1] - want)
    assert np.all(err <= 1e-12 * asum + 1e-300), (name, bits, err.max())
    ref = pp.apply(terms, pp.PreparedTargets(cfg_t, nq), cfg_t, vec, pp.MODE_ADJOINT, symmetric=False, pairs=False).cpu().numpy()
    got_a = pp.apply(terms, targets, cfg_t, vec, pp.MODE_ADJOINT, symmetric=False, pairs=True).cpu().numpy()
    scale = np.abs(ref).sum(axis=1) + np.abs(c).sum(axis=1).max() * np.abs(psi).sum()
    assert np.all(np.abs(got_a - ref).max(axis=1) <= 1e-12 * scale), (name, bits)

