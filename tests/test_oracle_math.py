"""
The oracle against what the mathematics fixes, independently of the reference's bit loop:

  - dense Jordan-Wigner matrices built from Kronecker products (a_j = Z_0 ... Z_{j-1} sigma^-_j for
    fermions, a_j = sigma^-_j for hard-core bosons), terms multiplied left to right;
  - Hermiticity of Hermitian term tables, linearity;
  - closed forms: free-fermion ground energy = sum of the lowest single-particle levels
    (qmb/free_fermion.py:68-73), Hubbard dimer E0 = (U - sqrt(U^2 + 16 t^2)) / 2, transverse-field
    Ising energies from Pauli Kronecker products;
  - diagonal_term(c) == apply_within([c], [1], [c]); single_relative by brute force.
"""

from __future__ import annotations

import functools
import math

import numpy as np
import pytest

from conftest import sorted_terms
from oracle import oracle as orc
from qmb_hip import synthetic as sy

SM = np.array([[0.0, 1.0], [0.0, 0.0]])  # |0><1|: annihilates an occupied site
ID = np.eye(2)
Z = np.diag([1.0, -1.0])


def _kron_sites(factors: list[np.ndarray]) -> np.ndarray:
    """factors[s] acts on site s; basis index = sum_s bit_s 2^s (site 0 least significant)."""
    return functools.reduce(np.kron, reversed(factors))


def _op(n: int, site: int, kind: int, fermion: bool) -> np.ndarray:
    f = [ID] * n
    for s in range(site):
        f[s] = Z if fermion else ID
    f[site] = SM if kind == 0 else SM.T
    return _kron_sites(f)


def dense_hamiltonian(n: int, site: np.ndarray, kind: np.ndarray, coef: np.ndarray, fermion: bool) -> np.ndarray:
    dim = 2**n
    h = np.zeros((dim, dim), dtype=np.complex128)
    cache: dict[tuple[int, int], np.ndarray] = {}
    for t in range(len(site)):
        m = np.eye(dim)
        for op in range(4):
            if kind[t, op] == 2:
                continue
            key = (int(site[t, op]), int(kind[t, op]))
            if key not in cache:
                cache[key] = _op(n, *key, fermion)
            m = m @ cache[key]
        h += (coef[t, 0] + 1j * coef[t, 1]) * m
    return h


def config_index(configs: np.ndarray) -> np.ndarray:
    idx = np.zeros(len(configs), dtype=np.int64)
    for b in range(configs.shape[1]):
        idx |= configs[:, b].astype(np.int64) << (8 * b)
    return idx


def all_configs(n: int) -> np.ndarray:
    x = np.arange(2**n, dtype=np.int64)
    nq = (n + 7) // 8
    return np.stack([(x >> (8 * b)) & 255 for b in range(nq)], axis=1).astype(np.uint8)


@pytest.mark.parametrize("fermion", [True, False])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_apply_within_equals_dense_jordan_wigner(fermion: bool, seed: int) -> None:
    n = 7
    s, k, c = (x.numpy() for x in sy.random_terms(n, 120, seed))
    h = dense_hamiltonian(n, s, k, c, fermion)
    cfg = all_configs(n)
    rng = np.random.default_rng(seed + 100)
    sub_i = cfg[np.sort(rng.choice(len(cfg), 70, replace=False))]
    sub_j = cfg[np.sort(rng.choice(len(cfg), 90, replace=False))]
    psi = sy.random_psi(len(sub_i), seed)
    got, _, _ = orc.apply_within(1 if fermion else 2, sub_i, psi, sub_j, s, k, c)
    want = h[np.ix_(config_index(sub_j), config_index(sub_i))] @ psi
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_molecular_table_is_hermitian_and_matches_dense() -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(3, 5))
    h = dense_hamiltonian(6, s, k, c, True)
    assert np.allclose(h, h.conj().T, atol=1e-12)
    cfg = all_configs(6)
    dense_from_oracle = np.zeros((64, 64), dtype=np.complex128)
    for i in range(64):
        e = np.zeros(64, dtype=np.complex128)
        e[i] = 1.0
        dense_from_oracle[:, i], _, _ = orc.apply_within(1, cfg, e, cfg, s, k, c)
    np.testing.assert_allclose(dense_from_oracle, h, rtol=0, atol=1e-12)


def _sector_matrix(pc: int, cfg: np.ndarray, s: np.ndarray, k: np.ndarray, c: np.ndarray) -> np.ndarray:
    n = len(cfg)
    m = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[i] = 1.0
        m[:, i], _, _ = orc.apply_within(pc, cfg, e, cfg, s, k, c)
    return m


def test_free_fermion_ground_energy_closed_form() -> None:
    m, n, ne = 3, 2, 3
    d = sy.free_fermion_dict(m, n)
    s, k, c = (x.numpy() for x in sy.terms_from_dict(d))
    sites = m * n
    single = np.zeros((sites, sites))
    for ((a, _), (b, _)), v in d.items():
        single[a, b] = v
    e_closed = np.linalg.eigvalsh(single)[:ne].sum()
    occ = np.array([[(x >> q) & 1 for q in range(sites)] for x in range(2**sites) if bin(x).count("1") == ne], dtype=np.uint8)
    cfg = sy.pack_occupation(occ)
    e_many = np.linalg.eigvalsh(_sector_matrix(1, cfg, s, k, c))[0]
    assert abs(e_many - e_closed) < 1e-10


@pytest.mark.parametrize("u", [0.0, 2.0, 4.0, 8.0])
def test_hubbard_dimer_closed_form(u: float) -> None:
    t = 1.0
    s, k, c = (x.numpy() for x in sy.terms_from_dict(sy.hubbard_dict(2, 1, t, u)))
    cfg = sy.sector_configs(2, 1, 1, 4, 0)
    e0 = np.linalg.eigvalsh(_sector_matrix(1, cfg, s, k, c))[0]
    assert abs(e0 - (u - np.sqrt(u * u + 16 * t * t)) / 2) < 1e-12


def test_transverse_field_ising_matches_pauli_kronecker() -> None:
    L, hx, jz = 6, 0.7, 1.3
    s, k, c = (x.numpy() for x in sy.terms_from_dict(sy.ising_dict(L, 1, x=hx, zh=jz)))
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    Zs = np.diag([-1.0, 1.0])  # a+a - a a+ = 2n - 1
    h = np.zeros((2**L, 2**L))
    for i in range(L):
        f = [ID] * L
        f[i] = X
        h += hx * _kron_sites(f)
        if i > 0:
            f = [ID] * L
            f[i], f[i - 1] = Zs, Zs
            h += jz * _kron_sites(f)
    cfg = all_configs(L)
    m = _sector_matrix(2, cfg, s, k, c)
    np.testing.assert_allclose(m, h, atol=1e-12)


def test_linearity() -> None:
    s, k, c = sorted_terms(*sy.molecular_terms(4, 11))
    cfg = sy.sector_configs(4, 2, 2, 36, 12)
    a, b = sy.random_psi(36, 1), sy.random_psi(36, 2)
    fa, _, _ = orc.apply_within(1, cfg, a, cfg, s, k, c)
    fb, _, _ = orc.apply_within(1, cfg, b, cfg, s, k, c)
    fab, _, _ = orc.apply_within(1, cfg, 2.5 * a - 1j * b, cfg, s, k, c)
    np.testing.assert_allclose(fab, 2.5 * fa - 1j * fb, atol=1e-12)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_apply_within_equals_serial(threads: int) -> None:
    """The multi-threaded timing variant equals the serial restatement: same M_gen / M_hit, the
    same values bit for bit on one thread and to rounding (regrouped sums) on more; configs_i is
    a different set from configs_j, with rows outside it, and smaller than some thread counts'
    slices would suggest (empty slices included)."""
    s, k, c = sorted_terms(*sy.molecular_terms(8, 21))
    cfg_j = sy.excitation_configs(8, 3, 3, 400, 22)
    cfg_i = np.concatenate([cfg_j[::7], sy.excitation_configs(8, 3, 3, 60, 23)])
    psi = sy.random_psi(len(cfg_i), 24)
    want, _, want_counts = orc.apply_within(1, cfg_i, psi, cfg_j, s, k, c)
    got, got_counts = orc.apply_within_threads(1, cfg_i, psi, cfg_j, s, k, c, threads)
    assert got_counts == want_counts and want_counts[1] > 0
    if threads == 1:
        assert np.array_equal(got, want)
    else:
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * float(np.abs(want).max()))
    few, few_counts = orc.apply_within_threads(1, cfg_i[:2], psi[:2], cfg_j, s, k, c, threads)
    ref, _, ref_counts = orc.apply_within(1, cfg_i[:2], psi[:2], cfg_j, s, k, c)
    assert few_counts == ref_counts
    np.testing.assert_allclose(few, ref, rtol=0, atol=1e-12 * max(float(np.abs(ref).max()), 1.0))


def test_diagonal_term_identity_and_dense() -> None:
    s, k, c = (x.numpy() for x in sy.random_terms(7, 150, 9))
    h = dense_hamiltonian(7, s, k, c, True)
    cfg = all_configs(7)
    d = orc.diagonal_term(1, cfg, s, k, c)
    np.testing.assert_allclose(d, np.diag(h), atol=1e-12)
    for i in range(0, 128, 17):
        single, _, _ = orc.apply_within(1, cfg[i:i + 1], np.array([1.0 + 0j]), cfg[i:i + 1], s, k, c)
        assert abs(single[0] - d[i]) < 1e-12


def test_single_relative_brute_force() -> None:
    s, k, c = (x.numpy() for x in sy.random_terms(8, 80, 4))
    cfg = sy.sort_unique_configs(np.random.default_rng(5).integers(0, 256, size=(40, 1), dtype=np.uint8))
    keys = np.random.default_rng(6).random(len(s))
    got = orc.single_relative(1, cfg, s, k, keys)
    inset = {bytes(r) for r in cfg}
    for i in range(len(cfg)):
        best, pick = -1.0, bytes(1)
        for t in range(len(s)):
            ok, out, _ = orc.apply_term(cfg[i], s[t], k[t], 1)
            if ok and bytes(out) not in inset and keys[t] > best:
                best, pick = keys[t], bytes(out)
        assert bytes(got[i]) == pick


def _molecular_m_gen_per_config(n: int, n_a: int, n_b: int) -> int:
    """Closed form of the connections (successful term applications) per config of a normal-ordered
    molecular table with every term present: the constant; a+_p a_q per spin (q occupied, p empty
    or p = q); same-spin a+a+aa (two occupied annihilated, two created among the empty sites and
    the two freed); opposite-spin a+a+aa (one of each spin annihilated, each created among its
    spin's empty sites and the freed one)."""
    one = sum(k * (n - k) + k for k in (n_a, n_b))
    same = sum(math.comb(k, 2) * math.comb(n - k + 2, 2) for k in (n_a, n_b))
    opposite = n_a * n_b * (n - n_a + 1) * (n - n_b + 1)
    return 1 + one + same + opposite


@pytest.mark.parametrize("n,n_a,n_b,count,seed", [(10, 5, 5, 10_000, 1002), (26, 7, 7, 60, 2002), (6, 3, 1, 80, 9)])
def test_count_connections_closed_form(n: int, n_a: int, n_b: int, count: int, seed: int) -> None:
    """oracle.count_connections gives M_gen = count * the closed form (1,381 per config at 20 spin
    orbitals 5a+5b, the 10,000-config plain workload; 28,701 at 52 spin orbitals 7a+7b) and M_hit =
    the contributions that land in the set (all of them with configs_j = the whole sector)."""
    s, k, _ = sy.molecular_terms(n, seed)
    assert len(s) == sy.molecular_term_count(n)
    if count <= 100:
        cfg = sy.excitation_configs(n, n_a, n_b, count, seed)
    else:
        cfg = sy.sector_configs(n, n_a, n_b, count, seed)
    m_gen, m_hit = orc.count_connections(1, cfg, cfg, s.numpy(), k.numpy())
    assert m_gen == count * _molecular_m_gen_per_config(n, n_a, n_b)
    if n == 10:
        assert _molecular_m_gen_per_config(n, n_a, n_b) == 1_381
    if n == 26:
        assert _molecular_m_gen_per_config(n, n_a, n_b) == 28_701
    full = sy.sector_configs(n, n_a, n_b, math.comb(n, n_a) * math.comb(n, n_b), seed) if count == 80 else None
    if full is not None:
        # configs_j = the whole sector: every connection stays in the sector, so every one is a hit
        gen_f, hit_f = orc.count_connections(1, cfg, full, s.numpy(), k.numpy())
        assert gen_f == hit_f == m_gen
    assert 0 < m_hit <= m_gen
