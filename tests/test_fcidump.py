"""
CPU checks of FCIDUMP ingestion (qmb_hip.fcidump, qmb/fcidump.py:50-199).

openfermion (which the reference uses for the normal ordering) is not installed, so the
ingestion is pinned by mathematics and by a literal restatement of openfermion's algorithm:
  - H2 in the minimal basis (Szabo & Ostlund's integrals at R = 1.4 bohr): the sector's lowest
    eigenvalue equals the closed form of the 2x2 CI (HF and the double excitation) and the
    published total energy -1.1373 Ha;
  - the dense matrix of the ingested terms (through the oracle) equals the Jordan-Wigner matrix of
    the textbook Hamiltonian E0 + sum h_pq a+_ps a_qs + 1/2 sum (pq|rs) a+_ps a+_rt a_st a_qs,
    assembled operator by operator from the integrals, over the whole Fock space;
  - diagonal elements equal the Slater-Condon rule for a determinant;
  - terms, values (bit for bit) and dict order equal a term-by-term transcription of openfermion's
    get_fermion_operator + normal_ordered (bubble-sort ladder ordering, additions that drop
    |c| < 1e-8), including a term whose partial sums vanish and re-insert it;
  - the synthetic molecular generator's terms agree with the ingestion of its integrals;
  - parsing: gzip, header forms, later lines overwrite earlier ones, blank lines, errors; the
    cache tuple round trip and FCIDUMP.yaml reference energy.
"""

from __future__ import annotations

import gzip
import itertools
import pathlib

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from qmb_hip import fcidump as fd
from qmb_hip import synthetic as sy
from test_oracle_math import _op, all_configs, dense_hamiltonian


def _random_integrals(n: int, seed: int, sparsity: float = 0.0) -> tuple[float, np.ndarray, np.ndarray]:
    h, g, e0 = sy.molecular_integrals(n, seed)
    if sparsity:
        # zero whole 8-fold classes: a symmetric mask over canonical orbital pairs
        rng = np.random.default_rng(seed + 1)
        idx = np.arange(n)
        hi, lo = np.maximum(idx[:, None], idx[None, :]), np.minimum(idx[:, None], idx[None, :])
        pair = hi * (hi + 1) // 2 + lo
        n_pair = n * (n + 1) // 2
        mask = rng.random((n_pair, n_pair)) < sparsity
        mask |= mask.T
        g = np.where(mask[pair[:, :, None, None], pair[None, None, :, :]], 0.0, g)
    return e0, h, g


def _write(tmp_path: pathlib.Path, name: str, e0: float, h: np.ndarray, g: np.ndarray, n_electron: int) -> pathlib.Path:
    path = tmp_path / name
    fd.write_fcidump(path, e0, h, g, n_electron)
    return path


def _h2_fcidump(tmp_path: pathlib.Path) -> tuple[pathlib.Path, dict[str, float]]:
    v = {"h11": -1.2528, "h22": -0.4756, "J11": 0.6746, "J22": 0.6975, "J12": 0.6636, "K12": 0.1813, "E_nuc": 1 / 1.4}
    text = (" &FCI NORB=2,NELEC=2,MS2=0,\n  ORBSYM=1,1,\n  ISYM=1,\n &END\n"
            f"{v['J11']!r} 1 1 1 1\n{v['K12']!r} 2 1 2 1\n{v['J12']!r} 2 2 1 1\n{v['J22']!r} 2 2 2 2\n"
            f"{v['h11']!r} 1 1 0 0\n{v['h22']!r} 2 2 0 0\n{v['E_nuc']!r} 0 0 0 0\n")
    path = tmp_path / "H2.FCIDUMP"
    path.write_text(text)
    return path, v


def _dense_terms(terms: tuple[np.ndarray, np.ndarray, np.ndarray], n_so: int) -> np.ndarray:
    """Matrix of a term table over the whole Fock space, from the oracle's apply_within."""
    cfg = all_configs(n_so)
    dim = len(cfg)
    site, kind, coef = terms
    m = np.zeros((dim, dim), dtype=np.complex128)
    for j in range(dim):
        e = np.zeros(dim, dtype=np.complex128)
        e[j] = 1.0
        m[:, j] = orc.apply_within(1, cfg, e, cfg, site, kind, coef)[0]
    return m


def test_h2_minimal_basis_closed_form(tmp_path: pathlib.Path) -> None:
    path, v = _h2_fcidump(tmp_path)
    (n_orbit, n_electron, n_spin), terms = fd.fcidump_terms(path)
    assert (n_orbit, n_electron, n_spin) == (2, 2, 0)
    m = _dense_terms(terms, 4)
    # the Sz = 0, two-electron sector: alpha in {0, 2}, beta in {1, 3} (even = alpha)
    idx = [(1 << a) | (1 << b) for a in (0, 2) for b in (1, 3)]
    e = np.linalg.eigvalsh(m[np.ix_(idx, idx)])[0]
    a = 2 * v["h11"] + v["J11"]
    b = 2 * v["h22"] + v["J22"]
    closed = (a + b) / 2 - np.sqrt(((a - b) / 2) ** 2 + v["K12"] ** 2) + v["E_nuc"]
    assert abs(e - closed) < 1e-12
    assert abs(e - (-1.1373)) < 1e-4  # Szabo & Ostlund, Table 4.x: E(FCI) = -1.1373 Ha at R = 1.4 bohr
    # the HF determinant's diagonal element: 2 h11 + J11 + E_nuc
    assert abs(m[idx[0], idx[0]].real - (a + v["E_nuc"])) < 1e-12


@pytest.mark.parametrize("seed,sparsity", [(11, 0.0), (12, 0.6)])
def test_terms_equal_textbook_hamiltonian(tmp_path: pathlib.Path, seed: int, sparsity: float) -> None:
    n = 3
    n_so = 2 * n
    e0, h, g = _random_integrals(n, seed, sparsity)
    path = _write(tmp_path, "rand.FCIDUMP", e0, h, g, 2)
    _, terms = fd.fcidump_terms(path)
    got = _dense_terms(terms, n_so)
    # textbook second quantisation, operator by operator
    dim = 2**n_so
    want = e0 * np.eye(dim, dtype=np.complex128)
    ops = {(s, k): _op(n_so, s, k, True) for s in range(n_so) for k in (0, 1)}
    for p, q in itertools.product(range(n), repeat=2):
        for sg in (0, 1):
            want += h[p, q] * ops[(2 * p + sg, 1)] @ ops[(2 * q + sg, 0)]
    for p, q, r, s in itertools.product(range(n), repeat=4):
        if g[p, q, r, s] == 0:
            continue
        for sg, tau in itertools.product((0, 1), repeat=2):
            want += 0.5 * g[p, q, r, s] * ops[(2 * p + sg, 1)] @ ops[(2 * r + tau, 1)] @ ops[(2 * s + tau, 0)] @ ops[(2 * q + sg, 0)]
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-13 * scale
    # and the oracle's matrix is the JW matrix of the term table itself
    assert np.abs(dense_hamiltonian(n_so, *terms, fermion=True) - got).max() <= 1e-13 * scale


def test_diagonal_is_slater_condon(tmp_path: pathlib.Path) -> None:
    n = 4
    e0, h, g = _random_integrals(n, 21)
    path = _write(tmp_path, "sc.FCIDUMP", e0, h, g, 4)
    _, (site, kind, coef) = fd.fcidump_terms(path)
    cfg = sy.sector_configs(n, 2, 2, 36, 0)
    diag = orc.diagonal_term(1, cfg, site, kind, coef)
    occ = sy.unpack_occupation(cfg, 2 * n)
    for row, d in zip(occ, diag):
        so = np.nonzero(row)[0]
        e = e0 + sum(h[i // 2, i // 2] for i in so)
        for a, b in itertools.combinations(so, 2):
            pa, pb = a // 2, b // 2
            e += g[pa, pa, pb, pb] - (g[pa, pb, pb, pa] if a % 2 == b % 2 else 0.0)
        assert abs(d - e) <= 1e-12 * (1 + abs(e)), (row, d, e)


def _openfermion_like(e0: float, h: np.ndarray, g: np.ndarray) -> dict[tuple[tuple[int, int], ...], float]:
    """Term-by-term transcription of openfermion's InteractionOperator -> get_fermion_operator ->
    normal_ordered for this operator form: python dicts, additions that delete |c| < 1e-8."""
    n = h.shape[0]
    big = 2 * n
    e1b = np.zeros((big, big))
    e1b[0::2, 0::2] = h
    e1b[1::2, 1::2] = h
    e2 = g.transpose(0, 2, 3, 1) / 2
    e2b = np.zeros((big,) * 4)
    e2b[0::2, 0::2, 0::2, 0::2] = e2
    e2b[0::2, 1::2, 1::2, 0::2] = e2
    e2b[1::2, 0::2, 0::2, 1::2] = e2
    e2b[1::2, 1::2, 1::2, 1::2] = e2

    def add(target: dict, term: tuple, value: float) -> None:
        target[term] = target.get(term, 0.0) + value
        if abs(target[term]) < 1e-8:
            del target[term]

    stage1: dict = {}
    add(stage1, (), e0)
    for idx in itertools.product(range(big), repeat=2):
        if e1b[idx]:
            add(stage1, ((idx[0], 1), (idx[1], 0)), e1b[idx])
    for idx in itertools.product(range(big), repeat=4):
        if e2b[idx]:
            add(stage1, tuple(zip(idx, (1, 1, 0, 0))), e2b[idx])

    def ladder(term: tuple, coefficient: float) -> dict:
        term = list(term)
        out: dict = {}
        for i in range(1, len(term)):
            for j in range(i, 0, -1):
                right, left = term[j], term[j - 1]
                if right[1] and not left[1]:
                    raise AssertionError("no annihilator precedes a creator in these terms")
                if right[1] == left[1]:
                    if right[0] == left[0]:
                        return out
                    if right[0] > left[0]:
                        term[j - 1], term[j] = right, left
                        coefficient *= -1
        add(out, tuple(term), coefficient)
        return out

    ordered: dict = {}
    for term, coefficient in stage1.items():
        for t, c in ladder(term, coefficient).items():
            add(ordered, t, c)
    return ordered


def _cancelling_integrals() -> tuple[float, np.ndarray, np.ndarray]:
    """(11|22) = J and (12|12) = J + 4e-9: the same-spin term (2s,1s,2s,1s) has partial sums of
    2e-9, each deleted on the way, so it is absent; 5e-9 integrals are dropped."""
    n = 3
    h = np.array([[-1.0, 0.2, 0.0], [0.2, -0.5, 5e-9], [0.0, 5e-9, 0.3]])
    g = np.zeros((n,) * 4)

    def put(i: int, j: int, k: int, l: int, v: float) -> None:
        for p, q, r, s in [(i, j, k, l), (i, j, l, k), (j, i, k, l), (j, i, l, k), (l, k, j, i), (k, l, j, i), (l, k, i, j), (k, l, i, j)]:
            g[p, q, r, s] = v

    put(0, 0, 1, 1, 0.3)
    put(0, 1, 0, 1, 0.3 + 4e-9)
    put(0, 0, 0, 0, 0.7)
    put(2, 2, 1, 1, 0.25)
    put(2, 2, 2, 0, 5e-9)
    # (20|12) = 2 (22|10): the same-spin term (2s,1s,2s,0s) sums A/2 - B/2 - B/2 + A/2 with A = 2B,
    # so its third partial sum is 0 (deleted) and the last contribution re-inserts it at the end
    put(2, 0, 1, 2, 0.25)
    put(2, 2, 1, 0, 0.125)
    return 0.6, h, g


@pytest.mark.parametrize("case", ["random", "sparse", "cancelling"])
def test_matches_openfermion_transcription(tmp_path: pathlib.Path, case: str) -> None:
    if case == "cancelling":
        e0, h, g = _cancelling_integrals()
    else:
        e0, h, g = _random_integrals(3, 31, 0.5 if case == "sparse" else 0.0)
    path = _write(tmp_path, f"{case}.FCIDUMP", e0, h, g, 2)
    meta, d = fd.read_fcidump(path)
    want = _openfermion_like(e0, h, g)
    assert list(d.keys()) == list(want.keys())  # same terms in the same dict order
    for k, v in want.items():
        assert d[k] == complex(v), k  # bit for bit
    if case == "cancelling":
        # orbitals 0, 1 -> spin orbitals alpha 0, 2 and beta 1, 3: the same-spin exchange-direct
        # term (J - K = -4e-9) is lost to the rule, the opposite-spin one (direct only) stays
        assert ((2, 1), (0, 1), (2, 0), (0, 0)) not in d and ((3, 1), (1, 1), (3, 0), (1, 0)) not in d
        assert d[((3, 1), (0, 1), (3, 0), (0, 0))] != 0
        assert ((2, 1), (4, 0)) not in d  # h between orbitals 1 and 2 is 5e-9
        assert all(abs(v) >= 1e-8 for v in d.values())
        keys = list(d)
        again = ((4, 1), (2, 1), (4, 0), (0, 0))
        # first contribution from a+_2 a+_4 a_0 a_4, i.e. before every term whose first one has P = 3; the
        # re-insertion by the last contribution (a+_4 a+_2 a_4 a_0) places it after them
        assert d[again] == 0.125 and keys.index(again) > keys.index(((5, 1), (3, 1), (5, 0), (3, 0)))
    assert meta == (h.shape[0], 2, 0)


def test_generator_agrees_with_ingestion(tmp_path: pathlib.Path) -> None:
    n = 5
    h, g, e0 = sy.molecular_integrals(n, 41)
    path = _write(tmp_path, "gen.FCIDUMP", e0, h, g, 4)
    _, terms = fd.fcidump_terms(path)
    got = fd.terms_to_dict(*terms)
    s, k, c = (x.numpy() for x in sy.molecular_terms(n, 41))
    want = fd.terms_to_dict(s, k, c)
    big = {key for key, v in want.items() if abs(v) >= 1e-7}
    assert big <= set(got)
    assert {key for key, v in got.items() if abs(v) >= 1e-7} <= set(want)
    for key in big:
        assert abs(got[key] - want[key]) <= 1e-15 * max(1.0, abs(want[key]))


def test_parsing_rules(tmp_path: pathlib.Path) -> None:
    text = (" &FCI NORB=  2,\n NELEC=2, MS2=0,\n  ORBSYM=1,1,\n ISYM=1\n&END\n"
            "0.5 1 1 1 1\n\n0.25 2 2 1 1\n0.4 1 1 2 2\n"  # the later (11|22) line overwrites (22|11)
            "-1.0 1 1 0 0\n0.125 1 2 0 0\n-0.5 2 2 0 0\n0.7 2 0 0 0\n0.3 0 0 0 0\n")
    p = tmp_path / "a.FCIDUMP"
    p.write_text(text)
    meta, e0, h, g = fd.read_integrals(p)
    assert meta == (2, 2, 0) and e0 == 0.3
    np.testing.assert_array_equal(h, [[-1.0, 0.125], [0.125, -0.5]])  # the Psi4 line "0.7 2 0 0 0" is skipped
    assert g[0, 0, 1, 1] == g[1, 1, 0, 0] == 0.4 and g[0, 0, 0, 0] == 0.5
    pz = tmp_path / "a.FCIDUMP.gz"
    with gzip.open(pz, "wt", encoding="utf-8") as f:
        f.write(text)
    assert fd.read_fcidump(pz) == fd.read_fcidump(p)
    assert fd.read_fcidump(p, cached=True) == ((2, 2, 0), {})
    bad = tmp_path / "bad.FCIDUMP"
    bad.write_text(" &FCI NORB=2,NELEC=2,MS2=0 &END\n0.5 1 1 2 0\n")
    with pytest.raises(ValueError):
        fd.read_fcidump(bad)
    bad.write_text(" &FCI NORB=2,NELEC=2 &END\n")
    with pytest.raises(AssertionError):
        fd.read_fcidump(bad)
    bad.write_text(" &FCI NORB=2,NELEC=2,MS2=0 &END\n0.5 3 1 1 1\n")
    with pytest.raises(IndexError):
        fd.read_fcidump(bad)


def test_model_cache_round_trip(tmp_path: pathlib.Path, monkeypatch: pytest.MonkeyPatch) -> None:
    monkeypatch.setenv("QMB_HIP_CACHE_DIR", str(tmp_path / "cache"))
    models = tmp_path / "models"
    models.mkdir()
    e0, h, g = _random_integrals(4, 51)
    fd.write_fcidump(models / "mol.FCIDUMP.gz", e0, h, g, 4)
    (models / "FCIDUMP.yaml").write_text("mol: -7.25\n")
    cfg = fd.ModelConfig("mol", models)
    first = fd.Model(cfg)
    cache = fd.cache_file_for(models / "mol.FCIDUMP.gz")
    assert cache.exists() and cache.name.endswith("v5") and len(cache.name) == 64 + 2
    second = fd.Model(cfg)  # from the cache
    for a, b in ((first.hamiltonian.site, second.hamiltonian.site), (first.hamiltonian.kind, second.hamiltonian.kind),
                 (first.hamiltonian.coef, second.hamiltonian.coef)):
        assert torch.equal(a, b)
    assert (first.n_qubit, first.n_electron, first.n_spin, first.ref_energy) == (8, 4, 0, -7.25)
    norms = first.hamiltonian.coef.norm(dim=1)
    assert torch.all(norms[:-1] >= norms[1:])  # sorted by |coef| descending, as the dict constructor does
    assert fd.Model(fd.ModelConfig("mol", models, ref_energy=1.5)).ref_energy == 1.5
    # the cached tuple is the reference's torch.save((site, kind, coef)) format
    site, kind, coef = fd.load_terms(cache)
    assert site.dtype == torch.int16 and kind.dtype == torch.uint8 and coef.dtype == torch.float64
    torch.save((site, kind), tmp_path / "broken")
    with pytest.raises(ValueError):
        fd.load_terms(tmp_path / "broken")
    cfg_row = torch.tensor([0b01100110], dtype=torch.uint8)
    assert first.show_config(cfg_row) == "[↓↑↓↑]"
