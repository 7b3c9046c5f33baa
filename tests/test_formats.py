"""
Config format and synthetic workloads: bit packing (the reference's own bitspack round-trip cases,
qmb tests/bitspack_test.py, plus the LSB-first bit order those tests leave open), term-table layout
of prepare<4> (qmb/_hamiltonian.cpp:29-64), the Hamiltonian class's host logic, and the generators
behind the benchmark configurations.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from qmb_hip import bitspack
from qmb_hip import synthetic as sy
from qmb_hip.hamiltonian import Hamiltonian

ROUND_TRIPS = [
    ([1, 0, 1, 1, 0, 0, 1, 0, 1, 1, 0, 1, 1, 0, 0, 1], 1),
    ([3, 2, 1, 0, 3, 2, 1, 0, 3, 2, 1, 0, 3, 2, 1, 0], 2),
    ([15, 10, 5, 0, 15, 10, 5, 0, 15, 10, 5, 0, 15, 10, 5, 0], 4),
    ([15, 210, 5, 123, 151, 130, 5, 0, 15, 10, 75, 0, 15, 10, 5, 0], 8),
    ([1, 0, 1, 1, 0, 0, 1, 0, 1, 1, 0, 1, 1, 0, 0], 1),
    ([3, 2, 1, 0, 3, 2, 1, 0, 3, 2, 1, 0, 3, 2], 2),
    ([15, 10, 5, 0, 15, 10, 5, 0, 15, 10, 5, 0, 15, 10], 4),
]


@pytest.mark.parametrize("values,size", ROUND_TRIPS)
def test_pack_unpack_round_trip(values: list[int], size: int) -> None:
    t = torch.tensor(values, dtype=torch.uint8)
    assert torch.equal(bitspack.unpack_int(bitspack.pack_int(t, size), size, len(values)), t)


@pytest.mark.parametrize("size,high", [(1, 2), (2, 4), (4, 16)])
def test_pack_unpack_large(size: int, high: int) -> None:
    t = torch.randint(0, high, (1000,), dtype=torch.uint8)
    assert torch.equal(bitspack.unpack_int(bitspack.pack_int(t, size), size, 1000), t)


def test_pack_multi_dimensional_and_bit_order() -> None:
    t = torch.tensor([[1, 0, 1, 1], [0, 0, 1, 0], [1, 1, 0, 1], [1, 0, 0, 1]], dtype=torch.uint8)
    assert torch.equal(bitspack.unpack_int(bitspack.pack_int(t, 1), 1, 4), t)
    # site s -> bit s % 8 of byte s // 8 (qmb/bitspack.py:50, weights 1 << i)
    occ = torch.zeros(12, dtype=torch.uint8)
    occ[0] = occ[3] = occ[9] = 1
    assert bitspack.pack_int(occ, 1).tolist() == [0b00001001, 0b00000010]
    assert sy.pack_occupation(occ.numpy()[None]).tolist() == [[0b00001001, 0b00000010]]


def test_terms_from_dict_layout() -> None:
    d = {((3, 1), (1, 0)): 0.5, (): 2.0, ((0, 1), (0, 0), (2, 1), (2, 0)): 1 + 2j}
    site, kind, coef = sy.terms_from_dict(d)
    assert site.dtype == torch.int16 and kind.dtype == torch.uint8 and coef.dtype == torch.float64
    assert kind.tolist() == [[1, 0, 2, 2], [2, 2, 2, 2], [1, 0, 1, 0]]
    assert site.tolist()[0][:2] == [3, 1]
    assert coef.tolist() == [[0.5, 0.0], [2.0, 0.0], [1.0, 2.0]]
    with pytest.raises(ValueError):
        sy.terms_from_dict({((0, 1),) * 5: 1.0})


def test_hamiltonian_sorts_by_norm_and_maps_kind() -> None:
    d = {((1, 1), (0, 0)): 0.1, ((0, 1), (1, 0)): -3.0, (): 1.0 + 1.0j}
    h = Hamiltonian(d, kind="bose2")
    assert h.particle_cut == 2
    assert h.coef.norm(dim=1).tolist() == sorted(h.coef.norm(dim=1).tolist(), reverse=True)
    with pytest.raises(ValueError, match="Unknown kind"):
        Hamiltonian(d, kind="spin")
    with pytest.raises(ValueError, match="Unsupported device type"):
        Hamiltonian._load_module("tpu", 1, 1)  # pylint: disable=protected-access


@pytest.mark.parametrize("n", [2, 4, 10])
def test_molecular_term_count_closed_form(n: int) -> None:
    site, kind, coef = sy.molecular_terms(n, 1)
    assert len(site) == sy.molecular_term_count(n)
    assert torch.all(coef[:, 1] == 0)
    two = kind[:, 2] == 0
    s = site[two]
    assert torch.all(s[:, 0] > s[:, 1]) and torch.all(s[:, 2] > s[:, 3])  # normal ordered


def test_molecular_terms_conserve_spin_and_are_hermitian() -> None:
    site, kind, coef = (x.numpy() for x in sy.molecular_terms(4, 3))
    table = {}
    for t in range(len(site)):
        ops = [(int(site[t, o]), int(kind[t, o])) for o in range(4) if kind[t, o] != 2]
        up = sum(1 if k else -1 for s, k in ops if s % 2 == 0)
        dn = sum(1 if k else -1 for s, k in ops if s % 2 == 1)
        assert up == 0 and dn == 0
        table[tuple(ops)] = coef[t, 0]
    for ops, v in table.items():
        if len(ops) == 4:
            (p, _), (q, _), (r, _), (s, _) = ops
            assert table[((r, 1), (s, 1), (p, 0), (q, 0))] == pytest.approx(v, abs=1e-14)


def test_sector_configs_unique_sorted_in_sector() -> None:
    cfg = sy.sector_configs(10, 5, 5, 10000, 1002)
    assert cfg.shape == (10000, 3)
    keys = [bytes(r) for r in cfg]
    assert keys == sorted(set(keys))
    occ = sy.unpack_occupation(cfg, 20)
    assert np.all(occ[:, 0::2].sum(1) == 5) and np.all(occ[:, 1::2].sum(1) == 5)
    big = sy.sector_configs(26, 7, 7, 2000, 5)
    occ = sy.unpack_occupation(big, 52)
    assert np.all(occ[:, 0::2].sum(1) == 7) and np.all(occ[:, 1::2].sum(1) == 7)


def test_excitation_configs_contain_all_singles_and_doubles() -> None:
    n, na, nb = 8, 3, 3
    cfg = sy.excitation_configs(n, na, nb, 1500, 7)
    occ = sy.unpack_occupation(cfg, 2 * n)
    assert np.all(occ[:, 0::2].sum(1) == na) and np.all(occ[:, 1::2].sum(1) == nb)
    hf = np.zeros(2 * n, dtype=np.uint8)
    hf[0:2 * na:2] = 1
    hf[1:2 * nb:2] = 1
    level = (occ != hf).sum(1) // 2
    singles = 2 * na * (n - na)
    doubles = 2 * math.comb(na, 2) * math.comb(n - na, 2) + (na * (n - na)) ** 2
    assert (level == 0).sum() == 1 and (level == 1).sum() == singles and (level == 2).sum() == doubles
    assert set(np.unique(level)) <= {0, 1, 2, 3, 4}
    keys = [bytes(r) for r in cfg]
    assert keys == sorted(set(keys)) and len(keys) == 1500
    with pytest.raises(ValueError):
        sy.excitation_configs(n, na, nb, 3136, 7)


def test_hubbard_chain_configs_half_filled() -> None:
    cfg = sy.hubbard_chain_configs(16, 5000, 3, n_seed_configs=300)
    occ = sy.unpack_occupation(cfg, 32)
    assert np.all(occ[:, 0::2].sum(1) == 8) and np.all(occ[:, 1::2].sum(1) == 8)
    keys = [bytes(r) for r in cfg]
    assert keys == sorted(set(keys))


def test_hubbard_term_table_matches_reference_builder() -> None:
    d = sy.hubbard_dict(64, 1, 1.0, 4.0)
    assert len(d) == 316  # 64 on-site + 63 bonds x 4
    site, kind, coef = sy.terms_from_dict(d)
    assert int(site.max()) == 127
